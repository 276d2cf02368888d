"""The two production workloads at their full sizes on the GPU, checked against
something other than the GPU itself.

* 200 prescribed-beat timesteps of the production body (640 filaments x 20
  segments + 9,000 blobs on R = 50a, dt = T/300, GMRES(50) to 1e-8 with FP64
  Arnoldi, quadratic warm start, filament/patch block preconditioner), three
  ways: the mixed-precision solve (FP32-tile inner matvecs inside FP64
  iterative refinement), the all-FP64 solve at the same tolerance, and an FP64
  truth at 1e-12.  Over the trajectory the solve tolerance sets the deviation
  (the phases integrate every step's error): mixed precision must be no
  further from the truth than FP64.  At one state (the truth's last) both
  solve within 1e-6 of the truth's solution, and the mixed solution is checked
  on the CPU: the oracle's <3,3> RPY direct sum plus the K_phi / K_theta
  columns and rows (the saddle operator of mobility_solver.cpp:1806-2090) at
  the geometry the oracle's beat gives for the step's phases, against the
  oracle's right-hand side (mobility_solver.cpp:240-440).
* The 64-body uneven ensemble (filaments U[100, 1000] x 20 segments, blobs
  U[2000, 9000]) as one block-diagonal 6-DOF matvec launch, in FP64 and with
  FP32 tiles, default (source split), autotuned, plain and persistent launch
  shapes and a hipGraph replay: the smallest, median and largest bodies and five seeded random ones
  against the oracle's evaluate_full_mobility of that body alone
  (rpy_mobility_solver.cu:330-337).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.ensemble import make_ensemble
from timfilsimmod_amd.mobility import RpyHandle
from timfilsimmod_amd.solver import PRECOND_BLOCKS, SOLVE_FP64, SOLVE_MIXED, BeatSimulation, beat_setup

pytestmark = pytest.mark.gpu

DL = 2.2
STEPS = 200


def saddle_apply(x_segs, x_blobs, kphi, ktheta, nfil, nseg, v):
    """A v of the prescribed-beat saddle system with the oracle's FP64 direct
    sum: [M lam - K_phi phi_dot - K_theta theta_dot; -K_phi^T lam; -K_theta^T lam]."""
    ns, nb = nfil * nseg, len(x_blobs) // 3
    lam_s, lam_b = v[:3 * ns], v[3 * ns:3 * ns + 3 * nb]
    pd, td = v[3 * (ns + nb):3 * (ns + nb) + nfil], v[3 * (ns + nb) + nfil:]
    f6 = np.zeros((ns, 6))
    f6[:, :3] = lam_s.reshape(-1, 3)
    vs, vb = O.full_mobility(x_segs, x_blobs, f6.reshape(-1), lam_b, seg_dof=3, bs_accumulate=True)
    kp, kt = kphi.reshape(nfil, -1), ktheta.reshape(nfil, -1)
    out_s = vs.reshape(-1, 6)[:, :3].reshape(nfil, -1) - kp * pd[:, None] - kt * td[:, None]
    ls = lam_s.reshape(nfil, -1)
    return np.concatenate([out_s.reshape(-1), vb, -(kp * ls).sum(1), -(kt * ls).sum(1)])


def make_sim(body, mode, tol):
    return BeatSimulation(body, mode=mode, tol=tol, restart=50, inner_tol=1e-6, max_cycles=60, max_refinements=20,
                          warm_start=2, precond=PRECOND_BLOCKS)


def run_beat(body, mode, tol):
    sim = make_sim(body, mode, tol)
    try:
        stats, states = [], []
        for k in range(STEPS):
            stats.append(sim.step())
            if k % 20 == 19:
                states.append(sim.state())
        return stats, states
    finally:
        sim.close()


def one_step_from(body, mode, tol, state, warm):
    """One solve at `state`'s phases, started from `warm` (tfs_beat_restore)."""
    sim = make_sim(body, mode, tol)
    try:
        sim.restore(state["phase"], state["angle"], warm)
        st = sim.step()
        return st, sim.state()
    finally:
        sim.close()


def test_production_beat_trajectory_mixed_against_fp64_truth():
    body = G.production_body()
    nfil, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    runs = {}
    for name, mode, tol in (("truth", SOLVE_FP64, 1e-12), ("fp64", SOLVE_FP64, 1e-8), ("mixed", SOLVE_MIXED, 1e-8)):
        print(f"case production body, {name}, tol {tol}, {STEPS} steps", flush=True)
        runs[name] = run_beat(body, mode, tol)
        st = runs[name][0]
        print(f"  iterations per step {np.mean([s['iterations'] for s in st]):.1f}, "
              f"fp32 matvecs {np.mean([s['fp32_matvecs'] for s in st]):.1f}, "
              f"max relres {max(s['relres'] for s in st):.2e}", flush=True)
        assert all(s["converged"] for s in st)
        assert max(s["relres"] for s in st) <= tol
    truth = runs["truth"][1]
    dev, err = {}, {}
    for name in ("fp64", "mixed"):
        states = runs[name][1]
        dev[name] = max(np.abs(a["x_segs"] - b["x_segs"]).max() for a, b in zip(states, truth)) / DL
        err[name] = max(np.linalg.norm(a["solution"] - b["solution"]) / np.linalg.norm(b["solution"])
                        for a, b in zip(states, truth))
    print(f"trajectory deviation from the truth / DL over {STEPS} steps: {dev}; solution relative error: {err}",
          flush=True)
    # Over a trajectory the solve tolerance, not the precision, sets the
    # deviation (the phases integrate every step's error): mixed precision is
    # no further from the truth than FP64 at the same tolerance.
    assert dev["mixed"] <= 1.5 * dev["fp64"] and err["mixed"] <= 1.5 * err["fp64"]
    assert dev["fp64"] < 0.05 and err["fp64"] < 0.05
    its = {k: np.mean([s["iterations"] for s in runs[k][0]]) for k in runs}
    assert its["mixed"] <= 1.3 * its["fp64"]

    # At one state (the truth's last), from the same warm start: the saddle
    # system is ill-conditioned (K_phi and K_theta nearly parallel in the power
    # stroke), so a solve to 1e-8 is ~1e-4 from the exact solution in either
    # precision; mixed precision must add nothing to that.
    fin, warm = truth[-1], truth[-2]["solution"]
    one = {}
    for name, mode in (("fp64", SOLVE_FP64), ("mixed", SOLVE_MIXED)):
        st, state = one_step_from(body, mode, 1e-8, fin, warm)
        assert st["converged"] and st["iterations"] > 10
        assert np.array_equal(state["phase"], fin["phase"])
        one[name] = (state, np.linalg.norm(state["solution"] - fin["solution"]) / np.linalg.norm(fin["solution"]))
    print(f"one step at the truth's state: relative solution error fp64 {one['fp64'][1]:.2e}, "
          f"mixed {one['mixed'][1]:.2e}", flush=True)
    assert one["mixed"][1] <= 1.5 * one["fp64"][1] + 1e-12
    assert one["fp64"][1] < 1e-2

    # That mixed solution, independently on the CPU: geometry from the oracle's
    # beat at the step's phases, A x by the oracle's FP64 direct sum.
    state = one["mixed"][0]
    su = beat_setup(body)
    geo = SO.beat_geometry(su, nseg, B.BEAT_A, B.BEAT_B, phase=state["phase"], angle=state["angle"])
    assert np.abs(geo["x_segs"] - state["x_segs"]).max() <= 1e-10 * np.abs(state["x_segs"]).max()
    rhs = SO.rhs_ref(su, nseg, nb, geo["x_segs"], body.x_blobs, body.centre, np.zeros(6), phase=state["phase"],
                     angle=state["angle"])
    ax = saddle_apply(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], nfil, nseg, state["solution"])
    relres = np.linalg.norm(rhs - ax) / np.linalg.norm(rhs)
    print(f"mixed solution: CPU oracle relative residual {relres:.3e}", flush=True)
    assert relres <= 2e-8


def ensemble_handle(e, precision):
    h = RpyHandle(e.num_segs, e.num_blobs, seg_dof=6, precision=precision, bs_accumulate=True)
    h.set_bodies(e.segs_per_body, e.blobs_per_body)
    for what, arr in ((L.SEG_POSITIONS, e.x_segs), (L.BLOB_POSITIONS, e.x_blobs), (L.SEG_FORCES, e.f_segs),
                      (L.BLOB_FORCES, e.f_blobs)):
        h.upload(what, arr)
    h.sync()
    return h


def read(h, e):
    vs, vb = np.zeros(6 * e.num_segs), np.zeros(3 * e.num_blobs)
    h.download(L.SEG_VELOCITIES, vs)
    h.download(L.BLOB_VELOCITIES, vb)
    h.sync()
    return vs, vb


def checked_bodies(e):
    cost = e.pairs_per_body()
    order = np.argsort(cost)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    rng = np.random.default_rng(64)
    for b in rng.permutation(len(cost)):
        if len(pick) == 8:
            break
        if int(b) not in pick:
            pick.append(int(b))
    return pick


def body_slice(e, b, vs, vb):
    s0, s1 = int(e.segs_per_body[:b].sum()), int(e.segs_per_body[:b + 1].sum())
    b0, b1 = int(e.blobs_per_body[:b].sum()), int(e.blobs_per_body[:b + 1].sum())
    return np.concatenate([vs[6 * s0:6 * s1], vb[3 * b0:3 * b1]])


def test_ensemble_64_bodies_match_per_body_oracle():
    e = make_ensemble(64)
    pick = checked_bodies(e)
    print(f"case 64-body ensemble: {e.num_segs} segments, {e.num_blobs} blobs; checked bodies {pick}", flush=True)
    ref = {}
    for b in pick:
        x = e.bodies[b]
        vs, vb = O.full_mobility(x.x_segs, x.x_blobs, x.f_segs, x.f_blobs, seg_dof=6, bs_accumulate=True)
        ref[b] = np.concatenate([vs, vb])
    for prec, tol, name in ((L.PRECISION_FP64, 1e-12, "fp64"), (L.PRECISION_FP32_ACC64, 1e-6, "fp32 tiles")):
        h = ensemble_handle(e, prec)
        try:
            results = {}
            h.evaluate()
            results["default"] = read(h, e)
            h.autotune(prec, reps=1)
            print(f"  {name}: autotuned shape {h.tile_shape(prec)}", flush=True)
            h.evaluate()
            results["autotuned"] = read(h, e)
            keep = h.tile_shape(prec)
            h.set_tile_shape(prec, 256, 2, 1024)
            h.evaluate()
            results["plain"] = read(h, e)
            h.set_persistent(prec, 2)
            h.evaluate()
            results["persistent"] = read(h, e)
            h.set_persistent(prec, 0)
            if keep["split"]:
                h.set_split_shape(prec, keep["block"], keep["per_thread"], keep["chunk"])
            else:
                h.set_tile_shape(prec, keep["block"], keep["per_thread"], keep["chunk"])
                h.set_persistent(prec, keep["persistent"])
            g = h.graph(L.BLOCK_ALL, 1)
            rng = np.random.default_rng(1)
            for k in range(3):
                h.upload(L.SEG_VELOCITIES, rng.normal(size=6 * e.num_segs))
                h.upload(L.BLOB_VELOCITIES, rng.normal(size=3 * e.num_blobs))
                h.sync()
                g.launch()
                got = read(h, e)
                assert all(np.array_equal(a, c) for a, c in zip(got, results["autotuned"])), (name, k)
            g.close()
            for shape, (vs, vb) in results.items():
                errs = []
                for b in pick:
                    got = body_slice(e, b, vs, vb)
                    errs.append(np.linalg.norm(got - ref[b]) / np.linalg.norm(ref[b]))
                print(f"  {name} {shape}: worst body error {max(errs):.2e}", flush=True)
                assert max(errs) <= tol, (name, shape, errs)
        finally:
            h.close()
