"""Free bodies in the Broyden-for-everything form on the GPU
(USE_BROYDEN_FOR_EVERYTHING with PRESCRIBED_BODY_VELOCITIES false) against the
oracle's restatement (oracle/swimmer_oracle.py bfe_coupling, BfeInverse,
free_bfe_loop; oracle/broyden_oracle.py compute_errors; pinned on the CPU in
tests/test_oracle_bfe_free.py):

* the residual's body rows (force and torque balance) and blob rows (rigid
  motion from the body's positions), Euler and BDF2, <= 1e-12 relative;
* the coupling blocks (tfs_filament_body_coupling_blobs: A, the balance's D,
  the Schur matrix's S0 with the blob terms, the blob displacements),
  <= 1e-11 relative;
* the engine's J0^{-1} with the blobs eliminated
  (tfs_broyden_set_body_coupling_blobs) against BfeInverse, <= 1e-12;
* two elastic implicit steps of a free swimmer (4 filaments x 20 segments on
  300 blobs) through tfs_elastic_step against the oracle loop: the same
  Broyden iteration counts (within 10 %), positions within 1e-6, the body
  within 1e-8."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import broyden_oracle as BO
from oracle import filament_oracle as F
from oracle import swimmer_oracle as S
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
import test_free_body_elastic_gpu as TF

pytestmark = pytest.mark.gpu
PI = np.pi


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nt", [0, 2])
def test_free_bfe_residual_and_coupling_match_oracle(nt):
    print(f"case free-body Broyden-for-everything rows and coupling, nt {nt}", flush=True)
    p, model, swimmers, rng = TF.swimmer_pair(nswim=1, nfil=3, nseg=8, nb=40, seed=7)
    body, fils, refs, blob_refs = swimmers[0]
    nfil, N, nb = 3, p.nseg, len(blob_refs)
    n6 = 6 * N
    n = n6 * nfil + 6
    try:
        for k in range(nt + 1):
            model.initial_guess(k)
            S.swimmer_initial_guess(body, fils, refs, k, p)
            du = rng.standard_normal(n) * 0.02
            du[n - 6:] *= 0.5
            model.update(du)
            S.swimmer_update(body, fils, refs, du, p)
        fb = rng.standard_normal(3 * nb) * 3.0
        forces = [F.internal_forces(fl, p) for fl in fils]
        f = model.forces(nt)
        assert rel(f, np.concatenate([fc.reshape(-1) for fc in forces])) < 1e-12
        xb = model.blob_positions()
        vs = rng.standard_normal(6 * nfil * N)
        vb = rng.standard_normal(3 * nb) * 0.1
        cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
        st = {"seg_x": cat("x"), "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"),
              "seg_um1": cat("um1"), "v_segs": vs, "body_x": body.x, "body_xm1": body.xm1, "body_xm2": body.xm2,
              "body_u": body.u, "body_um1": body.um1, "v_bodies": np.zeros(6), "x_blobs": xb, "v_blobs": vb,
              "f_segs": f, "f_blobs": fb}
        want, _ = BO.compute_errors(st, 1, nfil, N, nt >= 1, p.dt, p.dl, 1e-8, blobs_per_swim=nb, free_bodies=True)
        sizes = [n6] * nfil + [6] + [3] * nb
        eng = BroydenSolver(sizes)
        try:
            eng.compute_errors(st, 0, num_swim=1, num_fil=nfil, num_seg=N, bdf2=nt >= 1, dt=p.dt, dl=p.dl, tol=1e-8,
                               blobs_per_swim=nb, free_bodies=True)
            got = eng.read_error(0)
        finally:
            eng.close()
        print("residual rel err", rel(got, want), flush=True)
        assert rel(got, want) < 1e-12
        # the rows are not trivial: force and torque balance and rigid motion
        assert np.abs(want[n6 * nfil:n6 * nfil + 6]).max() > 1.0 and np.abs(want[n:]).max() > 1e-2
        blocks = np.zeros(nfil * n6 * n6 + 36)
        model.jacobian(nt, f, blocks)
        A, B, S0, disp = model.body_coupling_blobs(nt, fb, 1.0)
    finally:
        model.close()
    Smat, Ds, Bs, dref = S.bfe_coupling(body, fils, refs, blob_refs, fb, nt, p)
    S.bfe_clamping(fils, forces, nt, p, Bs)
    want_A = np.concatenate([b.reshape(-1) for b in Bs])
    want_B = np.concatenate([d.reshape(-1) for d in Ds])
    Smat, dref = Smat.reshape(-1), dref.reshape(-1)
    print("rel err A", rel(A, want_A), "B", rel(B, want_B), "S0", rel(S0, Smat), "disp", rel(disp, dref), flush=True)
    assert rel(A, want_A) < 1e-11 and rel(B, want_B) < 1e-11 and rel(S0, Smat) < 1e-11
    assert rel(disp, dref) < 1e-14


def test_engine_inverse_with_blobs_matches_oracle():
    print("case J0^{-1} with blob elimination, 3 filaments x 6 segments + 20 blobs", flush=True)
    rng = np.random.default_rng(9)
    nfil, N, nb = 3, 6, 20
    n6 = 6 * N
    p = F.Params(nseg=N, dt=0.2)
    Js = [np.eye(n6) * 3 + rng.standard_normal((n6, n6)) * 0.2 for _ in range(nfil)]
    Ds = [rng.standard_normal((6, n6)) * 0.3 for _ in range(nfil)]
    Bs = [rng.standard_normal((n6, 6)) * 0.3 for _ in range(nfil)]
    Smat = np.eye(6) * 40 + rng.standard_normal((6, 6))
    disp = rng.standard_normal((nb, 3)) * 5
    for nt in (0, 2):
        fac = -6.0 * PI * p.mu
        dt_fac = (1.5 if nt >= 1 else 1.0) / p.dt
        sizes = [n6] * nfil + [6] + [3] * nb
        eng = BroydenSolver(sizes, confidence=0.1)
        try:
            eng.set_jacobian(Js + [np.eye(6)] + [np.eye(3) / fac] * nb)
            eng.set_body_coupling_blobs(1, nfil, nb, np.concatenate([b.reshape(-1) for b in Bs]),
                                        np.concatenate([d.reshape(-1) for d in Ds]), Smat.reshape(-1),
                                        disp.reshape(-1), fac, dt_fac)
            e = rng.standard_normal(sum(sizes))
            eng.start(e)
            up = eng.find_update()
        finally:
            eng.close()
        want = -0.1 * S.BfeInverse(Js, Smat, Ds, Bs, disp, nt, p).apply(e)
        print("nt", nt, "rel err", rel(up, want), flush=True)
        assert rel(up, want) < 1e-12


STEPS = 2


@pytest.fixture(scope="module")
def bfe_case():
    nfil, nseg, nb, R = 4, 20, 300, 9.0
    body = G.make_body(nfil, nseg, nb, R)
    dl, kb = 2.2, 1800.0
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    p = F.Params(nseg=nseg, dl=dl, kb=kb, kt=kb, dt=0.1, end_force=efm)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    refs = [body.fil_base[i] - body.centre for i in range(nfil)]
    b = S.Body.at_rest(body.centre)
    fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (nseg, 1)), nseg, dl) for i in range(nfil)]
    runs = S.free_bfe_loop(b, fils, refs, body.blob_refs, p, STEPS, 1e-8, pert)
    return dict(body=body, p=p, pert=pert, refs=refs, q0=q0, runs=runs)


def test_free_swimmer_bfe_elastic_step_matches_oracle_loop(bfe_case):
    print("case free swimmer, Broyden for everything, 4x20 + 300 blobs, tfs_elastic_step", flush=True)
    c = bfe_case
    body, p, nb = c["body"], c["p"], c["body"].num_blobs
    nfil, nseg = body.num_fil, p.nseg
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    model = FM.FilamentModel(1, nfil, nseg, dl=p.dl, kb=p.kb, kt=p.kt, dt=p.dt, end_force=p.end_force,
                             memory=L.MEMORY_DEVICE)
    eng = BroydenSolver(model.block_sizes(blobs_per_swim=nb), memory=L.MEMORY_DEVICE)
    f_blobs = dev(np.zeros(3 * nb))
    step = None
    iters, xs, bx, fbs = [], [], [], []
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(np.repeat(c["q0"][:, None, :], nseg, axis=1).reshape(-1)))
        model.set_body(dev(body.centre), dev([1.0, 0.0, 0.0, 0.0]), dev(np.array(c["refs"]).reshape(-1)), nb,
                       dev(body.blob_refs.reshape(-1)))
        model.set_perturbation(4, 12, dev(c["pert"].reshape(-1)))
        step = FM.ElasticStep(model, None, eng, None, None, None, f_blobs, tol=1e-8, max_iter=400,
                              broyden_blobs=nb, blob_radius=1.0)
        for nt in range(STEPS):
            st = step.step(nt)
            print("step", nt, st, flush=True)
            assert st["converged"]
            iters.append(st["broyden_iterations"])
            xs.append(model.read(FM.X))
            bx.append(model.read(FM.BODY_X))
            fbs.append(f_blobs.cpu().numpy())
    finally:
        if step is not None:
            step.close()
        model.close()
        eng.close()
    ref = c["runs"]
    print("broyden iterations", iters, "oracle", [r["iters"] for r in ref],
          "max |dx|", [float(np.abs(a - r["x"]).max()) for a, r in zip(xs, ref)],
          "body", [float(np.abs(a - r["body"].x).max()) for a, r in zip(bx, ref)], flush=True)
    for s in range(STEPS):
        r = ref[s]
        assert r["iters"] >= 10 and abs(iters[s] - r["iters"]) <= max(2, r["iters"] // 10)
        assert np.abs(xs[s] - r["x"]).max() <= 1e-6
        assert np.abs(bx[s] - r["body"].x).max() <= 1e-8 + 1e-3 * np.abs(r["body"].x - body.centre).max()
        assert np.abs(fbs[s] - r["f_blobs"]).max() <= 1e-3 * np.abs(r["f_blobs"]).max()
    # the body swam
    assert np.abs(bx[-1] - body.centre).max() > 1e-6
