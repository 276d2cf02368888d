"""Free bodies in the elastic implicit timestep on the GPU against the oracle's
free swimmer (oracle/swimmer_oracle.py, pinned on the CPU):

* the blob-only solve with body rows (tfs_velocities_compute_free): segment
  velocities, blob forces and body velocities against the dense solve, with
  and without an external force on the body; the body mobility from six such
  solves against (K^T M_bb^{-1} K)^{-1};
* the swimmer's state kernels (body initial guess and update, anchored
  filaments, blob positions) and the body's part of J0 (A_i, B_i, S0 with and
  without the sphere correction) against the restatement, <= 1e-11 relative;
* the Broyden engine's coupled J0^{-1} (tfs_broyden_set_body_coupling)
  against the inverse of the assembled block Jacobian;
* two elastic implicit timesteps of a free swimmer (4 filaments x 20
  segments on 400 blobs) through tfs_elastic_step against the oracle's loop:
  the same Broyden iteration counts (within 10 %) and states within 1e-6."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import filament_oracle as F
from oracle import oracle as O
from oracle import swimmer_oracle as S
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
from timfilsimmod_amd.solver import BLOB_DIRECT, SOLVE_FP64, ElasticVelocities

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("external", [False, True])
def test_free_blob_solve_matches_dense_oracle(external):
    print(f"case free blob solve 3x6 segments + 150 blobs, external force {external}", flush=True)
    body = G.make_body(3, 6, 150, 6.0)
    X = np.array([0.3, -0.2, 0.1])
    fe = np.array([1.0, -2.0, 0.5, 3.0, 0.2, -1.0]) if external else None
    ev = ElasticVelocities(1, 18, 150, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-12, free_bodies=True)
    try:
        vs, fb, vb, st = ev.compute_free(body.x_segs, body.f_segs, body.x_blobs, X, fe)
    finally:
        ev.close()
    xs, xb, fs = body.x_segs, body.x_blobs, body.f_segs
    fils = [F.Filament.at_rest(np.zeros(3), np.tile([1.0, 0, 0, 0], (6, 1)), 6, 2.2) for _ in range(3)]
    for i, f in enumerate(fils):
        f.x = xs.reshape(3, 6, 3)[i]
    sf = S.body_force(S.Body.at_rest(X), fils, [fs.reshape(3, 6, 6)[i] for i in range(3)])
    if external:
        sf = sf + fe
    v_b = O.block("bs", xs, xb, fs, np.zeros(xb.size), seg_dof=6, bs_accumulate=False)
    lam, V = S.free_blob_solve(O.dense_mobility(np.zeros(0), xb), xb, X, v_b, sf)
    v_s = O.block("sb", xs, xb, fs, lam, O.block("ss", xs, xb, fs, lam, seg_dof=6), seg_dof=6)
    print("rel err f_blobs", rel(fb, lam), "V", rel(vb, V), "v_segs", rel(vs, v_s), st, flush=True)
    assert st["converged"]
    # GMRES stops at a relative residual of 1e-12; the blob forces carry it
    # times the condition number of the blob mobility (overlapping blobs,
    # ~1e3 here), the velocities do not
    assert rel(fb, lam) < 1e-8 and rel(vb, V) < 1e-10 and rel(vs, v_s) < 1e-10


def test_direct_free_blob_solve_matches_dense_oracle():
    print("case direct free-body solve 3x6 segments + 150 blobs, turned body", flush=True)
    body = G.make_body(3, 6, 150, 6.0)
    X = np.array([0.3, -0.2, 0.1])
    q = np.roll(Rotation.random(random_state=11).as_quat(), 1)
    R = F.rot_mat(q)
    xb = (X + body.blob_refs @ R.T).reshape(-1)
    fe = np.array([1.0, -2.0, 0.5, 3.0, 0.2, -1.0])
    ev = ElasticVelocities(1, 18, 150, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-12, free_bodies=True)
    try:
        ev.set_body_frame(body.blob_refs.reshape(-1))
        ev.set_preconditioner(BLOB_DIRECT)
        out = [ev.compute_free(body.x_segs, body.f_segs, xb, X, fe, q) for _ in range(2)]
    finally:
        ev.close()
    xs, fs = body.x_segs, body.f_segs
    fils = [F.Filament.at_rest(np.zeros(3), np.tile([1.0, 0, 0, 0], (6, 1)), 6, 2.2) for _ in range(3)]
    for i, f in enumerate(fils):
        f.x = xs.reshape(3, 6, 3)[i]
    sf = S.body_force(S.Body.at_rest(X), fils, [fs.reshape(3, 6, 6)[i] for i in range(3)]) + fe
    v_b = O.block("bs", xs, xb, fs, np.zeros(xb.size), seg_dof=6, bs_accumulate=False)
    lam, V = S.free_blob_solve(O.dense_mobility(np.zeros(0), xb), xb, X, v_b, sf)
    v_s = O.block("sb", xs, xb, fs, lam, O.block("ss", xs, xb, fs, lam, seg_dof=6), seg_dof=6)
    for vs, fb, vb, st in out:
        print("rel err f_blobs", rel(fb, lam), "V", rel(vb, V), "v_segs", rel(vs, v_s), flush=True)
        assert rel(fb, lam) < 1e-10 and rel(vb, V) < 1e-11 and rel(vs, v_s) < 1e-11


def test_body_mobility_matches_oracle():
    print("case body mobility, 300 blobs", flush=True)
    refs = G.fibonacci_sphere(300, 8.0)
    X = np.array([1.0, 2.0, -0.5])
    ev = ElasticVelocities(1, 2, 300, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-12, free_bodies=True)
    try:
        N = ev.body_mobility((refs + X).reshape(-1), X)[0]
    finally:
        ev.close()
    ref = S.body_mobility_reference(O.dense_mobility(np.zeros(0), refs.reshape(-1)), refs)
    print("rel err", rel(N, ref), flush=True)
    assert rel(N, ref) < 1e-9


def swimmer_pair(nswim=1, nfil=3, nseg=8, nb=40, seed=0):
    """The same free swimmers on the device (one handle) and on the oracle (one
    restatement per swimmer)."""
    rng = np.random.default_rng(seed)
    p = F.Params(nseg=nseg, dl=2.2, kb=1800.0, kt=1300.0, dt=0.2, end_force=150.0)
    xb0 = np.array([[0.5, -1.0, 2.0], [30.0, 4.0, -3.0]])[:nswim]
    qb0 = np.array([np.roll(Rotation.random(random_state=seed + s).as_quat(), 1) for s in range(nswim)])
    refs = [[7.0 * v / np.linalg.norm(v) for v in rng.standard_normal((nfil, 3))] for _ in range(nswim)]
    blob_refs = np.concatenate([G.fibonacci_sphere(nb, 5.0 + s) for s in range(nswim)])
    qs, bases = [], []
    for s in range(nswim):
        R0 = F.rot_mat(qb0[s])
        for i in range(nfil):
            q = [np.roll(Rotation.random(random_state=seed + 20 + 10 * s + i).as_quat(), 1)]
            for _ in range(nseg - 1):
                q.append(F.qmul(F.lie_exp(0.1 * rng.standard_normal(3)), q[-1]))
            qs.append(np.array(q))
            bases.append(xb0[s] + R0 @ refs[s][i])
    model = FM.FilamentModel(nswim, nfil, nseg, dl=p.dl, kb=p.kb, kt=p.kt, dt=p.dt, end_force=p.end_force)
    model.set_state(np.array(bases).reshape(-1), np.array(qs).reshape(-1))
    model.set_body(xb0.reshape(-1), qb0.reshape(-1), np.array(refs).reshape(-1), nb, blob_refs.reshape(-1))
    swimmers = []
    for s in range(nswim):
        fils = [F.Filament.at_rest(bases[s * nfil + i], qs[s * nfil + i], nseg, p.dl) for i in range(nfil)]
        swimmers.append((S.Body.at_rest(xb0[s], qb0[s]), fils, refs[s], blob_refs[s * nb:(s + 1) * nb]))
    return p, model, swimmers, rng


@pytest.mark.parametrize("nswim,sphere", [(1, 0.0), (1, 5.0), (2, 5.0)])
def test_swimmer_state_and_body_coupling_match_oracle(nswim, sphere):
    print(f"case swimmer state and J0 coupling {nswim} x 3 x 8, sphere {sphere}", flush=True)
    p, model, swimmers, rng = swimmer_pair(nswim=nswim, seed=3)
    nfil, N = 3, p.nseg
    n = 6 * (nfil * N + 1)
    N_ref = [np.diag([0.02, 0.021, 0.019, 0.001, 0.0011, 0.0009]) + 1e-4 * rng.standard_normal((6, 6))
             for _ in range(nswim)]
    cat = lambda name: np.concatenate([getattr(f, name).reshape(-1) for _, fils, _, _ in swimmers for f in fils])
    try:
        for nt in range(3):
            model.initial_guess(nt)
            du = rng.standard_normal(n * nswim) * 0.02
            for s, (body, fils, refs, _) in enumerate(swimmers):
                du[n * s + n - 6:n * (s + 1)] *= 0.5
                S.swimmer_initial_guess(body, fils, refs, nt, p)
                S.swimmer_update(body, fils, refs, du[n * s:n * (s + 1)], p)
            model.update(du)
            assert rel(model.read(FM.X), cat("x")) < 1e-12 and rel(model.read(FM.Q), cat("q")) < 1e-12
            bodies = [sw[0] for sw in swimmers]
            assert rel(model.read(FM.BODY_X), np.concatenate([b.x for b in bodies])) < 1e-14
            assert rel(model.read(FM.BODY_Q), np.concatenate([b.q for b in bodies])) < 1e-14
            assert rel(model.read(FM.BODY_U), np.concatenate([b.u for b in bodies])) < 1e-14
            want = np.concatenate([S.blob_positions(b, br) for b, _, _, br in swimmers])
            assert rel(model.blob_positions(), want) < 1e-13
        nt = 2
        f = model.forces(nt)
        forces = [[F.internal_forces(fl, p) for fl in fils] for _, fils, _, _ in swimmers]
        assert rel(f, np.concatenate([fc.reshape(-1) for fs in forces for fc in fs])) < 1e-12
        blocks = np.zeros(nswim * (nfil * (6 * N) ** 2 + 36))
        model.jacobian(nt, f, blocks)
        A, B, S0 = model.body_coupling(nt, np.array(N_ref).reshape(-1), sphere)
    finally:
        model.close()
    As, Bs, S0s = [], [], []
    for s, (body, fils, refs, _) in enumerate(swimmers):
        S0r, Ar, Br = S.body_coupling(body, fils, refs, forces[s], nt, N_ref[s], p, sphere)
        As += Ar
        Bs += Br
        S0s.append(S0r)
    want_A = np.concatenate([a.reshape(-1) for a in As])
    want_B = np.concatenate([b.reshape(-1) for b in Bs])
    want_S = np.concatenate([m.reshape(-1) for m in S0s])
    print("rel err A", rel(A, want_A), "B", rel(B, want_B), "S0", rel(S0, want_S), flush=True)
    assert rel(A, want_A) < 1e-11 and rel(B, want_B) < 1e-11 and rel(S0, want_S) < 1e-11
    assert A.size == nswim * nfil * 6 * N * 6


def test_coupled_broyden_inverse_matches_assembled_jacobian():
    print("case coupled J0^{-1}, 2 swimmers x 3 filaments x 6 segments", flush=True)
    rng = np.random.default_rng(4)
    nsw, nfil, N = 2, 3, 6
    n6 = 6 * N
    Js, As, Bs, S0s = [], [], [], []
    for s in range(nsw):
        Js.append([np.eye(n6) * 3 + rng.standard_normal((n6, n6)) * 0.2 for _ in range(nfil)])
        As.append([rng.standard_normal((n6, 6)) * 0.3 for _ in range(nfil)])
        Bs.append([rng.standard_normal((6, n6)) * 0.3 for _ in range(nfil)])
        S0s.append(np.eye(6) * 4 + rng.standard_normal((6, 6)) * 0.2)
    sizes = ([n6] * nfil + [6]) * nsw
    eng = BroydenSolver(sizes, confidence=0.1)
    try:
        eng.set_jacobian([b for s in range(nsw) for b in Js[s] + [np.eye(6)]])
        eng.set_body_coupling(nsw, nfil, np.concatenate([a.reshape(-1) for s in range(nsw) for a in As[s]]),
                              np.concatenate([b.reshape(-1) for s in range(nsw) for b in Bs[s]]),
                              np.concatenate([m.reshape(-1) for m in S0s]))
        e = rng.standard_normal(sum(sizes))
        eng.start(e)
        up = eng.find_update()
    finally:
        eng.close()
    want = np.concatenate([-0.1 * np.linalg.solve(S.assembled_jacobian(Js[s], S0s[s], As[s], Bs[s]),
                                                  e[s * (nfil * n6 + 6):(s + 1) * (nfil * n6 + 6)])
                           for s in range(nsw)])
    print("rel err", rel(up, want), flush=True)
    assert rel(up, want) < 1e-12


STEPS = 2


@pytest.fixture(scope="module")
def free_case():
    nfil, nseg, nb, R = 4, 20, 400, 10.0
    body = G.make_body(nfil, nseg, nb, R)
    dl, kb = 2.2, 1800.0
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    p = F.Params(nseg=nseg, dl=dl, kb=kb, kt=kb, dt=0.1, end_force=efm)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    refs = [body.fil_base[i] - body.centre for i in range(nfil)]
    N_ref = S.body_mobility_reference(O.dense_mobility(np.zeros(0), body.x_blobs), body.blob_refs)
    b = S.Body.at_rest(body.centre)
    fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (nseg, 1)), nseg, dl) for i in range(nfil)]
    runs = S.free_swimmer_loop(b, fils, refs, body.blob_refs, p, STEPS, 1e-8, pert, N_ref, sphere_radius=R)
    return dict(body=body, p=p, pert=pert, refs=refs, N_ref=N_ref, q0=q0, R=R, runs=runs)


@pytest.mark.parametrize("direct", [False, True])
def test_free_swimmer_elastic_step_matches_oracle_loop(free_case, direct):
    print(f"case free swimmer elastic step 4x20 + 400 blobs, tfs_elastic_step, direct {direct}", flush=True)
    c = free_case
    body, p, nb = c["body"], c["p"], c["body"].num_blobs
    nfil, nseg = body.num_fil, p.nseg
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    # the body mobility on the device (six free-body solves) agrees with the oracle's
    evh = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-13,
                            free_bodies=True)
    try:
        N_dev = evh.body_mobility(body.x_blobs, body.centre)[0]
    finally:
        evh.close()
    print("body mobility rel err", rel(N_dev, c["N_ref"]), flush=True)
    assert rel(N_dev, c["N_ref"]) < 1e-9
    model = FM.FilamentModel(1, nfil, nseg, dl=p.dl, kb=p.kb, kt=p.kt, dt=p.dt, end_force=p.end_force,
                             memory=L.MEMORY_DEVICE)
    ev = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                           memory=L.MEMORY_DEVICE, free_bodies=True)
    if direct:
        ev.set_preconditioner(BLOB_DIRECT)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
    f_blobs, nmob = dev(np.zeros(3 * nb)), dev(N_dev.reshape(-1))
    step = None
    iters, xs, bx = [], [], []
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(np.repeat(c["q0"][:, None, :], nseg, axis=1).reshape(-1)))
        model.set_body(dev(body.centre), dev([1.0, 0.0, 0.0, 0.0]), dev(np.array(c["refs"]).reshape(-1)), nb,
                       dev(body.blob_refs.reshape(-1)))
        model.set_perturbation(4, 12, dev(c["pert"].reshape(-1)))
        step = FM.ElasticStep(model, ev, eng, None, None, None, f_blobs, tol=1e-8, max_iter=400,
                              body_mobility=nmob, sphere_radius=c["R"])
        for nt in range(STEPS):
            st = step.step(nt)
            print("step", nt, st, flush=True)
            assert st["converged"]
            iters.append(st["broyden_iterations"])
            xs.append(model.read(FM.X))
            bx.append(model.read(FM.BODY_X))
    finally:
        if step is not None:
            step.close()
        model.close()
        ev.close()
        eng.close()
    ref = c["runs"]
    print("broyden iterations", iters, "oracle", [r["iters"] for r in ref],
          "max |dx|", [float(np.abs(a - r["x"]).max()) for a, r in zip(xs, ref)],
          "body", [float(np.abs(a - r["body"].x).max()) for a, r in zip(bx, ref)], flush=True)
    for s in range(STEPS):
        r = ref[s]
        assert abs(iters[s] - r["iters"]) <= max(2, r["iters"] // 10)
        assert np.abs(xs[s] - r["x"]).max() <= 1e-6
        assert np.abs(bx[s] - r["body"].x).max() <= 1e-8 + 1e-3 * np.abs(r["body"].x - body.centre).max()
