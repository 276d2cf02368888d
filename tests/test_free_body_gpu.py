"""Free bodies in the saddle-point solve (tfs_solver with free_bodies; the
reference's build without PRESCRIBED_BODY_VELOCITIES): the operator against
the oracle's dense matrix (one and two swimmers, with and without filaments),
the preconditioner as the exact inverse of [D -K; -K^T 0] with the body
columns, solves against dense direct solves in FP64 and mixed precision with
either preconditioner, and a rigid sphere of blobs pushed and twisted by an
external force and torque moving as a sphere of one hydrodynamic radius."""
import numpy as np
import pytest

from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import (PRECOND_BLOCKS, PRECOND_REFERENCE, SOLVE_FP64, SOLVE_MIXED, Solver,
                                     beat_setup, solver_desc)

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def case(nsw, nfil, nseg, nblob, **kw):
    """nsw swimmers side by side, each nfil filaments and nblob blobs."""
    xs, xb, kp, kt, xbod = [], [], [], [], []
    for w in range(nsw):
        centre = (25.0 * w, 3.0 * w, -2.0 * w)
        body = G.make_body(nfil, nseg, nblob, 8.0, centre=centre)
        if nfil:
            geo = SO.beat_geometry(beat_setup(body), nseg, B.BEAT_A, B.BEAT_B)
            xs.append(geo["x_segs"])
            kp.append(geo["kphi"])
            kt.append(geo["ktheta"])
        xb.append(body.x_blobs)
        xbod.append(np.array(centre) + np.array([0.3, -0.2, 0.1]))   # any reference point
    cat = lambda a: np.concatenate(a) if a else np.zeros(0)
    xs, xb, kp, kt, xbod = cat(xs), cat(xb), cat(kp), cat(kt), cat(xbod)
    dyn = nfil > 0
    s = Solver(solver_desc(nsw * nfil, nseg, nsw * nblob, num_swim=nsw, free_bodies=True, dynamic_phase=dyn,
                           dynamic_angle=dyn, **kw))
    s.set_geometry(xs if nfil else None, xb if nblob else None, kp if nfil else None, kt if nfil else None)
    s.set_bodies(xbod)
    A = SO.saddle_matrix(xs, xb, kp, kt, nsw * nfil, nseg, x_bodies=xbod, num_swim=nsw, dyn_phase=dyn,
                         dyn_angle=dyn)
    P = SO.precond_matrix(kp, kt, nsw * nfil, nseg, nsw * nblob, x_segs=xs, x_blobs=xb, x_bodies=xbod, num_swim=nsw,
                          dyn_phase=dyn, dyn_angle=dyn)
    return s, A, P


@pytest.mark.parametrize("nsw,nfil,nseg,nblob", [(1, 3, 6, 40), (2, 4, 20, 150), (1, 0, 2, 200), (2, 2, 5, 0)])
def test_operator_and_preconditioner(nsw, nfil, nseg, nblob):
    print(f"case free body nsw={nsw} nfil={nfil} nseg={nseg} nblob={nblob}", flush=True)
    s, A, P = case(nsw, nfil, nseg, nblob)
    try:
        assert s.n == A.shape[0]
        v = np.random.default_rng(2).normal(size=s.n)
        av = s.system_mult(v)
        assert rel(av, A @ v) <= 1e-12
        assert rel(s.system_mult(v, L.PRECISION_FP32_ACC64), A @ v) <= 1e-6
        pv = s.precondition(v)
        assert rel(P @ pv, v) <= 1e-11
    finally:
        s.close()


@pytest.mark.parametrize("mode", [SOLVE_FP64, SOLVE_MIXED])
@pytest.mark.parametrize("precond", [PRECOND_REFERENCE, PRECOND_BLOCKS])
def test_solve_matches_dense(mode, precond):
    print(f"case free body solve mode={mode} precond={precond}", flush=True)
    s, A, P = case(2, 4, 20, 150, mode=mode, restart=50, tol=1e-10, max_cycles=200, inner_tol=1e-3,
                   max_refinements=20)
    try:
        s.set_preconditioner(precond)
        rhs = np.random.default_rng(7).normal(size=s.n)
        x, st = s.solve(rhs)
        print(st, flush=True)
        assert st["converged"] == 1
        assert rel(A @ x, rhs) <= 2e-10
        assert rel(x, np.linalg.solve(A, rhs)) <= 1e-6
    finally:
        s.close()


def test_rigid_sphere_under_force_and_torque():
    """Body rows -K^T lambda = -F_ext (force and torque balance with an external
    load): the sphere translates with |V| = |F|/(6 pi mu R_h) and turns with
    |Omega| = |T|/(8 pi mu R_h^3), one R_h between R and R + a for both."""
    print("case free rigid sphere", flush=True)
    R, nb = 12.0, 800
    body = G.make_body(0, 0, nb, R)
    s = Solver(solver_desc(0, 1, nb, free_bodies=True, dynamic_phase=False, dynamic_angle=False, tol=1e-11,
                           restart=60, max_cycles=200))
    try:
        s.set_geometry(None, body.x_blobs)
        s.set_bodies(np.zeros(3))
        F, T = np.array([0.4, -0.1, 0.7]), np.array([-0.3, 0.6, 0.2])
        rhs = np.zeros(s.n)
        off = 3 * nb
        rhs[off:off + 3], rhs[off + 3:off + 6] = -F, -T
        x, st = s.solve(rhs)
        V, W = x[off:off + 3], x[off + 3:off + 6]
        lam = x[:off].reshape(-1, 3)
        np.testing.assert_allclose(lam.sum(0), F, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(np.cross(body.x_blobs.reshape(-1, 3), lam).sum(0), T, rtol=1e-8, atol=1e-9)
        rh_drag = np.linalg.norm(F) / (6 * np.pi * np.linalg.norm(V))
        rh_rot = (np.linalg.norm(T) / (8 * np.pi * np.linalg.norm(W))) ** (1 / 3)
        print("hydrodynamic radius", rh_drag, rh_rot, flush=True)
        assert np.linalg.norm(np.cross(V, F)) <= 2e-3 * np.linalg.norm(V) * np.linalg.norm(F)
        assert np.linalg.norm(np.cross(W, T)) <= 2e-3 * np.linalg.norm(W) * np.linalg.norm(T)
        assert R <= rh_drag <= R + 1.0 and R <= rh_rot <= R + 1.0
        assert abs(rh_drag - rh_rot) <= 0.02 * R
    finally:
        s.close()


def test_swimming_beat_matches_oracle_loop():
    """A swimming body in the prescribed-beat loop (free_bodies): each step the
    body moves by the last solve's V, Omega (cilia_sim_main.cu:677-693), its
    filaments and blobs with it; against the same loop on the oracle (dense
    free-body solves, the beat geometry of the moved frames)."""
    from scipy.spatial.transform import Rotation
    from timfilsimmod_amd.solver import BeatSimulation
    print("case swimming beat", flush=True)
    body = G.make_body(6, 20, 200, 8.0, centre=(1.0, -2.0, 0.5))
    steps, dt = 4, 1.0 / 300.0
    sim = BeatSimulation(body, tol=1e-11, restart=60, max_cycles=50, free_body=True)
    su = sim.setup
    try:
        gpu = []
        for _ in range(steps):
            st = sim.step()
            assert st["converged"] == 1
            x, q, v = sim.body()
            gpu.append((sim.state()["phase"].copy(), x, q, v))
    finally:
        sim.close()
    # oracle loop
    nf, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    X0 = np.array(body.centre)
    X, qb = X0.copy(), np.array([1.0, 0.0, 0.0, 0.0])
    phase, angle = su.phase0.copy(), su.angle0.copy()
    base_ref = su.bases.reshape(nf, 3) - X0
    frames_ref = su.frames.reshape(nf, 3, 3)
    blob_ref = su.x_blobs.reshape(nb, 3) - X0
    sol = None
    for step in range(steps):
        if step > 0:
            off = 3 * (nf * nseg + nb)
            V, W = sol[off:off + 3], sol[off + 3:off + 6]
            phase = phase + dt * sol[off + 6:off + 6 + nf]
            angle = angle + dt * sol[off + 6 + nf:off + 6 + 2 * nf]
            X = X + dt * V
            e = Rotation.from_rotvec(dt * W)
            qb = np.roll((e * Rotation.from_quat(np.roll(qb, -1))).as_quat(), 1)
            if qb[0] * gpu[step][2][0] < 0:
                qb = -qb
        R = Rotation.from_quat(np.roll(qb, -1)).as_matrix()
        moved = type(su)(**{**su.__dict__, "bases": (X + base_ref @ R.T).reshape(-1),
                            "frames": (frames_ref @ R.T).reshape(-1), "x_blobs": (X + blob_ref @ R.T).reshape(-1)})
        geo = SO.beat_geometry(moved, nseg, B.BEAT_A, B.BEAT_B, phase=phase, angle=angle)
        xb = moved.x_blobs
        A = SO.saddle_matrix(geo["x_segs"], xb, geo["kphi"], geo["ktheta"], nf, nseg, x_bodies=X)
        r0 = SO.rhs_ref(moved, nseg, nb, geo["x_segs"], xb, X, np.zeros(6), phase=phase, angle=angle)
        off = 3 * (nf * nseg + nb)
        rhs = np.concatenate([r0[:off], np.zeros(6), r0[off:]])
        sol = np.linalg.solve(A, rhs)
        gp, gx, gq, gv = gpu[step]
        np.testing.assert_allclose(gp, phase, rtol=0, atol=1e-9)
        np.testing.assert_allclose(gx, X, rtol=0, atol=1e-9)
        np.testing.assert_allclose(gq, qb, rtol=0, atol=1e-9)
        off = 3 * (nf * nseg + nb)
        assert np.linalg.norm(gv - sol[off:off + 6]) <= 1e-7 * np.linalg.norm(sol[off:off + 6])
    assert np.linalg.norm(gpu[-1][1] - X0) > 0   # it swam
