"""GPU parity of the elastic-filament compute_velocities sequence
(tfs_velocities_*, include/tfs_solver.h) against a dense FP64 solve built from
the oracle's grand mobility.

Bars: FP64 mode with the blob solve at tol 1e-11 agrees with the dense
solution to 1e-8 relative (the solve tolerance times a modest condition
factor); mixed mode (FP32-tile segment blocks, FP32-inner blob GMRES with FP64
refinement) to 1e-6.  Physical pin: with no filaments, a rigid sphere of 2,000
blobs (R = 20a) pulled at V has total blob force 6 pi mu R_h V and, rotated at
Omega, total torque 8 pi mu R_h^3 Omega with one hydrodynamic radius R_h in
[R, R + a] (the reference's own Stokes-drag check, mobility_solver.cpp:3017-3033,
reports the percentage deviation from R_h = R)."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import SOLVE_FP64, SOLVE_MIXED, ElasticVelocities

pytestmark = pytest.mark.gpu


def _bodies(nswim, nfil, nseg, nblob, radius):
    parts = [G.make_body(nfil, nseg, nblob, radius, centre=(3.0 * radius * k, 0.0, 0.0), force_seed=10 + k,
                         perturb_seed=20 + k) for k in range(nswim)]
    cat = lambda name: np.concatenate([getattr(b, name) for b in parts]) if parts else np.zeros(0)
    return (cat("x_segs"), cat("f_segs"), cat("x_blobs"), np.concatenate([b.centre for b in parts]))


def _dense(x_segs, f_segs, x_blobs, x_bodies, v_bodies, blobs_per_swim):
    ns, nb = x_segs.size // 3, x_blobs.size // 3
    M = O.dense_mobility(x_segs, x_blobs, seg_dof=6)
    s6 = 6 * ns
    Mss, Msb, Mbs, Mbb = M[:s6, :s6], M[:s6, s6:], M[s6:, :s6], M[s6:, s6:]
    xb = x_blobs.reshape(-1, 3)
    body = np.arange(nb) // blobs_per_swim
    r = xb - x_bodies.reshape(-1, 3)[body]
    V = v_bodies.reshape(-1, 6)[body]
    kv = V[:, :3] + np.cross(V[:, 3:], r)
    rhs = kv.reshape(-1) - (Mbs @ f_segs if ns else 0.0)
    lam = np.linalg.solve(Mbb, rhs)
    v = Mss @ f_segs + Msb @ lam if ns else np.zeros(0)
    return v, lam


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("mode,bar", [(SOLVE_FP64, 1e-8), (SOLVE_MIXED, 1e-6)])
@pytest.mark.parametrize("nswim,nfil,nseg,nblob,radius", [(1, 6, 8, 300, 8.0), (2, 3, 5, 150, 6.0),
                                                          (1, 16, 20, 2000, 20.0)])
def test_elastic_velocities_match_dense_solve(mode, bar, nswim, nfil, nseg, nblob, radius):
    print(f"case velocities mode={mode} nswim={nswim} nfil={nfil} nseg={nseg} nblob={nblob}", flush=True)
    xs, fs, xb, xbody = _bodies(nswim, nfil, nseg, nblob, radius)
    vbody = np.random.default_rng(4).normal(size=6 * nswim) * 0.1
    ref_v, ref_lam = _dense(xs, fs, xb, xbody, vbody, nblob)
    ev = ElasticVelocities(nswim, nfil * nseg, nblob, mode=mode, restart=30, max_cycles=200, tol=1e-11,
                           inner_tol=1e-3, max_refinements=20)
    try:
        v, lam, st = ev.compute(xs, fs, xb, xbody, vbody)
        print(st, _rel(v, ref_v), _rel(lam, ref_lam))
        assert st["converged"] == 1
        assert _rel(v, ref_v) <= bar
        assert _rel(lam, ref_lam) <= 100 * bar
        # warm start: the same inputs again need (almost) no iterations
        v2, lam2, st2 = ev.compute(xs, fs, xb, xbody, vbody)
        assert st2["iterations"] <= 10 and _rel(v2, ref_v) <= bar
    finally:
        ev.close()


def test_rigid_sphere_drag_and_torque():
    """The blob shell acts as a sphere of one hydrodynamic radius R_h, between
    R and R + a: the drag gives R_h = |F| / (6 pi mu |V|), the torque gives
    R_h^3 = |T| / (8 pi mu |Omega|); both parallel to the motion and equal to
    within 1%."""
    print("case velocities stokes drag", flush=True)
    R, nb = 20.0, 2000
    body = G.make_body(0, 0, nb, R)
    ev = ElasticVelocities(1, 0, nb, restart=30, max_cycles=200, tol=1e-10)
    try:
        V = np.array([0.3, -0.2, 0.5])
        _, lam, st = ev.compute(np.zeros(0), np.zeros(0), body.x_blobs, np.zeros(3),
                                np.concatenate([V, np.zeros(3)]))
        F = lam.reshape(-1, 3).sum(axis=0)
        W = np.array([0.0, 0.1, -0.05])
        ev.f_blobs[:] = 0.0
        _, lam, st = ev.compute(np.zeros(0), np.zeros(0), body.x_blobs, np.zeros(3),
                                np.concatenate([np.zeros(3), W]))
        T = np.cross(body.x_blobs.reshape(-1, 3), lam.reshape(-1, 3)).sum(axis=0)
        r_drag = np.linalg.norm(F) / (6 * np.pi * np.linalg.norm(V))
        r_rot = (np.linalg.norm(T) / (8 * np.pi * np.linalg.norm(W))) ** (1 / 3)
        print("hydrodynamic radius from drag", r_drag, "from torque", r_rot)
        assert np.linalg.norm(np.cross(F, V)) <= 1e-3 * np.linalg.norm(F) * np.linalg.norm(V)
        assert np.linalg.norm(np.cross(T, W)) <= 1e-3 * np.linalg.norm(T) * np.linalg.norm(W)
        assert R <= r_drag <= R + 1.0 and R <= r_rot <= R + 1.0
        assert abs(r_drag - r_rot) <= 0.01 * R
    finally:
        ev.close()


@pytest.mark.parametrize("nfil,nb", [(1, 10), (4, 200), (16, 2000)])
def test_direct_blob_solve_matches_dense_oracle(nfil, nb):
    """TFS_BLOB_DIRECT (dense inverse of the blob mobility by blocked
    Gauss-Jordan on the device, one product per solve) against an LU solve
    of the oracle's dense M_bb; sizes below, at and above one 64-row tile."""
    import scipy.linalg as sla
    from timfilsimmod_amd.solver import BLOB_DIRECT
    print(f"case direct blob solve nfil={nfil} nb={nb}", flush=True)
    body = G.make_body(nfil, 8, nb, 6.0 + nb ** 0.5 * 0.45)
    ns = nfil * 8
    ev = ElasticVelocities(1, ns, nb)
    try:
        ev.set_preconditioner(BLOB_DIRECT)
        f = body.f_segs
        V = np.array([0.1, -0.2, 0.05, 0.01, 0.02, -0.03])
        for rep in range(2):          # the second call reuses the stored inverse
            vs, lam, st = ev.compute(body.x_segs, f, body.x_blobs, body.centre, V)
        vb = O.block("bs", body.x_segs, body.x_blobs, f, np.zeros(3 * nb), seg_dof=6, bs_accumulate=False)
        r = body.x_blobs.reshape(-1, 3) - body.centre
        kv = V[:3] + np.cross(V[3:], r)
        want_lam = sla.solve(O.dense_mobility(np.zeros(0), body.x_blobs), -vb + kv.reshape(-1), assume_a="pos")
        vs_o = O.block("ss", body.x_segs, body.x_blobs, f, want_lam, seg_dof=6)
        vs_o = O.block("sb", body.x_segs, body.x_blobs, f, want_lam, vs_o, seg_dof=6)
        assert np.linalg.norm(lam - want_lam) <= 1e-10 * np.linalg.norm(want_lam)
        assert np.linalg.norm(vs - vs_o) <= 1e-10 * np.linalg.norm(vs_o)
        assert st["converged"] == 1 and st["iterations"] == 0
    finally:
        ev.close()
