"""Pins the oracle's free swimmer (oracle/swimmer_oracle.py) on the CPU:

* the coupled J0^{-1} (body Schur complement, jacobian_inv_mult) is the
  inverse of the assembled block Jacobian [[diag J_i, A], [B, S0]];
* the body update keeps every filament anchored: base at X + R(q) ref, base
  rotation equal to the body's, robot-arm chain intact;
* the free blob solve is force- and torque-free (blobs and segments together)
  and moves the blobs rigidly;
* the body mobility of a sphere of blobs against Stokes drag (the reference's
  own self-check, mobility_solver.cpp:3017-3033);
* a free swimmer's implicit step converges, fewer Broyden iterations with the
  body's Schur complement than with an identity body block, and its converged
  state is force-free;
* the Broyden-for-everything form (blob forces as Broyden unknowns) reaches
  the trajectory of the form that solves for them at every evaluation."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import filament_oracle as F
from oracle import oracle as O
from oracle import swimmer_oracle as S
from timfilsimmod_amd import geometry as G


def quat(r):
    return np.roll(r.as_quat(), 1)


def random_swimmer(nfil=3, nseg=6, seed=0):
    rng = np.random.default_rng(seed)
    p = F.Params(nseg=nseg, end_force=40.0, dt=0.2, kt=1500.0)
    body = S.Body.at_rest(rng.standard_normal(3), quat(Rotation.random(random_state=seed)))
    body.u = rng.standard_normal(3) * 0.05
    body.q = F.qmul(F.lie_exp(body.u), body.qm1)
    body.Q_init = F.rot_mat(body.q)
    refs = [8.0 * v / np.linalg.norm(v) for v in rng.standard_normal((nfil, 3))]
    fils = []
    for i in range(nfil):
        q0 = quat(Rotation.random(random_state=seed + 10 + i))
        qs = [q0]
        for _ in range(nseg - 1):
            qs.append(F.qmul(F.lie_exp(0.05 * rng.standard_normal(3)), qs[-1]))
        f = F.Filament.at_rest(body.x + F.rot_mat(body.q) @ refs[i], np.array(qs), nseg, p.dl)
        f.lam = rng.standard_normal((nseg - 1, 3)) * 20
        f.tether = rng.standard_normal(3) * 10
        f.clamp = rng.standard_normal(3) * 10
        f.u[0] = body.u
        fils.append(f)
    return p, body, fils, refs, rng


@pytest.mark.parametrize("nt,sphere", [(0, 0.0), (3, 0.0), (3, 7.0)])
def test_coupled_inverse_is_the_inverse_of_the_assembled_jacobian(nt, sphere):
    p, body, fils, refs, rng = random_swimmer(seed=nt)
    forces = [F.internal_forces(f, p) for f in fils]
    Js = [F.approx_jacobian(f, forces[i], nt >= 1, p) for i, f in enumerate(fils)]
    N_ref = np.diag([0.02, 0.02, 0.02, 0.001, 0.001, 0.001]) + 1e-4 * rng.standard_normal((6, 6))
    S0, As, Bs = S.body_coupling(body, fils, refs, forces, nt, N_ref, p, sphere)
    inv = S.CoupledInverse(Js, S0, As, Bs)
    J = S.assembled_jacobian(Js, S0, As, Bs)
    for _ in range(3):
        x = rng.standard_normal(J.shape[0])
        y = inv.apply(x)
        assert np.abs(J @ y - x).max() <= 1e-10 * np.abs(x).max() * np.abs(J).max() * J.shape[0]
        assert np.allclose(y, np.linalg.solve(J, x), rtol=1e-9, atol=1e-12 * np.abs(y).max())
    # the coupling is not trivial: the body rows see the filaments
    assert all(np.abs(B).max() > 0 for B in Bs) and np.abs(S0 - np.eye(6)).max() > 0
    # clamping blocks: the Lie-algebra block's first column before the overwrite
    c1, c2 = S.clamping_blocks(fils[0], forces[0], nt >= 1, p)
    N = p.nseg
    assert np.allclose(As[0][3 * N:3 * N + 3, 3:], c1) and np.allclose(As[0][3 * N + 3:3 * N + 6, 3:], c2)
    assert not np.allclose(c1, Js[0][3 * N:3 * N + 3, 3 * N:3 * N + 3])


def test_body_update_keeps_filaments_anchored():
    p, body, fils, refs, rng = random_swimmer(seed=5)
    n = 6 * p.nseg * len(fils) + 6
    for _ in range(3):
        du = rng.standard_normal(n) * 0.01
        S.swimmer_update(body, fils, refs, du, p)
        R = F.rot_mat(body.q)
        assert np.allclose(body.q, F.qmul(F.lie_exp(body.u), body.qm1), atol=1e-15)
        for f, ref in zip(fils, refs):
            assert np.allclose(f.x[0], body.x + R @ ref, atol=1e-13)
            assert np.allclose(f.u[0], body.u, atol=1e-15)
            assert np.allclose(f.q[0], F.qmul(F.lie_exp(body.u), f.qm1[0]), atol=1e-15)
            t = np.array([F.frame(q)[0] for q in f.q])
            assert np.allclose(f.x[1:] - f.x[:-1], 0.5 * p.dl * (t[1:] + t[:-1]), atol=1e-12)
    # the initial guess of the body extrapolates linearly, then quadratically
    b = S.Body.at_rest(np.zeros(3))
    for k in range(4):
        S.body_initial_guess(b, k)
        b.x = np.array([1.0, 2.0, 3.0]) * (k + 1) ** 2 * 0.1     # the converged position of step k
    S.body_initial_guess(b, 4)
    assert np.allclose(b.x, np.array([1.0, 2.0, 3.0]) * 25 * 0.1, atol=1e-12)


def test_free_blob_solve_is_force_free_and_rigid():
    body = G.make_body(3, 6, 150, 6.0)
    xs, xb, fs = body.x_segs, body.x_blobs, body.f_segs
    X = np.array([0.3, -0.2, 0.1])
    vb = O.block("bs", xs, xb, fs, np.zeros(xb.size), seg_dof=6, bs_accumulate=False)
    fils = [F.Filament.at_rest(np.zeros(3), np.tile([1.0, 0, 0, 0], (6, 1)), 6, 2.2) for _ in range(3)]
    for i, f in enumerate(fils):
        f.x = xs.reshape(3, 6, 3)[i]
    sf = S.body_force(S.Body.at_rest(X), fils, [fs.reshape(3, 6, 6)[i] for i in range(3)])
    M = O.dense_mobility(np.zeros(0), xb)
    lam, V = S.free_blob_solve(M, xb, X, vb, sf)
    Fs, Ts = fs.reshape(-1, 6)[:, :3], fs.reshape(-1, 6)[:, 3:]
    lb = lam.reshape(-1, 3)
    total_f = Fs.sum(0) + lb.sum(0)
    total_t = (Ts + np.cross(xs.reshape(-1, 3) - X, Fs)).sum(0) + np.cross(xb.reshape(-1, 3) - X, lb).sum(0)
    assert np.abs(total_f).max() <= 1e-10 * np.abs(Fs).sum()
    assert np.abs(total_t).max() <= 1e-9 * np.abs(Fs).sum() * 10
    # blob velocities (segments' flow plus the blobs' own forces) are the rigid motion V + Omega x r
    v_blobs = vb + M @ lam
    r = xb.reshape(-1, 3) - X
    rigid = V[:3] + np.cross(V[3:], r)
    assert np.allclose(v_blobs.reshape(-1, 3), rigid, atol=1e-10 * np.abs(rigid).max())
    assert np.abs(V).max() > 1e-4


def test_body_mobility_of_a_sphere_of_blobs_is_stokes_drag():
    R = 10.0
    refs = G.fibonacci_sphere(800, R)
    N = S.body_mobility_reference(O.dense_mobility(np.zeros(0), refs.reshape(-1)), refs)
    # the reference's check: within a few percent of 1/(6 pi mu R), 1/(8 pi mu R^3)
    t, r = 1.0 / (6 * np.pi * R), 1.0 / (8 * np.pi * R ** 3)
    assert np.allclose(np.diag(N)[:3], t, rtol=0.05) and np.allclose(np.diag(N)[3:], r, rtol=0.1)
    assert np.abs(N[:3, 3:]).max() < 0.01 * np.sqrt(t * r) and np.allclose(N, N.T, atol=1e-12 * t)


def free_case(nfil=4, nseg=20, nb=400, R=10.0):
    body = G.make_body(nfil, nseg, nb, R)
    dl, kb = 2.2, 1800.0
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    p = F.Params(nseg=nseg, dl=dl, kb=kb, kt=kb, dt=0.1, end_force=efm)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = [np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal]
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    refs = [body.fil_base[i] - body.centre for i in range(nfil)]
    N_ref = S.body_mobility_reference(O.dense_mobility(np.zeros(0), body.x_blobs), body.blob_refs)

    def fresh():
        b = S.Body.at_rest(body.centre)
        fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (nseg, 1)), nseg, dl) for i in range(nfil)]
        return b, fils

    return dict(body=body, p=p, pert=pert, refs=refs, N_ref=N_ref, fresh=fresh, R=R, tol=1e-8)


def test_free_swimmer_step_converges_and_is_force_free():
    c = free_case()
    runs = {}
    for coupled in (True, False):
        b, fils = c["fresh"]()
        runs[coupled] = S.free_swimmer_loop(b, fils, c["refs"], c["body"].blob_refs, c["p"], 2, c["tol"], c["pert"],
                                            c["N_ref"], sphere_radius=c["R"], coupled=coupled)
    it = {k: [r["iters"] for r in v] for k, v in runs.items()}
    print("Broyden iterations, coupled", it[True], "identity body block", it[False])
    assert all(r["converged"] for v in runs.values() for r in v)
    assert sum(it[True]) < sum(it[False])
    # the same solution either way (to the tolerance) and the body moved
    for a, bb in zip(runs[True], runs[False]):
        assert np.abs(a["x"] - bb["x"]).max() < 1e-5
    last = runs[True][-1]
    assert np.abs(last["body"].x - c["body"].centre).max() > 1e-6
    # force-free: the blob forces balance the segments' forces
    Fseg = np.concatenate([f[:, :3] for f in last["forces"]]).sum(0)
    assert np.abs(Fseg + last["f_blobs"].reshape(-1, 3).sum(0)).max() <= 1e-8 * np.abs(Fseg).max() + 1e-9


def test_broyden_for_everything_reaches_the_gmres_form_trajectory():
    """Two solver forms of the same implicit step (blob forces solved for each
    evaluation, or carried as Broyden unknowns) converge to the same state."""
    c = free_case()
    body, p = c["body"], c["p"]
    runs = {}
    for mode in (False, True):
        q0 = [f.q[0] for f in c["fresh"]()[1]]
        fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (p.nseg, 1)), p.nseg, p.dl)
                for i in range(body.num_fil)]
        runs[mode] = S.held_still_loop(fils, body.x_blobs, p, 2, c["tol"], c["pert"], mode)
    print("Broyden iterations, GMRES form", [r["iters"] for r in runs[False]], "for everything",
          [r["iters"] for r in runs[True]])
    for a, b in zip(runs[False], runs[True]):
        assert a["converged"] and b["converged"]
        assert np.abs(a["x"] - b["x"]).max() < 1e-7
        assert np.linalg.norm(a["f_blobs"] - b["f_blobs"]) < 1e-2 * np.linalg.norm(a["f_blobs"])
    assert all(b["iters"] > a["iters"] for a, b in zip(runs[False], runs[True]))
