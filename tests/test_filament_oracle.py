"""Pins the oracle's Cosserat filament model (oracle/filament_oracle.py) on the
CPU, independently of how it is written:

* quaternion algebra against scipy's Rotation (product, frame, exp map,
  midpoint as the half-way slerp) and dexpinv_transpose as the transpose of
  dexpinv's matrix;
* the elastic moment of a uniformly bent and twisted filament against its
  closed form K * 4 sin(theta/4) / DL per body axis;
* force and moment balance of the internal forces (the constraint and elastic
  terms are internal: only the tether, clamp and end forces remain);
* the Lie derivative of the moment against finite differences (exact for
  aligned neighbours, first-order accurate in the bend otherwise);
* the whole approximate Jacobian against central differences of the
  self-mobility residual it linearises (exact at u = 0 for a straight
  filament, backward Euler and BDF2);
* the robot arm and the multipliers' extrapolation on closed forms."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation, Slerp

from oracle import broyden_oracle as BO
from oracle import filament_oracle as F


def rot(q):
    return Rotation.from_quat(np.roll(q, -1))


def quat(r):
    return np.roll(r.as_quat(), 1)


def test_quaternion_algebra_matches_scipy():
    rng = np.random.default_rng(0)
    for _ in range(20):
        r1, r2 = Rotation.random(random_state=rng), Rotation.random(random_state=rng)
        q1, q2 = quat(r1), quat(r2)
        assert np.allclose(rot(F.qmul(q1, q2)).as_matrix(), (r1 * r2).as_matrix(), atol=1e-14)
        assert np.allclose(F.rot_mat(q1), r1.as_matrix(), atol=1e-14)
        u = rng.standard_normal(3)
        assert np.allclose(rot(F.lie_exp(u)).as_matrix(), Rotation.from_rotvec(u).as_matrix(), atol=1e-14)
        # the midpoint is the half-way point of the shorter geodesic from q1 to
        # q2 (for quaternions in the same hemisphere, as neighbours are)
        q2h = q2 if q1 @ q2 >= 0 else -q2
        half = Slerp([0.0, 1.0], Rotation.concatenate([r1, r2]))([0.5])[0]
        assert np.allclose(F.rot_mat(F.midpoint(q1, q2h)), half.as_matrix(), atol=1e-12)
        # psi: d/d(eps) exp(eps d) q at 0
        d, eps = rng.standard_normal(3), 1e-4   # lie_exp is the identity below |u| = 1e-6
        fd = (F.qmul(F.lie_exp(eps * d), q1) - F.qmul(F.lie_exp(-eps * d), q1)) / (2 * eps)
        assert np.allclose(F.psi_mat(q1) @ d, fd, atol=1e-7)
        # left multiplication matrix
        assert np.allclose(F.left_mult_mat(q1) @ q2, F.qmul(q1, q2), atol=1e-15)
        # rcross(v) x = x cross v
        x = rng.standard_normal(3)
        assert np.allclose(F.rcross(u) @ x, np.cross(x, u), atol=1e-15)


def test_dexpinv_transpose_is_the_transpose():
    rng = np.random.default_rng(1)
    for scale in (1e-7, 0.3, 1.5):
        u = scale * rng.standard_normal(3)
        A = np.stack([F.dexpinv(u, e) for e in np.eye(3)], axis=1)
        At = np.stack([F.dexpinv_transpose(u, e) for e in np.eye(3)], axis=1)
        assert np.allclose(At, A.T, atol=1e-15)
        v = rng.standard_normal(3)
        assert np.allclose(F.dexpinv(u, v), BO.dexpinv(u, v), atol=1e-14)


@pytest.mark.parametrize("axis,k", [(0, "kt"), (1, "kb"), (2, "kb")])
def test_moment_of_uniform_bend_closed_form(axis, k):
    p = F.Params(nseg=2, dl=2.2, kb=1800.0, kt=1200.0)
    q1 = quat(Rotation.random(random_state=3))
    for theta in (1e-3, 0.2, 1.0):
        e = np.eye(3)[axis]
        q2 = F.qmul(q1, F.lie_exp(theta * e))          # turn about a body axis
        M = F.body_frame_moment(q1, q2, p)
        want = getattr(p, k) * 4.0 * np.sin(theta / 4.0) / p.dl * e
        assert np.allclose(M, want, rtol=1e-13, atol=1e-12)


def bent_filament(nseg, bend, seed, dl=2.2):
    rng = np.random.default_rng(seed)
    q0 = quat(Rotation.random(random_state=seed))
    qs = [q0]
    for _ in range(nseg - 1):
        qs.append(F.qmul(F.lie_exp(bend * rng.standard_normal(3)), qs[-1]))
    f = F.Filament.at_rest(rng.standard_normal(3) * 5, np.array(qs), nseg, dl)
    return f, rng


def test_internal_forces_balance():
    p = F.Params(nseg=12, end_force=37.0)
    f, rng = bent_filament(12, 0.3, 4)
    f.lam = rng.standard_normal((11, 3)) * 50
    f.tether = rng.standard_normal(3) * 20
    f.clamp = rng.standard_normal(3) * 20
    f.u[0] = rng.standard_normal(3) * 0.1
    out = F.internal_forces(f, p, perturbation={4: np.array([1.0, 2.0, 3.0])})
    tN = F.frame(f.q[-1])[0]
    total_f = out[:, :3].sum(0)
    assert np.allclose(total_f, f.tether - p.end_force * tN + [1.0, 2.0, 3.0], atol=1e-10)
    moment = (np.cross(f.x, out[:, :3]) + out[:, 3:]).sum(0)
    want = np.cross(f.x[0], f.tether) + F.dexpinv_transpose(f.u[0], f.clamp) + \
        np.cross(f.x[-1], -p.end_force * tN) + np.cross(f.x[4], [1.0, 2.0, 3.0])
    assert np.allclose(moment, want, atol=1e-8)
    # straight, unstressed, no multipliers: no forces at all
    g = F.Filament.at_rest(np.zeros(3), np.tile(f.q[0], (12, 1)), 12, p.dl)
    assert np.allclose(F.internal_forces(g, F.Params(nseg=12)), 0.0, atol=1e-9)


@pytest.mark.parametrize("upper", [False, True])
def test_lie_derivative_against_finite_differences(upper):
    p = F.Params(nseg=2, kt=1300.0)
    errs = []
    for bend in (0.0, 0.02, 0.04, 0.08):
        f, rng = bent_filament(2, bend, 5)
        q1, q2 = f.q
        D = F.body_frame_moment_lie_derivative(q1, q2, upper, p)
        fd = np.zeros((3, 3))
        h = 1e-5
        for c in range(3):
            d = h * np.eye(3)[c]
            if upper:
                mp = F.body_frame_moment(q1, F.qmul(F.lie_exp(d), q2), p)
                mm = F.body_frame_moment(q1, F.qmul(F.lie_exp(-d), q2), p)
            else:
                mp = F.body_frame_moment(F.qmul(F.lie_exp(d), q1), q2, p)
                mm = F.body_frame_moment(F.qmul(F.lie_exp(-d), q1), q2, p)
            fd[:, c] = (mp - mm) / (2 * h)
        errs.append(np.abs(D - fd).max() / np.abs(fd).max())
    assert errs[0] < 1e-8
    # first order in the bend: halving it halves the error
    assert errs[1] < 0.02 and errs[2] / errs[1] < 2.3 and errs[3] / errs[2] < 2.3


def jacobian_fd(f, p, bdf2):
    N = p.nseg
    h = 1e-5          # above lie_exp's identity threshold |u| < 1e-6
    base = f.copy()
    forces = F.internal_forces(base, p)
    J = F.approx_jacobian(base, forces, bdf2, p)
    fd = np.zeros_like(J)
    for c in range(6 * N):
        cols = []
        for s in (1.0, -1.0):
            g = base.copy()
            dz = np.zeros(6 * N)
            dz[c] = s * h
            F.update(g, dz, p)
            cols.append(F.self_mobility_residual(g, base.xm1, base.xm2, base.um1, bdf2, p))
        fd[:, c] = (cols[0] - cols[1]) / (2 * h)
    return J, fd


@pytest.mark.parametrize("bdf2", [False, True])
def test_approximate_jacobian_is_the_self_mobility_jacobian(bdf2):
    # straight (tilted) filament at u = 0 with multipliers and an end force:
    # the approximate Jacobian is exact there
    N = 6
    p = F.Params(nseg=N, end_force=55.0, dt=0.3, kt=1500.0)
    q0 = quat(Rotation.random(random_state=7))
    f = F.Filament.at_rest(np.array([1.0, -2.0, 0.5]), np.tile(q0, (N, 1)), N, p.dl)
    rng = np.random.default_rng(8)
    f.lam = rng.standard_normal((N - 1, 3)) * 30
    f.tether = rng.standard_normal(3) * 10
    f.clamp = rng.standard_normal(3) * 10
    J, fd = jacobian_fd(f, p, bdf2)
    scale = np.abs(fd).max()
    assert np.abs(J - fd).max() <= 2e-7 * scale


def test_approximate_jacobian_near_straight_bent_filament():
    # a bent filament: the differences come only from the moment's Lie
    # derivative approximation and shrink with the bend
    N = 6
    p = F.Params(nseg=N, end_force=20.0, dt=0.3)
    errs = []
    for bend in (0.005, 0.01, 0.02):
        f, rng = bent_filament(N, bend, 9)
        f.lam = rng.standard_normal((N - 1, 3)) * 30
        J, fd = jacobian_fd(f, p, False)
        errs.append(np.abs(J - fd).max() / np.abs(fd).max())
    # first order in the bend (0.35 x the bend per segment here)
    assert errs[0] < 5e-3 and 1.8 < errs[1] / errs[0] < 2.2 and 1.8 < errs[2] / errs[1] < 2.2


def test_robot_arm_and_update_closed_forms():
    N, dl = 5, 2.2
    q0 = quat(Rotation.from_rotvec([0.0, 0.0, 0.4]))
    f = F.Filament.at_rest(np.array([1.0, 2.0, 3.0]), np.tile(q0, (N, 1)), N, dl)
    t = F.frame(q0)[0]
    assert np.allclose(f.x, np.array([1.0, 2.0, 3.0]) + np.outer(np.arange(N) * dl, t), atol=1e-14)
    # rotating segment 3 (and only it) moves segments 3.. by the chord rule
    p = F.Params(nseg=N, dl=dl)
    dz = np.zeros(6 * N)
    dz[3 * N + 9:3 * N + 12] = [0.0, 0.0, 0.5]
    g = f.copy()
    F.update(g, dz, p)
    t3 = F.frame(F.qmul(F.lie_exp(np.array([0.0, 0.0, 0.5])), q0))[0]
    assert np.allclose(g.x[:3], f.x[:3], atol=1e-14)
    assert np.allclose(g.x[3], f.x[2] + 0.5 * dl * (t + t3), atol=1e-14)
    assert np.allclose(g.x[4], g.x[3] + 0.5 * dl * (t3 + t), atol=1e-14)


def test_multiplier_extrapolation_is_exact_on_polynomials():
    N = 4
    p = F.Params(nseg=N)
    f = F.Filament.at_rest(np.zeros(3), np.tile([1.0, 0, 0, 0], (N, 1)), N, p.dl)
    seq = lambda n: np.full((N - 1, 3), 2.0 + 0.5 * n - 0.25 * n * n)      # quadratic in n
    for nt in range(1, 6):
        f.lam = seq(nt - 1)                     # the converged multipliers of step nt-1
        F.initial_guess(f, nt, p)
        if nt >= 3:
            assert np.allclose(f.lam, seq(nt), atol=1e-12)
    lin = lambda n: np.full((N - 1, 3), 1.0 + 3.0 * n)
    g = F.Filament.at_rest(np.zeros(3), np.tile([1.0, 0, 0, 0], (N, 1)), N, p.dl)
    for nt in (1, 2):
        g.lam = lin(nt - 1)
        F.initial_guess(g, nt, p)
    assert np.allclose(g.lam, lin(2), atol=1e-12)


# ---- the base segment's other builds (CILIA_TYPE 1 and 2) --------------------------
def test_driven_base_forces_balance():
    p = F.Params(nseg=10, base="driven")
    f, rng = bent_filament(10, 0.3, 14)
    f.lam = rng.standard_normal((9, 3)) * 50
    f.tether = rng.standard_normal(3) * 20
    f.clamp = rng.standard_normal(3) * 20          # ignored by a driven base
    f.drive = rng.standard_normal(3) * 40
    out = F.internal_forces(f, p)
    assert np.allclose(out[:, :3].sum(0), f.tether, atol=1e-10)
    moment = (np.cross(f.x, out[:, :3]) + out[:, 3:]).sum(0)
    assert np.allclose(moment, np.cross(f.x[0], f.tether) + f.drive, atol=1e-8)


@pytest.mark.parametrize("bdf2", [False, True])
def test_driven_base_jacobian_is_the_self_mobility_jacobian(bdf2):
    # a free, driven base: its Lie-algebra element is the unknown in the
    # clamp's slot and moves every later segment through the robot arm
    N = 6
    p = F.Params(nseg=N, dt=0.3, kt=1500.0, base="driven")
    q0 = quat(Rotation.random(random_state=17))
    f = F.Filament.at_rest(np.array([0.5, 1.0, -2.0]), np.tile(q0, (N, 1)), N, p.dl)
    rng = np.random.default_rng(18)
    f.lam = rng.standard_normal((N - 1, 3)) * 30
    f.tether = rng.standard_normal(3) * 10
    f.drive = rng.standard_normal(3) * 25
    J, fd = jacobian_fd(f, p, bdf2)
    scale = np.abs(fd).max()
    assert np.abs(J - fd).max() <= 2e-7 * scale
    # the base column is not the clamp's: rows of later segments' positions see it
    assert np.abs(J[3:3 * N, 3 * N:3 * N + 3]).max() > 0.1
    # a planted sign flip of the robot-arm column is seen
    Jbad = J.copy()
    Jbad[3:3 * N, 3 * N:3 * N + 3] *= -1
    assert np.abs(Jbad - fd).max() > 1e-2 * scale


def test_driven_base_update_turns_the_whole_filament():
    N, dl = 5, 2.2
    p = F.Params(nseg=N, dl=dl, base="driven")
    q0 = quat(Rotation.from_rotvec([0.0, 0.3, 0.0]))
    f = F.Filament.at_rest(np.array([1.0, 2.0, 3.0]), np.tile(q0, (N, 1)), N, dl)
    t = F.frame(q0)[0]
    dz = np.zeros(6 * N)
    dz[3 * N:3 * N + 3] = [0.0, 0.0, 0.4]           # turn the base segment only
    g = f.copy()
    F.update(g, dz, p)
    t0 = F.frame(F.qmul(F.lie_exp(np.array([0.0, 0.0, 0.4])), q0))[0]
    assert np.allclose(g.u[0], [0.0, 0.0, 0.4]) and np.allclose(g.clamp, 0.0)
    assert np.allclose(g.x[0], f.x[0], atol=1e-14)
    for i in range(1, N):
        assert np.allclose(g.x[i], f.x[0] + 0.5 * dl * (t0 + t) + (i - 1) * dl * t, atol=1e-13)
    # the clamped build puts the same slot into the clamp multiplier
    h = f.copy()
    F.update(h, dz, F.Params(nseg=N, dl=dl))
    assert np.allclose(h.clamp, [0.0, 0.0, 0.4]) and np.allclose(h.x, f.x, atol=1e-14)


def test_constant_base_rotation_turns_the_base_every_step():
    N = 6
    p = F.Params(nseg=N)
    q0 = quat(Rotation.random(random_state=23))
    f = F.Filament.at_rest(np.zeros(3), np.tile(q0, (N, 1)), N, p.dl)
    t, n, _ = F.frame(q0)
    theta = 0.05
    f.u0 = theta * t                                 # BASE_ROTATION_RATE DT about the tangent
    x0 = f.x.copy()
    for nt in range(4):
        F.initial_guess(f, nt, p)
        rot_k = Rotation.from_rotvec((nt + 1) * theta * t)
        tk, nk, _ = F.frame(f.q[0])
        assert np.allclose(tk, t, atol=1e-13) and np.allclose(nk, rot_k.apply(n), atol=1e-12)
        assert np.allclose(f.u[0], theta * t) and np.allclose(f.x, x0, atol=1e-12)
        for i in range(1, N):                        # the others stay at rest (no forces acted)
            assert np.allclose(f.q[i], q0, atol=1e-14)


def test_constant_base_rotation_clamp_columns_of_the_jacobian():
    # with the base turned (u_0 != 0) the clamp block fac dexpinv dexpinv^T is
    # the series of the exact derivative to fourth order in |u_0|; the rest of
    # J0 is first order in the twist the turn puts between segments 0 and 1
    N = 5
    errs = []
    for theta in (0.02, 0.04):
        p = F.Params(nseg=N, dt=0.25)
        q0 = quat(Rotation.random(random_state=29))
        f = F.Filament.at_rest(np.zeros(3), np.tile(q0, (N, 1)), N, p.dl)
        f.u0 = theta * F.frame(q0)[0]
        F.initial_guess(f, 0, p)
        rng = np.random.default_rng(30)
        f.lam = rng.standard_normal((N - 1, 3)) * 30
        f.clamp = rng.standard_normal(3) * 10
        J, fd = jacobian_fd(f, p, False)
        scale = np.abs(fd).max()
        cl = slice(3 * N, 3 * N + 3)
        assert np.abs(J[:, cl] - fd[:, cl]).max() <= 1e-7 * np.abs(fd[:, cl]).max()
        errs.append(np.abs(J - fd).max() / scale)
    assert errs[0] < 2e-3 and 1.7 < errs[1] / errs[0] < 2.5
