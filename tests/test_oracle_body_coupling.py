"""Pins the free swimmer's body coupling of J0 (oracle/swimmer_oracle.py:
body_coupling, restating swimmer.cpp:688-880) and the blob forces'
start-of-step extrapolation (blob_force_guess, rigid_body.cpp:447-478) on the CPU.

J0 = [[diag J_i, A_i], [B_i, S0]] linearises the residual of one swimmer (the
Broyden vector: NFIL filament blocks of 6 NSEG, then the body's position and
Lie-algebra element) for a model in which every segment moves with its
self-mobility (v = F / 6 pi mu a, omega = T / 8 pi mu a^3, the model the
filament blocks J_i linearise) and the body with its 6 x 6 mobility N, rotated
to the step's initial orientation, applied to the total force and torque the
filaments exert on it; optionally (a sphere of radius R) plus the RPY coupling
between the sphere and each segment.  Here that residual is evaluated directly
(swimmer_update, internal_forces, the Euler/BDF2 rows of compute_errors) and
differentiated by central differences; the sphere-segment coupling comes from
the oracle's pair tensor (oracle.rpy_pair, pinned in test_oracle_pair.py).

At a state with straight filaments along the body normal and the body's Lie
element zero, the entries of A, B and S0 are exact derivatives, except for four
simplifications the reference writes into J0 and the oracle keeps; each is
checked as the exact closed-form difference it makes:

* A, position rows of segment 0 against the body's rotation: the reference
  moves every segment as if it sat 0.5 DL beyond the anchor
  (rcross((0.5 DL + |d|) d/|d|), swimmer.cpp:860-866), right for segments
  1..NSEG-1, off by rcross(0.5 DL d/|d|) for segment 0;
* A, Lie rows of segment 0 (elastic_clamping_block1): the Lie-algebra block's
  column of u_0 before the clamp overwrites it (filament.cpp:1883-1884) leaves
  out how the clamp torque dexpinv^T(u_0, clamp) turns with u_0: the
  derivative is larger by dt Tfac rcross(clamp)/2 (dt -> 2 dt/3 in BDF2);
* S0, the clamp term of the body's torque derivative (swimmer.cpp:712-714):
  +rcross(clamp)/2 where the torque -dexpinv^T(u_b, clamp) the body takes
  gives -rcross(clamp)/2, so S0 differs by dt_fac N_rr rcross(sum of clamps);
* S0, its own Lie-algebra term (swimmer.cpp:690-697): the exact Jacobian of
  u - |dt_fac| dexpinv(u, Omega) at Omega = -u / |dt_fac| (the reference's
  cross_body_omega = -body_lie_cross / dt_fac), where the consistent state of
  the residual is Omega = +u / |dt_fac|;
* the sphere term (swimmer.cpp:813-862): the translation-rotation blocks carry
  the opposite sign of the pair tensor's for the body as target (the reference
  forms rcross(r_hat) with r_hat from the body to the segment).

A deliberate sign flip in any block makes the checks fail (planted on the
oracle's output below).
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import filament_oracle as F
from oracle import oracle as O
from oracle import swimmer_oracle as S

PI = np.pi
DT = 0.2


def quat(r):
    return np.roll(r.as_quat(), 1)


def along(v):
    return quat(Rotation.align_vectors([v / np.linalg.norm(v)], [[1.0, 0.0, 0.0]])[0])


def swimmer(seed, nfil=2, nseg=5, lam=True, tether=True, clamp=True, end_force=30.0, kb=1800.0):
    """Straight filaments along the body normal at their anchors, every Lie
    element zero (the base segments turn with the body), random multipliers,
    and a body mobility without translation-rotation coupling (a sphere's)."""
    rng = np.random.default_rng(seed)
    p = F.Params(nseg=nseg, end_force=end_force, dt=DT, kb=kb, kt=0.8 * kb)
    b = S.Body.at_rest(rng.standard_normal(3), quat(Rotation.random(random_state=seed)))
    b.xm1 = b.x - 0.01 * rng.standard_normal(3)
    b.xm2 = b.xm1 - 0.01 * rng.standard_normal(3)
    b.um1 = 0.01 * rng.standard_normal(3)
    refs = [8.0 * v / np.linalg.norm(v) for v in rng.standard_normal((nfil, 3))]
    fils = []
    for i in range(nfil):
        d = b.Q_init @ refs[i]
        f = F.Filament.at_rest(b.x + d, np.tile(along(d), (nseg, 1)), nseg, p.dl)
        f.xm1 = f.x - 0.01 * rng.standard_normal(f.x.shape)
        f.xm2 = f.xm1 - 0.01 * rng.standard_normal(f.x.shape)
        f.um1 = 0.01 * rng.standard_normal(f.u.shape)
        f.lam = rng.standard_normal((nseg - 1, 3)) * 20 if lam else f.lam
        f.tether = rng.standard_normal(3) * 10 if tether else f.tether
        f.clamp = rng.standard_normal(3) * 10 if clamp else f.clamp
        fils.append(f)
    G1, G2 = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
    N_ref = np.zeros((6, 6))
    N_ref[:3, :3] = 0.02 * np.eye(3) + 0.002 * G1 @ G1.T
    N_ref[3:, 3:] = 0.001 * np.eye(3) + 1e-4 * G2 @ G2.T
    return p, b, fils, refs, N_ref


def sphere_coupling(x_body, R, x_seg, p):
    """6 x 6 velocity of the sphere (target) per force/torque on the segment
    (source), the oracle's RPY pair tensor, translation-rotation blocks negated
    as the reference orients them."""
    M = np.column_stack([O.rpy_pair(6, 6, e, x_body, R, x_seg, p.a, p.mu) for e in np.eye(6)])
    M[:3, 3:] *= -1.0
    M[3:, :3] *= -1.0
    return M


def residual(b0, fils0, refs, N_ref, p, bdf2, R, dz):
    """Residual of the model swimmer after the Broyden update dz, and the
    body's velocity (V, Omega)."""
    b, fils = b0.copy(), [f.copy() for f in fils0]
    S.swimmer_update(b, fils, refs, dz, p)
    forces = [F.internal_forces(fl, p) for fl in fils]
    e = [F.self_mobility_residual(fl, f0.xm1, f0.xm2, f0.um1, bdf2, p) for fl, f0 in zip(fils, fils0)]
    vel = S.rotate_mobility(N_ref, b0.Q_init) @ S.body_force(b, fils, forces)
    if R > 0:
        for fl, Fi in zip(fils, forces):
            for n in range(p.nseg):
                vel += sphere_coupling(b.x, R, fl.x[n], p) @ Fi[n]
    W = F.dexpinv(b.u, vel[3:])
    if bdf2:
        ex = b.x - 4.0 / 3.0 * b0.xm1 + 1.0 / 3.0 * b0.xm2 - 2.0 / 3.0 * p.dt * vel[:3]
        eu = b.u - 1.0 / 3.0 * b0.um1 - 2.0 / 3.0 * p.dt * W
    else:
        ex = b.x - b0.xm1 - p.dt * vel[:3]
        eu = b.u - p.dt * W
    return np.concatenate(e + [ex, eu]), vel


def central_differences(b, fils, refs, N_ref, p, bdf2, R, h=1e-6):
    """d(residual)/dz and d(V, Omega)/dz at dz = 0."""
    n = 6 * p.nseg * len(fils) + 6
    Jr, Jv = [], []
    for c in range(n):
        dz = np.zeros(n)
        dz[c] = h
        rp, vp = residual(b, fils, refs, N_ref, p, bdf2, R, dz)
        rm, vm = residual(b, fils, refs, N_ref, p, bdf2, R, -dz)
        Jr.append((rp - rm) / (2 * h))
        Jv.append((vp - vm) / (2 * h))
    return np.array(Jr).T, np.array(Jv).T


def blocks_against_fd(nt, R=0.0, **kw):
    p, b, fils, refs, N_ref = swimmer(seed=nt + (11 if R else 0), **kw)
    bdf2 = nt >= 1
    forces = [F.internal_forces(f, p) for f in fils]
    S0, As, Bs = S.body_coupling(b, fils, refs, forces, nt, N_ref, p, R)
    Jr, Jv = central_differences(b, fils, refs, N_ref, p, bdf2, R)
    n6, nf = 6 * p.nseg, len(fils)
    A_fd = [Jr[n6 * i:n6 * (i + 1), n6 * nf:] for i in range(nf)]
    B_fd = [Jr[n6 * nf:, n6 * i:n6 * (i + 1)] for i in range(nf)]
    dt_fac = -2.0 * p.dt / 3.0 if bdf2 else -p.dt
    return dict(p=p, b=b, fils=fils, refs=refs, N_ref=N_ref, S0=S0, As=As, Bs=Bs, A_fd=A_fd, B_fd=B_fd,
                S0_fd=np.eye(6) + dt_fac * Jv[:, n6 * nf:], dt_fac=dt_fac)


def a_deviation(c, i):
    """The closed-form simplifications the reference applies to A_i."""
    p, f = c["p"], c["fils"][i]
    N = p.nseg
    d = c["b"].Q_init @ c["refs"][i]
    dev = np.zeros((6 * N, 6))
    dev[0:3, 3:6] = F.rcross(0.5 * p.dl * d / np.linalg.norm(d))
    tsf = -c["dt_fac"]
    dev[3 * N:3 * N + 3, 3:6] = tsf / (8 * PI * p.mu * p.a ** 3) * 0.5 * F.rcross(f.clamp)
    return dev


def s0_deviation(c):
    dev = np.zeros((6, 6))
    Nb = S.rotate_mobility(c["N_ref"], c["b"].Q_init)
    dev[3:, 3:] = c["dt_fac"] * Nb[3:, 3:] @ F.rcross(sum(f.clamp for f in c["fils"]))
    return dev


@pytest.mark.parametrize("nt", [0, 3])
def test_body_coupling_is_the_jacobian_of_the_swimmer_residual(nt):
    c = blocks_against_fd(nt)
    for i in range(len(c["fils"])):
        A, B = c["As"][i], c["Bs"][i]
        assert np.abs(A - c["A_fd"][i] - a_deviation(c, i)).max() <= 1e-7 * np.abs(A).max()
        assert np.abs(B - c["B_fd"][i]).max() <= 1e-7 * np.abs(B).max()
        # the simplifications are real (not zero) in this state
        assert np.abs(a_deviation(c, i)).max() > 1e-2 * np.abs(A).max()
    S0 = c["S0"]
    assert np.abs(S0 - c["S0_fd"] - s0_deviation(c)).max() <= 1e-7 * np.abs(S0 - np.eye(6)).max()
    assert np.abs(s0_deviation(c)).max() > 1e-2 * np.abs(S0 - np.eye(6)).max()


@pytest.mark.parametrize("nt", [0, 3])
def test_body_coupling_without_clamp_is_exact(nt):
    """With the clamp multipliers zero, only the segment-0 anchor offset is left."""
    c = blocks_against_fd(nt, clamp=False)
    for i in range(len(c["fils"])):
        dev = a_deviation(c, i)
        assert np.abs(dev[3:]).max() == 0.0
        assert np.abs(c["As"][i] - c["A_fd"][i] - dev).max() <= 1e-7 * np.abs(c["As"][i]).max()
    assert np.abs(c["S0"] - c["S0_fd"]).max() <= 1e-7 * np.abs(c["S0"] - np.eye(6)).max()


@pytest.mark.parametrize("nt", [0, 3])
def test_sphere_term_is_the_jacobian_of_the_sphere_coupling(nt):
    """The sphere term (swimmer.cpp:813-862) is exact where the reference's
    coupling has no position or direction dependence left out: multipliers
    lambda zero (their values do not enter the derivative), no end force, no
    elastic moments; tether and clamp act on the anchored segment 0."""
    c = blocks_against_fd(nt, R=6.0, lam=False, end_force=0.0, kb=0.0)
    for i in range(len(c["fils"])):
        B = c["Bs"][i]
        assert np.abs(B - c["B_fd"][i]).max() <= 1e-7 * np.abs(B).max()
    # the sphere adds to B only, and it is not small here
    no_sphere = S.body_coupling(c["b"], c["fils"], c["refs"], [F.internal_forces(f, c["p"]) for f in c["fils"]], nt,
                                c["N_ref"], c["p"], 0.0)
    assert np.array_equal(no_sphere[0], c["S0"])
    assert np.abs(no_sphere[2][0] - c["Bs"][0]).max() > 0.05 * np.abs(c["Bs"][0]).max()
    # the pair tensor's own translation-rotation coupling has the other sign
    p, b, f = c["p"], c["b"], c["fils"][0]
    rv = f.x[2] - b.x
    r = np.linalg.norm(rv)
    phys = np.column_stack([O.rpy_pair(6, 6, e, b.x, 6.0, f.x[2], p.a, p.mu) for e in np.eye(6)])
    ref_tr = F.rcross(rv / r) / (8 * PI * p.mu * r * r)
    assert np.allclose(phys[3:, :3], -ref_tr, rtol=1e-12, atol=1e-15)


def test_sphere_end_force_columns_approach_the_jacobian_far_from_the_body():
    """The end-force entries of the sphere term (swimmer.cpp:836-840) leave out
    how the distal segment's position enters the coupling: rotating it moves it
    by 0.5 DL, so the relative difference is about 0.5 DL / r at distance r
    from the sphere's centre and vanishes far from it."""
    errs, dist = [], []
    for scale in (1.0, 2.0, 4.0, 8.0):
        p, b, fils, refs, N_ref = swimmer(seed=21, lam=False, kb=0.0, end_force=30.0, nfil=1)
        refs = [scale * r for r in refs]
        fils = []
        d = b.Q_init @ refs[0]
        f = F.Filament.at_rest(b.x + d, np.tile(along(d), (p.nseg, 1)), p.nseg, p.dl)
        f.tether = np.array([1.0, -2.0, 0.5])
        fils.append(f)
        forces = [F.internal_forces(fl, p) for fl in fils]
        _, _, Bs = S.body_coupling(b, fils, refs, forces, 3, N_ref, p, 6.0)
        Jr, _ = central_differences(b, fils, refs, N_ref, p, True, 6.0)
        n6 = 6 * p.nseg
        cols = slice(n6 - 3, n6)
        # the sphere's share of the end-force columns: with and without the sphere
        _, _, B0 = S.body_coupling(b, fils, refs, forces, 3, N_ref, p, 0.0)
        Jr0, _ = central_differences(b, fils, refs, N_ref, p, True, 0.0)
        sph, sph_fd = Bs[0][:, cols] - B0[0][:, cols], Jr[n6:, cols] - Jr0[n6:, cols]
        errs.append(np.abs(sph - sph_fd).max() / np.abs(sph_fd).max())
        dist.append(np.linalg.norm(f.x[-1] - b.x))
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:]))
    ratio = np.array(errs) * np.array(dist) / (0.5 * p.dl)
    assert np.all(np.abs(ratio - 1.0) < 0.25) and abs(ratio[-1] - 1.0) < 0.05


@pytest.mark.parametrize("nt", [0, 3])
def test_body_lie_term_uses_the_reference_omega_substitution(nt):
    """S0 of a body alone: the exact Jacobian of u - |dt_fac| dexpinv(u, Omega)
    at Omega = -u / |dt_fac| (swimmer.cpp:690-697), not at +u / |dt_fac|."""
    p = F.Params(nseg=5, dt=DT)
    dt_fac = -2.0 * DT / 3.0 if nt >= 1 else -DT
    b = S.Body.at_rest(np.zeros(3))
    b.u = 0.01 * np.random.default_rng(nt).standard_normal(3)
    b.q = F.qmul(F.lie_exp(b.u), b.qm1)
    b.Q_init = F.rot_mat(b.q)
    S0, _, _ = S.body_coupling(b, [], [], [], nt, np.eye(6), p, 0.0)
    assert np.array_equal(S0[:3], np.eye(6)[:3]) and np.array_equal(S0[3:, :3], np.zeros((3, 3)))

    def jac(om):
        h, J = 1e-6, np.zeros((3, 3))
        for k in range(3):
            e = h * np.eye(3)[k]
            J[:, k] = ((b.u + e) + dt_fac * F.dexpinv(b.u + e, om) - (b.u - e) - dt_fac * F.dexpinv(b.u - e, om)) / (2 * h)
        return J

    scale = np.abs(S0[3:, 3:] - np.eye(3)).max()
    assert scale > 1e-3
    assert np.abs(S0[3:, 3:] - jac(-b.u / abs(dt_fac))).max() <= 1e-6 * scale
    assert np.abs(S0[3:, 3:] - jac(b.u / abs(dt_fac))).max() > 0.5 * scale


def test_planted_sign_flips_fail_the_checks():
    """The checks above catch a sign slip in any block (planted in copies of
    the oracle's output)."""
    c = blocks_against_fd(3)
    i = 0
    A, B, S0 = c["As"][i], c["Bs"][i], c["S0"]
    ok = lambda A_, B_, S0_: (np.abs(A_ - c["A_fd"][i] - a_deviation(c, i)).max() <= 1e-7 * np.abs(A).max() and
                              np.abs(B_ - c["B_fd"][i]).max() <= 1e-7 * np.abs(B).max() and
                              np.abs(S0_ - c["S0_fd"] - s0_deviation(c)).max() <= 1e-7 * np.abs(S0 - np.eye(6)).max())
    assert ok(A, B, S0)
    S0_flip = S0.copy()
    S0_flip[3:, 3:] = 2 * np.eye(3) - S0[3:, 3:]
    assert not ok(A, B, S0_flip)
    B_flip = B.copy()
    B_flip[3:] *= -1
    assert not ok(A, B_flip, S0)
    A_flip = A.copy()
    A_flip[:, 3:] *= -1
    assert not ok(A_flip, B, S0)
    A_idx = A.copy()
    A_idx[3 * c["p"].nseg:3 * c["p"].nseg + 3, 3:] = A[3 * c["p"].nseg + 3:3 * c["p"].nseg + 6, 3:]
    assert not ok(A_idx, B, S0)


def test_blob_force_guess_is_exact_on_polynomial_histories():
    """rigid_body::initial_guess in the Broyden-for-everything form
    (rigid_body.cpp:447-478): the converged blob forces of the last steps are
    extrapolated linearly at step 2 and quadratically from step 3 on, so a
    linear history is continued exactly from step 2 and a quadratic one from
    step 3; step 1 only stores the history."""
    rng = np.random.default_rng(3)
    c0, c1, c2 = rng.standard_normal((3, 12))
    for poly, exact_from in ((lambda n: c0 + c1 * n, 2), (lambda n: c0 + c1 * n + c2 * n * n, 3)):
        fb, fbm1, fbm2 = poly(0), np.zeros(12), np.zeros(12)
        for nt in range(1, 7):
            fb, fbm1, fbm2 = S.blob_force_guess(fb, fbm1, fbm2, nt)
            if nt == 1:
                assert np.array_equal(fb, poly(0)) and np.array_equal(fbm1, poly(0))
            if nt >= exact_from:
                assert np.allclose(fb, poly(nt), rtol=1e-12, atol=1e-12)
            fb = poly(nt)          # the step converges to the history's next value
    # a quadratic history is not continued exactly by the linear rule of step 2
    q = lambda n: c0 + c1 * n + c2 * n * n
    g, _, _ = S.blob_force_guess(q(1), q(0), np.zeros(12), 2)
    assert not np.allclose(g, q(2))
