"""Pins the Broyden oracle (oracle/broyden_oracle.py) without the reference.

* secant condition H_{k+1} (e_{k+1} - e_k) = -(x_{k+1} - x_k) after every update;
* the low-rank C/D bookkeeping equals Broyden's second method written as a dense
  rank-one recurrence on H;
* the no-update rules of end_of_iter (error small enough, memory full);
* convergence: exact Jacobian and confidence 1 solve a linear problem in one
  step, confidence 0.1 converges within 2n steps on a linear problem (finite
  termination of Broyden's methods, Gay 1979), and a mildly nonlinear problem
  converges to its root;
* dexp equals the right-trivialised derivative of the matrix exponential,
  dexpinv inverts it, and the small-angle branches join the large-angle ones;
* compute_errors vanishes on exact backward-Euler / BDF2 states and follows the
  reference's storage order.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import broyden_oracle as B


def _blocks(rng, sizes, diag=4.0):
    return [rng.normal(size=(s, s)) * 0.3 + diag * np.eye(s) for s in sizes]


def _block_diag(blocks):
    n = sum(b.shape[0] for b in blocks)
    M = np.zeros((n, n))
    r = 0
    for b in blocks:
        s = b.shape[0]
        M[r:r + s, r:r + s] = b
        r += s
    return M


def _nonlinear(A, b, eps):
    return lambda x: A @ x + eps * np.sin(x) - b


def test_secant_condition_and_dense_recurrence():
    rng = np.random.default_rng(3)
    sizes = [6, 6, 4]
    n = sum(sizes)
    J0 = _blocks(rng, sizes)
    A = _block_diag(J0) + 0.2 * rng.normal(size=(n, n))
    F = _nonlinear(A, rng.normal(size=n), 0.3)
    Jinv = [np.linalg.inv(b) for b in J0]
    br = B.BroydenRef(n, 50, 0.1, Jinv)
    x = rng.normal(size=n)
    e = F(x)
    br.start(e)
    H = -0.1 * np.linalg.inv(_block_diag(J0))
    for _ in range(8):
        assert np.allclose(br.dense_inverse(), H, rtol=1e-12, atol=1e-12)
        dx = br.find_update()
        assert np.allclose(dx, H @ e, rtol=1e-12, atol=1e-13)
        x = x + dx
        e_new = F(x)
        br.end_of_iter(e_new, True)
        de = e_new - e
        H = H - np.outer(H @ e_new, de) / (de @ de)
        Hk = br.dense_inverse()
        np.testing.assert_allclose(Hk @ de, -dx, rtol=1e-10, atol=1e-12)
        e = e_new


def test_no_update_when_small_or_full():
    rng = np.random.default_rng(4)
    n = 5
    br = B.BroydenRef(n, 3, 0.1, [np.eye(n)])
    br.start(rng.normal(size=n))
    e0 = br.error.copy()
    br.end_of_iter(rng.normal(size=n), False)
    assert br.iter == 1 and not br.C.any() and np.array_equal(br.error, e0)
    br.end_of_iter(rng.normal(size=n), True)     # iter 1 < max_iter - 1: stored
    assert br.iter == 2 and br.C[:, 1].any() and not br.C[:, 2].any()
    e2 = br.error.copy()
    br.end_of_iter(rng.normal(size=n), True)     # iter 2 = max_iter - 1: nothing stored
    assert br.iter == 3 and not br.C[:, 2].any() and np.array_equal(br.error, e2)


def test_exact_jacobian_linear_problem_one_step():
    rng = np.random.default_rng(5)
    sizes = [3, 5]
    J = _blocks(rng, sizes)
    A = _block_diag(J)
    b = rng.normal(size=8)
    x, iters, norms = B.solve(lambda x: A @ x - b, np.zeros(8), [np.linalg.inv(m) for m in J],
                              confidence=1.0, tol=1e-12)
    assert iters == 1 and norms[-1] < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_linear_problem_finite_termination(seed):
    rng = np.random.default_rng(seed)
    sizes = [4, 4]
    n = 8
    J0 = _blocks(rng, sizes)
    A = _block_diag(J0) + 0.3 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    x, iters, norms = B.solve(lambda x: A @ x - b, np.zeros(n), [np.linalg.inv(m) for m in J0],
                              confidence=0.1, tol=1e-10)
    assert iters <= 2 * n + 2, (iters, norms)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_nonlinear_problem_converges():
    rng = np.random.default_rng(7)
    sizes = [6, 6, 6]
    n = 18
    J0 = _blocks(rng, sizes)
    A = _block_diag(J0) + 0.1 * rng.normal(size=(n, n))
    F = _nonlinear(A, rng.normal(size=n), 0.2)
    x, iters, norms = B.solve(F, np.zeros(n), [np.linalg.inv(m) for m in J0], confidence=0.1, tol=1e-10)
    assert norms[-1] <= 1e-10 and iters < 100
    assert np.linalg.norm(F(x)) <= 1e-10


def _hat(u):
    return np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])


@pytest.mark.parametrize("scale", [1e-3, 0.3, 1.0, 2.5])
def test_dexp_is_derivative_of_exponential(scale):
    rng = np.random.default_rng(11)
    u = rng.normal(size=3) * scale
    v = rng.normal(size=3)
    h = 1e-6
    D = (expm(_hat(u + h * v)) - expm(_hat(u - h * v))) / (2 * h)
    W = D @ expm(_hat(u)).T
    np.testing.assert_allclose([W[2, 1], W[0, 2], W[1, 0]], B.dexp(u, v), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(B.dexp(u, B.dexpinv(u, v)), v, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(B.dexpinv(u, B.dexp(u, v)), v, rtol=1e-12, atol=1e-13)


def test_small_angle_branches_are_continuous():
    """Just above the threshold the trigonometric branch equals the series
    branch written out at the same u."""
    rng = np.random.default_rng(12)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    v = rng.normal(size=3)
    u = d * (B.SMALL_ANGLE * 1.001)
    c = np.cross(u, v)
    w = np.cross(u, c)
    np.testing.assert_allclose(B.dexpinv(u, v), v - 0.5 * c + w / 12.0, rtol=1e-13)
    t2 = u @ u
    beta = (1.0 / 3.0 + t2 / 72.0) / (2.0 + t2 / 6.0 + t2 * t2 / 72.0)
    alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - t2 / 12.0)
    # alpha of the series branch is 1/2 + O(theta^2): the reference's series is
    # built to invert dexpinv's series, so compare through the inverse property
    np.testing.assert_allclose(B.dexp(u, B.dexpinv(u, v)), v, rtol=1e-13)
    assert abs(alpha - 0.5) < 1e-9 and abs(beta - 1.0 / 6.0) < 1e-9
    np.testing.assert_allclose(B.dexpinv(np.zeros(3), v), v, rtol=0, atol=0)


def _state(rng, ns, nb, dt, bdf2, exact):
    """Random state; `exact` makes positions and Lie elements satisfy the step equations."""
    c1, c2, c3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
    st = {}
    for pre, m in (("seg_", ns), ("body_", nb)):
        st[pre + "xm1"] = rng.normal(size=3 * m)
        st[pre + "xm2"] = rng.normal(size=3 * m)
        st[pre + "um1"] = rng.normal(size=3 * m) * 0.1
        st[pre + "u"] = rng.normal(size=3 * m) * 0.1
    st["v_segs"] = rng.normal(size=6 * ns)
    st["v_bodies"] = rng.normal(size=6 * nb)
    for pre, m, vk in (("seg_", ns, "v_segs"), ("body_", nb, "v_bodies")):
        v = st[vk].reshape(m, 6)
        if exact:
            x = np.empty((m, 3))
            for k in range(m):
                xm1, xm2 = st[pre + "xm1"][3 * k:3 * k + 3], st[pre + "xm2"][3 * k:3 * k + 3]
                x[k] = (-(c1 * xm1 + c2 * xm2 + c3 * dt * v[k, :3])) if bdf2 else xm1 + dt * v[k, :3]
                # omega such that the Lie residual vanishes: dexpinv(u, omega) = W, omega = dexp(u, W)
                u = st[pre + "u"][3 * k:3 * k + 3]
                um1 = st[pre + "um1"][3 * k:3 * k + 3]
                W = (u - c2 * um1) / (-c3 * dt) if bdf2 else u / dt
                v[k, 3:] = B.dexp(u, W)
            st[pre + "x"] = x.reshape(-1)
        else:
            st[pre + "x"] = rng.normal(size=3 * m)
        st[vk] = v.reshape(-1)
    return st


@pytest.mark.parametrize("bdf2", [False, True])
def test_compute_errors_vanish_on_exact_states(bdf2):
    rng = np.random.default_rng(20)
    nswim, nfil, nseg, dt = 2, 3, 4, 0.01
    st = _state(rng, nswim * nfil * nseg, nswim, dt, bdf2, exact=True)
    err, big = B.compute_errors(st, nswim, nfil, nseg, bdf2, dt, dl=2.2, tol=1e-4)
    assert np.abs(err).max() < 1e-13 and not big


def test_compute_errors_storage_order():
    rng = np.random.default_rng(21)
    nswim, nfil, nseg, dt = 2, 2, 3, 0.05
    st = _state(rng, nswim * nfil * nseg, nswim, dt, True, exact=True)
    # perturb one segment position and one body's previous Lie element (both enter
    # their residual linearly); see where they land
    s, i, j = 1, 1, 2
    g = j + nseg * (i + s * nfil)
    st["seg_x"][3 * g + 1] += 1.0
    st["body_um1"][3 * 0 + 2] += 0.6
    err, big = B.compute_errors(st, nswim, nfil, nseg, True, dt, dl=1.0, tol=1e-4)
    hot = np.flatnonzero(np.abs(err) > 1e-8)
    pid = s * 6 * (nfil * nseg + 1) + 6 * i * nseg + 3 * j + 1
    body_lie = 0 * 6 * (nfil * nseg + 1) + 6 * nfil * nseg + 3 + 2
    assert set(hot) == {pid, body_lie}, hot
    assert big
    np.testing.assert_allclose(err[pid], 1.0, rtol=1e-12)
    np.testing.assert_allclose(err[body_lie], -0.2, rtol=1e-9)
