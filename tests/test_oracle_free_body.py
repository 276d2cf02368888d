"""Pins the oracle's free-body columns (oracle/solver_oracle.py k_body_matrix,
mobility_solver.cpp:1850-1908) on the CPU: K_b [V; Omega] is the rigid-body
velocity V + Omega x (x - x_body) of each particle of its swimmer, K_b^T is
total force and torque about the body centre, and the free-body saddle matrix
is symmetric."""
import numpy as np

from oracle import solver_oracle as SO


def test_body_columns_are_rigid_motion_and_their_transpose_force_and_torque():
    rng = np.random.default_rng(0)
    ns, nb, nsw = 12, 18, 3
    xs, xb = rng.standard_normal(3 * ns) * 5, rng.standard_normal(3 * nb) * 5
    xbod = rng.standard_normal(3 * nsw)
    K = SO.k_body_matrix(xs, xb, xbod, nsw)
    VW = rng.standard_normal(6 * nsw)
    got = (K @ VW).reshape(-1, 3)
    x = np.concatenate([xs, xb]).reshape(-1, 3)
    owner = np.concatenate([np.arange(ns) // (ns // nsw), np.arange(nb) // (nb // nsw)])
    VWr = VW.reshape(nsw, 6)
    want = VWr[owner, :3] + np.cross(VWr[owner, 3:], x - xbod.reshape(nsw, 3)[owner])
    np.testing.assert_allclose(got, want, atol=1e-13)
    lam = rng.standard_normal(x.shape)
    ft = (K.T @ lam.reshape(-1)).reshape(nsw, 6)
    for s in range(nsw):
        sel = owner == s
        np.testing.assert_allclose(ft[s, :3], lam[sel].sum(0), atol=1e-12)
        np.testing.assert_allclose(ft[s, 3:], np.cross(x[sel] - xbod.reshape(nsw, 3)[s], lam[sel]).sum(0), atol=1e-12)


def test_free_body_saddle_matrix_is_symmetric():
    rng = np.random.default_rng(1)
    nfil, nseg, nb = 2, 4, 10
    xs = rng.standard_normal(3 * nfil * nseg) * 6
    xb = rng.standard_normal(3 * nb) * 6
    kp, kt = rng.standard_normal(3 * nfil * nseg), rng.standard_normal(3 * nfil * nseg)
    A = SO.saddle_matrix(xs, xb, kp, kt, nfil, nseg, x_bodies=np.zeros(3))
    assert A.shape == (3 * (nfil * nseg + nb) + 6 + 2 * nfil,) * 2
    np.testing.assert_allclose(A, A.T, atol=1e-14)
