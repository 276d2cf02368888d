"""The generalised-force tables of the prescribed beat (beat.py), built as the
reference's WRITE_GENERALISED_FORCES run builds them (cilia_sim_main.cu:370-380:
q = K^T f for the segment forces f that drive one filament at its natural
frequency).  Pinned against the oracle's RPY mobility: with the tables as
generalised forces, the isolated filament's own 2 x 2 generalised resistance
K^T M^-1 K returns phi_dot = 2 pi and theta_dot = 0 at every phase -- the
defining property that keeps the dynamic shape angle bounded."""
import numpy as np

from oracle import oracle as O
from timfilsimmod_amd import beat as B

NSEG, DL = 20, 2.2


def _filament(phase, length):
    s = B.uniform_s(NSEG)
    pos, der = B.shape(s, phase)
    x = np.zeros((NSEG, 3))
    x[:, :2] = length * pos
    kphi = np.zeros((NSEG, 3))
    kphi[1:, :2] = length * der[1:]
    ktheta = np.zeros((NSEG, 3))
    ktheta[1:, 0] = -x[1:, 1]
    ktheta[1:, 1] = x[1:, 0]
    return x, kphi.reshape(-1), ktheta.reshape(-1)


def test_tables_drive_isolated_filament_at_natural_frequency():
    length = DL * (NSEG - 1)
    samples = 60
    qp, qa = B.generalised_force_tables(NSEG, length, samples=samples)
    assert qp.shape == qa.shape == (samples,)
    for k in range(0, samples, 7):
        x, kp, kt = _filament(2 * np.pi * k / samples, length)
        M = O.dense_mobility(x.reshape(-1), np.zeros(0), seg_dof=3)
        K = np.stack([kp, kt], axis=1)
        R = K.T @ np.linalg.solve(M, K)
        rates = np.linalg.solve(R, np.array([qp[k], qa[k]]))
        np.testing.assert_allclose(rates, [2 * np.pi, 0.0], rtol=0, atol=1e-8 * 2 * np.pi)


def test_tables_are_periodic_and_cached_copies():
    length = DL * (NSEG - 1)
    qp, qa = B.generalised_force_tables(NSEG, length, samples=40)
    qp2, _ = B.generalised_force_tables(NSEG, length, samples=40)
    qp2[:] = 0.0                     # the caller's copy, not the cache
    qp3, _ = B.generalised_force_tables(NSEG, length, samples=40)
    np.testing.assert_array_equal(qp, qp3)
    # sample k is phase 2 pi k / samples: a finer table holds the coarse one at its even entries
    qp80, qa80 = B.generalised_force_tables(NSEG, length, samples=80)
    np.testing.assert_allclose(qp80[::2], qp, rtol=1e-12)
    np.testing.assert_allclose(qa80[::2], qa, rtol=1e-12, atol=1e-12 * np.abs(qa).max())
    assert np.all(qp > 0)            # driving the beat forward costs positive work
