"""Pins the oracle's barrier forces (cuda_functions.cu:1240-1436) on the CPU:
the force law at points where it has simple values (saturation at contact, a
sixteenth at the half-way gap, zero at the 10% gap and beyond), its direction,
the neighbour exclusion along a filament, Newton's third law on a random
cluster, rotation covariance, the wall term's values, and the write rules."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O

S = 7.5


def pair(d, seg_i, seg_j, a_seg=1.0, a_blob=1.3):
    """Two particles on the x axis, i at 0 and j at d; returns (f_i, f_j)."""
    xs, xb = [], []
    for seg, x in ((seg_i, 0.0), (seg_j, d)):
        (xs if seg else xb).append([x, 0.0, 0.0])
    fs, rep = O.barrier_forces(np.array(xs).reshape(-1), np.array(xb).reshape(-1),
                               np.zeros(6 * len(xs)), nseg=1, strength=S, a_seg=a_seg, a_blob=a_blob)
    out = [fs.reshape(-1, 6)[:, :3][k] for k in range(len(xs))] + [rep.reshape(-1, 3)[k] for k in range(len(xb))]
    if not seg_i and seg_j:
        out = out[::-1]
    return out[0], out[1]


@pytest.mark.parametrize("seg_i,seg_j", [(True, True), (True, False), (False, True), (False, False)])
def test_force_law_values(seg_i, seg_j):
    a_sum = (1.0 if seg_i else 1.3) + (1.0 if seg_j else 1.3)
    cases = [(0.6 * a_sum, S), (a_sum, S), (1.05 * a_sum, S / 16.0), (1.075 * a_sum, S / 256.0),
             (1.1 * a_sum, 0.0), (1.3 * a_sum, 0.0)]
    for d, mag in cases:
        fi, fj = pair(d, seg_i, seg_j)
        # i sits at -x of j: pushed towards -x, j towards +x, equal and opposite
        np.testing.assert_allclose(fi, [-mag, 0.0, 0.0], rtol=1e-13, atol=1e-13 * S)
        np.testing.assert_allclose(fj, -fi, rtol=0, atol=0)
    # continuous at the cut-off
    fi, _ = pair(1.1 * a_sum * (1 - 1e-9), seg_i, seg_j)
    assert 0 < -fi[0] < 1e-30 * S + 1e-20


def test_filament_neighbours_are_excluded():
    # one filament of three segments 1.0 apart: 0-1 and 1-2 are neighbours,
    # 0-2 (2.0 apart, inside 2.2) repel with full strength
    xs = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]]).reshape(-1)
    fs, _ = O.barrier_forces(xs, np.zeros(0), np.zeros(18), nseg=3, strength=S)
    f = fs.reshape(3, 6)[:, :3]
    np.testing.assert_allclose(f, [[-S, 0, 0], [0, 0, 0], [S, 0, 0]], atol=1e-13)
    # the same chain as three one-segment filaments: neighbours now repel too
    fs, _ = O.barrier_forces(xs, np.zeros(0), np.zeros(18), nseg=1, strength=S)
    f = fs.reshape(3, 6)[:, :3]
    np.testing.assert_allclose(f, [[-2 * S, 0, 0], [0, 0, 0], [2 * S, 0, 0]], atol=1e-13)


def cluster(seed=0, nfil=12, nseg=6, nb=80, box=9.0):
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, box, (nfil, 3))
    dirs = rng.standard_normal((nfil, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    xs = (base[:, None, :] + 1.9 * np.arange(nseg)[None, :, None] * dirs[:, None, :]).reshape(-1)
    xb = rng.uniform(0, box, 3 * nb)
    return xs, xb, nseg


def test_newton_third_law_and_write_rules():
    xs, xb, nseg = cluster()
    ns = xs.size // 3
    rng = np.random.default_rng(1)
    f0 = rng.standard_normal(6 * ns)
    fs, rep = O.barrier_forces(xs, xb, f0, nseg=nseg, strength=S, a_blob=1.2)
    add = (fs - f0).reshape(ns, 6)
    assert np.all(add[:, 3:] == 0.0)                      # torques untouched
    total = add[:, :3].sum(0) + rep.reshape(-1, 3).sum(0)
    assert np.abs(add).max() > S                          # the cluster does interact
    assert np.linalg.norm(total) <= 1e-12 * S * ns


def test_rotation_covariance():
    xs, xb, nseg = cluster(seed=2)
    R = Rotation.from_rotvec([0.3, -1.1, 0.7]).as_matrix()
    fs, rep = O.barrier_forces(xs, xb, np.zeros(2 * xs.size), nseg=nseg, strength=S)
    fs_r, rep_r = O.barrier_forces((xs.reshape(-1, 3) @ R.T).reshape(-1), (xb.reshape(-1, 3) @ R.T).reshape(-1),
                                   np.zeros(2 * xs.size), nseg=nseg, strength=S)
    np.testing.assert_allclose(fs_r.reshape(-1, 6)[:, :3], fs.reshape(-1, 6)[:, :3] @ R.T, atol=1e-11 * S)
    np.testing.assert_allclose(rep_r.reshape(-1, 3), rep.reshape(-1, 3) @ R.T, atol=1e-11 * S)


def test_wall_term_values():
    # DL = 2.2 a: the wall force acts below z = 1.1 a, saturating at z <= a
    for z, fz in [(0.5, S), (1.0, S), (1.05, S / 16.0), (1.1, 0.0), (3.0, 0.0)]:
        fs, _ = O.barrier_forces(np.array([0.0, 0.0, z]), np.zeros(0), np.zeros(6), nseg=1, strength=S,
                                 wall=True, dl=2.2)
        np.testing.assert_allclose(fs[:3], [0.0, 0.0, fz], atol=1e-13 * S)
    fs, _ = O.barrier_forces(np.array([0.0, 0.0, 1.05]), np.zeros(0), np.zeros(6), nseg=1, strength=S)
    assert np.all(fs == 0.0)                              # no wall, no force
