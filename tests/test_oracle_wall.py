"""Pins of the oracle's half-space pair tensor (no-slip wall at z = 0,
cuda_functions.cu:12-203, Swan & Brady 2007; radius 1 and viscosity 1 as the
reference assumes):

* symmetry of the grand mobility, M_ij = M_ji^T for 6-DOF particles -- this
  fixes the signs of the rotation-translation wall terms, which the reference
  flips relative to the paper's printed convention;
* positive definiteness for particles above the wall;
* the unbounded far field recovered far from the wall;
* the self-mobility against the classical wall expansions in a/h to the
  order the RPY-level theory keeps: Faxen's parallel 1 - 9/16 (a/h) +
  1/8 (a/h)^3, and Lorentz/Brenner's normal 1 - 9/8 (a/h) + 1/2 (a/h)^3 (both
  times 1/(6 pi)), errors O((a/h)^4);
* invariance under translations parallel to the wall and rotations about the
  wall normal."""
import numpy as np

from oracle import oracle as O

rng = np.random.default_rng(11)


def particles(n, zmin=1.5, span=6.0):
    x = rng.uniform(-span, span, size=(n, 3))
    x[:, 2] = rng.uniform(zmin, zmin + span, size=n)
    # keep distinct particles at least 2 radii apart (the far-field form has no overlap branch)
    for i in range(n):
        for j in range(i):
            if np.linalg.norm(x[i] - x[j]) < 2.05:
                x[i, 2] += 3.0
    return x


def test_half_space_grand_mobility_is_symmetric_positive_definite():
    x = particles(9)
    with O.domain(O.HALF_SPACE):
        M = O.dense_mobility(x.reshape(-1), np.zeros(0), seg_dof=6)
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-13 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    # mixed classes (segments with 6 DOF, blobs with 3): still symmetric
    xb = particles(5) + np.array([20.0, 0.0, 0.0])
    with O.domain(O.HALF_SPACE):
        M2 = O.dense_mobility(x.reshape(-1), xb.reshape(-1), seg_dof=6)
    np.testing.assert_allclose(M2, M2.T, rtol=0, atol=1e-13 * np.abs(M2).max())
    assert np.linalg.eigvalsh(0.5 * (M2 + M2.T)).min() > 0


def test_far_from_wall_recovers_unbounded_rpy():
    xi = np.array([0.3, -0.4, 1e5])
    xj = np.array([2.9, 1.7, 1e5 + 1.1])
    f = rng.normal(size=6)
    for vd, fd in ((3, 3), (6, 3), (3, 6), (6, 6)):
        w = O.rpy_pair_wall(vd, fd, f[:fd], xi, xj)
        u = O.rpy_pair(vd, fd, f[:fd], xi, 1.0, xj, 1.0)
        assert np.abs(w - u).max() <= 1e-4 * np.abs(u).max()   # image terms O(r/z) here


def test_self_mobility_matches_wall_expansions():
    c6 = 1.0 / (6.0 * np.pi)
    for h in (4.0, 8.0, 16.0, 32.0):
        x = np.array([0.0, 0.0, h])
        vx = O.rpy_pair_wall(3, 3, [1.0, 0.0, 0.0], x, x, same=True)
        vz = O.rpy_pair_wall(3, 3, [0.0, 0.0, 1.0], x, x, same=True)
        par = c6 * (1 - 9 / 16 / h + 1 / 8 / h**3)
        perp = c6 * (1 - 9 / 8 / h + 1 / 2 / h**3)
        assert abs(vx[0] - par) <= c6 * 0.2 / h**4
        assert abs(vz[2] - perp) <= c6 * 0.2 / h**4
        assert abs(vx[1]) == 0 and abs(vx[2]) == 0
    # far from the wall: the free-space self mobilities 1/(6 pi) and 1/(8 pi)
    x = np.array([0.0, 0.0, 1e6])
    w = O.rpy_pair_wall(6, 6, [1, 2, 3, 4, 5, 6], x, x, same=True)
    np.testing.assert_allclose(w, np.concatenate([c6 * np.array([1, 2, 3.0]), np.array([4, 5, 6.0]) / (8 * np.pi)]),
                               rtol=1e-5)


def test_invariant_under_wall_parallel_translation_and_normal_rotation():
    xi, xj = np.array([0.5, 1.0, 2.5]), np.array([-2.0, 0.7, 3.1])
    f = rng.normal(size=6)
    base = O.rpy_pair_wall(6, 6, f, xi, xj)
    shift = np.array([13.0, -7.0, 0.0])
    np.testing.assert_allclose(O.rpy_pair_wall(6, 6, f, xi + shift, xj + shift), base, rtol=1e-12, atol=1e-15)
    t = 0.7
    R = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1.0]])
    fr = np.concatenate([R @ f[:3], R @ f[3:]])
    rot = O.rpy_pair_wall(6, 6, fr, R @ xi, R @ xj)
    np.testing.assert_allclose(rot, np.concatenate([R @ base[:3], R @ base[3:]]), rtol=1e-11, atol=1e-14)


def test_domain_switch_restores_unbounded():
    x = particles(4)
    before = O.dense_mobility(x.reshape(-1), np.zeros(0), seg_dof=3)
    with O.domain(O.HALF_SPACE):
        wall = O.dense_mobility(x.reshape(-1), np.zeros(0), seg_dof=3)
    after = O.dense_mobility(x.reshape(-1), np.zeros(0), seg_dof=3)
    np.testing.assert_array_equal(before, after)
    assert np.abs(wall - before).max() > 1e-3 * np.abs(before).max()
