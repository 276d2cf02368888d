"""Pins the FP64 oracle's pair tensor (oracle/rpy_oracle.c, restating
src/cilia/cuda_functions.cu:207-354) against facts it does not share code with:

* closed-form self mobilities 1/(6 pi mu a) and 1/(8 pi mu a^3);
* far field from its definition: tt = (1 + (ai^2+aj^2)/6 Lap) G, tr = rt =
  (1/2) curl G, rr = (1/4) curl curl G, with G the Oseen tensor and every
  derivative taken by 4th-order finite differences;
* equal-radius overlap blocks against the published Rotne-Prager (1969) /
  Wajnryb et al. (2013) closed forms;
* continuity across r = ai+aj and r = |ai-aj| for unequal radii (Zuk et al. 2014);
* symmetry and positive definiteness of the assembled grand mobility;
* rotation covariance and translation invariance.
"""
import numpy as np
import pytest

from oracle import oracle as O

MU = 1.3


def pair_matrix(vd, fd, xi, ai, xj, aj, mu=MU):
    m = np.zeros((vd, fd))
    for c in range(fd):
        f = np.zeros(fd)
        f[c] = 1.0
        m[:, c] = O.rpy_pair(vd, fd, f, xi, ai, xj, aj, mu)
    return m


def oseen(r):
    r = np.asarray(r, dtype=float)
    n = np.linalg.norm(r)
    return (np.eye(3) / n + np.outer(r, r) / n**3) / (8 * np.pi * MU)


H = 0.02
STENCIL = [(-2, 1.0 / 12), (-1, -8.0 / 12), (1, 8.0 / 12), (2, -1.0 / 12)]


def grad_fd(field, x):
    """Jacobian d field_a / d x_b by 4th-order central differences."""
    x = np.asarray(x, dtype=float)
    f0 = np.asarray(field(x))
    jac = np.zeros(f0.shape + (3,))
    for b in range(3):
        e = np.zeros(3)
        e[b] = H
        jac[..., b] = sum(w * np.asarray(field(x + k * e)) for k, w in STENCIL) / H
    return jac


def curl_fd(field, x):
    j = grad_fd(field, x)
    return np.array([j[2, 1] - j[1, 2], j[0, 2] - j[2, 0], j[1, 0] - j[0, 1]])


def laplacian_fd(field, x):
    x = np.asarray(x, dtype=float)
    acc = -2.5 * 3 * np.asarray(field(x))
    for b in range(3):
        e = np.zeros(3)
        e[b] = H
        acc = acc + sum(w * np.asarray(field(x + k * e)) for k, w in
                        [(-2, -1.0 / 12), (-1, 16.0 / 12), (1, 16.0 / 12), (2, -1.0 / 12)])
    return acc / H**2


def test_self_mobility_closed_forms():
    a = 1.7
    m = pair_matrix(6, 6, [0.3, -2, 5], a, [0.3, -2, 5], a)
    np.testing.assert_allclose(m[:3, :3], np.eye(3) / (6 * np.pi * MU * a), rtol=1e-15)
    np.testing.assert_allclose(m[3:, 3:], np.eye(3) / (8 * np.pi * MU * a**3), rtol=1e-15)
    assert np.all(m[:3, 3:] == 0) and np.all(m[3:, :3] == 0)


def test_inner_branch_uses_larger_radius():
    # r = 0 with unequal radii: rigid-body limit of the larger sphere.
    m = pair_matrix(6, 6, [0, 0, 0], 0.5, [0, 0, 0], 2.0)
    np.testing.assert_allclose(m[:3, :3], np.eye(3) / (6 * np.pi * MU * 2.0), rtol=1e-15)
    np.testing.assert_allclose(m[3:, 3:], np.eye(3) / (8 * np.pi * MU * 8.0), rtol=1e-15)


@pytest.mark.parametrize("ai,aj", [(1.0, 1.0), (0.7, 1.9)])
@pytest.mark.parametrize("d", [[4.0, 1.0, -2.5], [-3.0, 5.0, 0.5]])
def test_far_field_tt_is_faxen_laplacian_of_oseen(ai, aj, d):
    xi = np.array([0.4, -0.2, 1.1])
    xj = xi - np.array(d)
    m = pair_matrix(3, 3, xi, ai, xj, aj)
    lap = laplacian_fd(lambda x: oseen(x - xj), xi)
    expect = oseen(xi - xj) + (ai**2 + aj**2) / 6.0 * lap
    np.testing.assert_allclose(m, expect, rtol=0, atol=1e-8 * np.abs(expect).max())


@pytest.mark.parametrize("d", [[4.0, 1.0, -2.5], [-3.0, 5.0, 0.5]])
def test_far_field_rotational_blocks_are_curls_of_oseen(d):
    ai, aj = 1.0, 1.4
    xi = np.array([0.4, -0.2, 1.1])
    xj = xi - np.array(d)
    m = pair_matrix(6, 6, xi, ai, xj, aj)
    scale = np.abs(m).max()
    for c in range(3):
        e = np.zeros(3)
        e[c] = 1.0
        # rotation from a force on j: (1/2) curl of the Stokeslet field at xi
        rt = 0.5 * curl_fd(lambda x: oseen(x - xj) @ e, xi)
        np.testing.assert_allclose(m[3:, c], rt, atol=1e-8 * scale)
        # translation from a torque on j: the rotlet (1/2) curl(G tau)
        tr = 0.5 * curl_fd(lambda x: oseen(x - xj) @ e, xi)
        np.testing.assert_allclose(m[:3, 3 + c], tr, atol=1e-8 * scale)
        # rotation from a torque on j: (1/4) curl curl (G tau)
        rr = 0.5 * curl_fd(lambda y: 0.5 * curl_fd(lambda x: oseen(x - xj) @ e, y), xi)
        np.testing.assert_allclose(m[3:, 3 + c], rr, atol=1e-6 * scale)


@pytest.mark.parametrize("r", [0.3, 1.0, 1.55, 1.99])
def test_equal_radius_overlap_matches_rotne_prager_closed_forms(r):
    a = 1.0
    rhat = np.array([2.0, -1.0, 0.5])
    rhat /= np.linalg.norm(rhat)
    xj = np.array([1.0, 2.0, 3.0])
    xi = xj + r * rhat
    m = pair_matrix(6, 6, xi, a, xj, a)
    rr_t = np.outer(rhat, rhat)
    tt = ((1 - 9 * r / (32 * a)) * np.eye(3) + 3 * r / (32 * a) * rr_t) / (6 * np.pi * MU * a)
    rot = ((1 - 27 * r / (32 * a) + 5 * r**3 / (64 * a**3)) * np.eye(3)
           + (9 * r / (32 * a) - 3 * r**3 / (64 * a**3)) * rr_t) / (8 * np.pi * MU * a**3)
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    cross_rhat = np.einsum("ijk,k->ij", eps, rhat)   # (cross_rhat @ v) = v x rhat
    c_tr = r / (16 * np.pi * MU * a**3) * (1 - 3 * r / (8 * a))
    np.testing.assert_allclose(m[:3, :3], tt, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(m[3:, 3:], rot, rtol=1e-13, atol=1e-16)
    # v = c (tau x rhat);  w = c (F x rhat)
    np.testing.assert_allclose(m[:3, 3:], c_tr * cross_rhat, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(m[3:, :3], c_tr * cross_rhat, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("ai,aj", [(1.0, 1.7), (1.7, 1.0), (0.6, 1.0)])
def test_continuity_across_branch_boundaries(ai, aj):
    rhat = np.array([0.3, -0.8, 0.52])
    rhat /= np.linalg.norm(rhat)
    xj = np.zeros(3)
    for r0 in (ai + aj, abs(ai - aj)):
        lo = pair_matrix(6, 6, (r0 * (1 - 1e-9)) * rhat, ai, xj, aj)
        hi = pair_matrix(6, 6, (r0 * (1 + 1e-9)) * rhat, ai, xj, aj)
        assert np.abs(lo - hi).max() < 1e-7 * np.abs(hi).max(), (r0, np.abs(lo - hi).max())


def cluster(n_seg, n_blob, box, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-box, box, 3 * n_seg), rng.uniform(-box, box, 3 * n_blob)


@pytest.mark.parametrize("seg_dof", [6, 3])
def test_grand_mobility_symmetric_and_positive_definite(seg_dof):
    xs, xb = cluster(14, 11, 2.5, 3)   # dense: many overlapping pairs
    m = O.dense_mobility(xs, xb, seg_dof=seg_dof, a_seg=1.0, a_blob=0.8, mu=MU)
    assert np.abs(m - m.T).max() < 1e-14 * np.abs(m).max()
    np.linalg.cholesky(m)
    assert np.linalg.eigvalsh(m).min() > 0


def test_rotation_covariance_and_translation_invariance():
    xs, xb = cluster(6, 5, 3.0, 5)
    rng = np.random.default_rng(9)
    fs, fb = rng.normal(size=36), rng.normal(size=15)
    vs, vb = O.full_mobility(xs, xb, fs, fb, seg_dof=6, a_seg=1.0, a_blob=1.3, mu=MU)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    shift = np.array([10.0, -4.0, 7.0])
    rot = lambda a, k: (a.reshape(-1, k).reshape(-1, 3) @ q.T).reshape(-1)
    vs2, vb2 = O.full_mobility(rot(xs, 3) + np.tile(shift, 6), rot(xb, 3) + np.tile(shift, 5),
                               rot(fs, 6), rot(fb, 3), seg_dof=6, a_seg=1.0, a_blob=1.3, mu=MU)
    np.testing.assert_allclose(vs2, rot(vs, 6), atol=1e-13)
    np.testing.assert_allclose(vb2, rot(vb, 3), atol=1e-13)


@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (6, True), (3, True)])
def test_block_sums_match_dense_matrix_and_write_rules(seg_dof, bs_acc):
    xs, xb = cluster(9, 7, 3.0, 11)
    ns, nb = 9, 7
    rng = np.random.default_rng(2)
    fs, fb = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    m = O.dense_mobility(xs, xb, seg_dof=seg_dof, a_seg=1.0, a_blob=1.2, mu=MU)
    fvec = np.concatenate([fs.reshape(ns, 6)[:, :seg_dof].reshape(-1), fb])
    nsd = seg_dof * ns
    kw = dict(seg_dof=seg_dof, bs_accumulate=bs_acc, a_seg=1.0, a_blob=1.2, mu=MU)
    pre_s, pre_b = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    ss = O.block("ss", xs, xb, fs, fb, pre_s, **kw).reshape(ns, 6)
    sb = O.block("sb", xs, xb, fs, fb, pre_s, **kw).reshape(ns, 6)
    bb = O.block("bb", xs, xb, fs, fb, pre_b, **kw)
    bs = O.block("bs", xs, xb, fs, fb, pre_b, **kw)
    m_ss, m_sb = m[:nsd, :nsd], m[:nsd, nsd:]
    m_bs, m_bb = m[nsd:, :nsd], m[nsd:, nsd:]
    f_s = fvec[:nsd]
    np.testing.assert_allclose(ss[:, :seg_dof].reshape(-1), m_ss @ f_s, atol=1e-14)
    if seg_dof == 3:
        assert np.all(ss[:, 3:] == 0)       # omega written as 0 ('=' of a zero)
    np.testing.assert_allclose((sb - pre_s.reshape(ns, 6))[:, :seg_dof].reshape(-1), m_sb @ fb, atol=1e-14)
    np.testing.assert_allclose(bb, m_bb @ fb, atol=1e-14)
    expect_bs = m_bs @ f_s + (pre_b if bs_acc else 0)
    np.testing.assert_allclose(bs, expect_bs, atol=1e-14)
