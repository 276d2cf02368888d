"""GPU parity of the half-space matvec (tfs_rpy_set_domain, TFS_DOMAIN_HALF_SPACE:
no-slip wall at z = 0, cuda_functions.cu:12-203) against the oracle's
half-space block sums (pinned in tests/test_oracle_wall.py).

Bar: FP64 tiles <= 1e-12 relative L2 over the stacked velocities, for every
block combination's write rule; FP32 tiles (the wall tensor in float from
exact fixed-point separations, FP64 sums across source batches) <= 2e-6, the
FP32 rounding of the tensor's terms, whose image parts partly cancel near the
wall; radius / viscosity other than 1 are rejected, as the reference's wall
tensor assumes a = mu = 1."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.mobility import RpyHandle, matvec

pytestmark = pytest.mark.gpu


def above_wall(n, seed, box=8.0):
    """n particles with z in [1.5, 1.5 + 2 box), pairwise at least 2.05 apart."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = np.array([rng.uniform(-box, box), rng.uniform(-box, box), rng.uniform(1.5, 1.5 + 2 * box)])
        if all(np.linalg.norm(p - q) >= 2.05 for q in pts):
            pts.append(p)
    return np.array(pts).reshape(-1)


def rel(a, b):
    a, b = np.concatenate([np.ravel(x) for x in a]), np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (3, True)])
@pytest.mark.parametrize("ns,nb", [(40, 0), (0, 60), (30, 45), (300, 257)])
def test_half_space_matvec_matches_oracle(seg_dof, bs_acc, ns, nb):
    print(f"case wall seg_dof={seg_dof} ns={ns} nb={nb}", flush=True)
    x = above_wall(ns + nb, seed=ns * 3 + nb, box=8.0 if ns + nb < 200 else 18.0)
    xs, xb = x[:3 * ns], x[3 * ns:]
    rng = np.random.default_rng(7)
    fs, fb = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    vs, vb = matvec(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc, domain=L.DOMAIN_HALF_SPACE)
    with O.domain(O.HALF_SPACE):
        rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    assert rel([vs, vb], [rs, rb]) <= 1e-12
    # the wall changes the answer (unless the write rules leave nothing: with no
    # segments an overwriting BS replaces BB by M_bs 0)
    us, ub = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    if np.linalg.norm(np.concatenate([rs, rb])) > 0:
        assert rel([vs, vb], [us, ub]) > 1e-3


@pytest.mark.parametrize("blocks", [L.BLOCK_SS, L.BLOCK_SB, L.BLOCK_BB, L.BLOCK_BS, L.BLOCK_SB | L.BLOCK_BS])
def test_half_space_block_write_rules(blocks):
    print(f"case wall blocks={blocks}", flush=True)
    ns, nb = 25, 35
    x = above_wall(ns + nb, seed=5)
    xs, xb = x[:3 * ns], x[3 * ns:]
    rng = np.random.default_rng(8)
    fs, fb = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    pre_s, pre_b = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    vs, vb = matvec(xs, xb, fs, fb, seg_dof=6, bs_accumulate=False, blocks=blocks, v_segs=pre_s, v_blobs=pre_b,
                    domain=L.DOMAIN_HALF_SPACE)
    rs, rb = pre_s.copy(), pre_b.copy()
    with O.domain(O.HALF_SPACE):
        if blocks & L.BLOCK_SS:
            rs = O.block("ss", xs, xb, fs, fb, rs, seg_dof=6)
        if blocks & L.BLOCK_SB:
            rs = O.block("sb", xs, xb, fs, fb, rs, seg_dof=6)
        if blocks & L.BLOCK_BB:
            rb = O.block("bb", xs, xb, fs, fb, rb, seg_dof=6)
        if blocks & L.BLOCK_BS:
            rb = O.block("bs", xs, xb, fs, fb, rb, seg_dof=6, bs_accumulate=False)
    assert rel([vs, vb], [rs, rb]) <= 1e-12


@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (3, True)])
@pytest.mark.parametrize("ns,nb", [(40, 0), (30, 45), (300, 257)])
def test_half_space_fp32_tiles_match_oracle(seg_dof, bs_acc, ns, nb):
    print(f"case wall fp32 seg_dof={seg_dof} ns={ns} nb={nb}", flush=True)
    x = above_wall(ns + nb, seed=ns * 3 + nb, box=8.0 if ns + nb < 200 else 18.0)
    xs, xb = x[:3 * ns], x[3 * ns:]
    rng = np.random.default_rng(9)
    fs, fb = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    vs, vb = matvec(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc, precision="fp32",
                    domain=L.DOMAIN_HALF_SPACE)
    with O.domain(O.HALF_SPACE):
        rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    err = rel([vs, vb], [rs, rb])
    print(f"fp32 wall rel err {err:.3e}")
    assert err <= 2e-6
    # and they agree with the FP64 tiles to the same bar
    v64 = matvec(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc, domain=L.DOMAIN_HALF_SPACE)
    assert rel([vs, vb], list(v64)) <= 2e-6


def test_half_space_rejects_other_radii():
    h = RpyHandle(4, 4, seg_dof=6, seg_radius=1.3)
    try:
        with pytest.raises(L.TfsError):
            h.set_domain(L.DOMAIN_HALF_SPACE)
    finally:
        h.close()
    h = RpyHandle(4, 4, seg_dof=6)
    try:
        with pytest.raises(L.TfsError):
            h.set_domain(7)
    finally:
        h.close()
