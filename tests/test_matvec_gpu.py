"""GPU parity of the fused RPY matvec (libtfs_rpy.so through the C-ABI) against
the FP64 oracle (oracle/rpy_oracle.c) on the same seeded inputs.

Bars: FP64 tiles <= 1e-12 relative L2 over the stacked velocity vector;
FP32 tiles with FP64 accumulation <= 1e-6 relative L2 (a stated tolerance of
the mixed path, not bit parity).  Write rules per block are checked with
pre-filled velocity arrays.
"""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.mobility import RpyMobilitySolver, matvec

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-6


def rel(a, b):
    a, b = np.concatenate([np.ravel(x) for x in a]), np.concatenate([np.ravel(x) for x in b])
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def cluster(ns, nb, box, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-box, box, 3 * ns), rng.uniform(-box, box, 3 * nb),
            rng.normal(size=6 * ns), rng.normal(size=3 * nb))


@pytest.mark.parametrize("seg_dof", [6, 3])
@pytest.mark.parametrize("ns,nb", [(37, 53), (300, 1), (1, 300), (0, 129), (129, 0), (257, 511)])
def test_full_matvec_fp64_matches_oracle_small(seg_dof, ns, nb):
    print(f"case seg_dof={seg_dof} ns={ns} nb={nb}", flush=True)
    xs, xb, fs, fb = cluster(ns, nb, 6.0, ns * 7 + nb)
    kw = dict(seg_dof=seg_dof, bs_accumulate=(seg_dof == 3), seg_radius=1.0, blob_radius=1.3, mu=0.9)
    vs, vb = matvec(xs, xb, fs, fb, **kw)
    rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=(seg_dof == 3),
                             a_seg=1.0, a_blob=1.3, mu=0.9)
    assert rel([vs, vb], [rs, rb]) <= FP64_TOL
    if seg_dof == 3 and ns:
        assert np.all(vs.reshape(ns, 6)[:, 3:] == 0.0)


@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (6, True), (3, True)])
@pytest.mark.parametrize("blocks", [L.BLOCK_SS, L.BLOCK_SB, L.BLOCK_BB, L.BLOCK_BS,
                                    L.BLOCK_SS | L.BLOCK_SB, L.BLOCK_BB | L.BLOCK_BS,
                                    L.BLOCK_SB | L.BLOCK_BS])
def test_block_write_rules_match_reference_sequence(seg_dof, bs_acc, blocks):
    print(f"case blocks={blocks} seg_dof={seg_dof} bs_acc={bs_acc}", flush=True)
    ns, nb = 70, 90
    xs, xb, fs, fb = cluster(ns, nb, 5.0, 21)
    rng = np.random.default_rng(4)
    pre_s, pre_b = rng.normal(size=6 * ns), rng.normal(size=3 * nb)
    vs, vb = matvec(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc, blocks=blocks,
                    v_segs=pre_s, v_blobs=pre_b)
    # The reference sequence: Mss '=' then Msb '+='; Mbb '=' then Mbs '+=' or '='.
    kw = dict(seg_dof=seg_dof, bs_accumulate=bs_acc)
    es, eb = pre_s.copy(), pre_b.copy()
    if blocks & L.BLOCK_SS:
        es = O.block("ss", xs, xb, fs, fb, es, **kw)
    if blocks & L.BLOCK_SB:
        es = O.block("sb", xs, xb, fs, fb, es, **kw)
    if blocks & L.BLOCK_BB:
        eb = O.block("bb", xs, xb, fs, fb, eb, **kw)
    if blocks & L.BLOCK_BS:
        eb = O.block("bs", xs, xb, fs, fb, eb, **kw)
    assert rel([vs, vb], [es, eb]) <= FP64_TOL


def test_small_body_fp64_and_fp32():
    b = G.small_body()
    ref = O.full_mobility(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=6)
    v64 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=6, precision="fp64")
    v32 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=6, precision="fp32")
    e64, e32 = rel(v64, ref), rel(v32, ref)
    print(f"small body: fp64 rel err {e64:.3e}, fp32 rel err {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32_TOL


@pytest.mark.parametrize("seg_dof", [6, 3])
def test_production_body_fp64_and_fp32(seg_dof):
    b = G.production_body()
    ref = O.full_mobility(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=seg_dof,
                          bs_accumulate=(seg_dof == 3))
    kw = dict(seg_dof=seg_dof, bs_accumulate=(seg_dof == 3))
    v64 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, precision="fp64", **kw)
    v32 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, precision="fp32", **kw)
    e64, e32 = rel(v64, ref), rel(v32, ref)
    print(f"production body seg_dof={seg_dof}: fp64 rel err {e64:.3e}, fp32 rel err {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32_TOL


@pytest.mark.parametrize("shape", [(64, 1, 64), (64, 2, 128), (128, 1, 384), (128, 2, 256),
                                   (256, 1, 256), (256, 2, 1024)])
def test_every_tile_shape_agrees(shape):
    xs, xb, fs, fb = cluster(333, 444, 9.0, 5)
    ref = O.full_mobility(xs, xb, fs, fb, seg_dof=6)
    for prec, tol in (("fp64", FP64_TOL), ("fp32", FP32_TOL)):
        v = matvec(xs, xb, fs, fb, seg_dof=6, precision=prec, tile_shape=shape)
        assert rel(v, ref) <= tol, (shape, prec)


def test_deterministic_run_to_run():
    b = G.small_body()
    a1 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs)
    a2 = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs)
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])


def test_reference_shaped_solver_surface():
    """copy_* / evaluate_* / wait_for_device on the pinned host arrays."""
    b = G.small_body()
    m = RpyMobilitySolver(b.num_segs, b.num_blobs, seg_dof=6)
    m.initialise()
    try:
        m.x_segs_host[:] = b.x_segs
        m.f_segs_host[:] = b.f_segs
        m.x_blobs_host[:] = b.x_blobs
        m.f_blobs_host[:] = b.f_blobs
        m.copy_segment_positions_to_device()
        m.copy_segment_forces_to_device()
        m.copy_blob_positions_to_device()
        m.copy_blob_forces_to_device()
        m.evaluate_segment_segment_mobility()
        m.evaluate_segment_blob_mobility()
        m.copy_segment_velocities_to_host()
        m.evaluate_blob_segment_mobility()   # '=' in this (elastic, GMRES) build
        m.copy_blob_velocities_to_host()
        m.wait_for_device()
        rs = O.block("ss", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs)
        rs = O.block("sb", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, rs)
        rb = O.block("bs", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs)
        assert rel([m.v_segs_host, m.v_blobs_host], [rs, rb]) <= FP64_TOL
    finally:
        m.finalise()


def test_linearity_and_symmetry_at_production_size():
    """Size-independent properties: M(a f + g) = a Mf + Mg and <u, M v> = <M u, v>."""
    b = G.production_body()
    rng = np.random.default_rng(12)
    g_s, g_b = rng.normal(size=b.f_segs.size), rng.normal(size=b.f_blobs.size)
    # the full grand mobility: Mbs accumulates onto Mbb (cuda_functions.cu:739-743)
    mf = matvec(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, bs_accumulate=True)
    mg = matvec(b.x_segs, b.x_blobs, g_s, g_b, bs_accumulate=True)
    mc = matvec(b.x_segs, b.x_blobs, 2.5 * b.f_segs + g_s, 2.5 * b.f_blobs + g_b, bs_accumulate=True)
    assert rel(mc, [2.5 * mf[0] + mg[0], 2.5 * mf[1] + mg[1]]) <= 1e-13
    lhs = np.dot(b.f_segs, mg[0]) + np.dot(b.f_blobs, mg[1])
    rhs = np.dot(g_s, mf[0]) + np.dot(g_b, mf[1])
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
