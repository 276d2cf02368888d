"""The source-split launch shapes of the pair kernel (tfs_rpy_set_split_shape:
the waves of a workgroup share one tile of 64 x per_thread targets and split
every LDS batch of sources, adding their sums in LDS in wave order; the
default of both precisions) against the FP64 oracle (oracle/rpy_oracle.c
full_mobility) on the same seeded inputs:

* one body, 6-DOF and 3-DOF, every split shape (block 256/512, 1 or 2 targets
  per lane) and the plain form, FP64 tiles <= 1e-12 and FP32 tiles <= 1e-6
  relative L2, with chunks that leave later waves of the last batch without
  sources;
* independent bodies (block-diagonal mobility) in one launch;
* repeated evaluations are bitwise identical (a fixed summation order);
* the argument rules: no persistent form, block 256 or 512 with the split,
  64/128/256 without;
* the default shapes: the split for a body that fills the device (12,800
  segments + 9,000 blobs; with 3-DOF segments the symmetric form comes
  first), one wave per 64 x 32 item for one that does not."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.ensemble import make_ensemble
from timfilsimmod_amd.mobility import RpyHandle

pytestmark = pytest.mark.gpu

TOL = {L.PRECISION_FP64: 1e-12, L.PRECISION_FP32_ACC64: 1e-6}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cluster(ns, nb, box, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-box, box, 3 * ns), rng.uniform(-box, box, 3 * nb),
            rng.normal(size=6 * ns), rng.normal(size=3 * nb))


def handle(ns, nb, xs, xb, fs, fb, seg_dof, precision):
    h = RpyHandle(ns, nb, seg_dof=seg_dof, precision=precision, bs_accumulate=True)
    h.upload(L.SEG_POSITIONS, xs)
    h.upload(L.BLOB_POSITIONS, xb)
    h.upload(L.SEG_FORCES, fs)
    h.upload(L.BLOB_FORCES, fb)
    return h


def read(h, ns, nb):
    vs, vb = np.zeros(6 * ns), np.zeros(3 * nb)
    h.download(L.SEG_VELOCITIES, vs)
    h.download(L.BLOB_VELOCITIES, vb)
    h.sync()
    return np.concatenate([vs, vb])


@pytest.mark.parametrize("precision", [L.PRECISION_FP64, L.PRECISION_FP32_ACC64])
@pytest.mark.parametrize("seg_dof", [6, 3])
def test_split_shapes_match_oracle(seg_dof, precision):
    ns, nb = 700, 530
    xs, xb, fs, fb = cluster(ns, nb, 9.0, 11 + seg_dof)
    rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=True)
    ref = np.concatenate([rs, rb])
    h = handle(ns, nb, xs, xb, fs, fb, seg_dof, precision)
    try:
        h.evaluate()
        assert rel(read(h, ns, nb), ref) <= TOL[precision]
        h.set_tile_shape(precision, 256, 2, 256)
        assert h.tile_shape(precision)["split"] == 0
        h.evaluate()
        assert rel(read(h, ns, nb), ref) <= TOL[precision]
        # chunks of 768 (block 256) and 1536 (block 512): the last batch of the
        # 700-segment range holds 188 sources, so wave 3 (block 256) or waves
        # 3..7 (block 512) get none of it
        for block, r, chunk in ((256, 1, 256), (256, 2, 512), (512, 1, 512), (512, 2, 1024), (256, 2, 768),
                                (512, 2, 1536)):
            print(f"case split block {block} per_thread {r} chunk {chunk} seg_dof {seg_dof} precision {precision}",
                  flush=True)
            h.set_split_shape(precision, block, r, chunk)
            assert h.tile_shape(precision)["split"] == 1
            h.evaluate()
            got = read(h, ns, nb)
            err = rel(got, ref)
            print(f"  rel err {err:.2e}", flush=True)
            assert err <= TOL[precision]
            h.evaluate()
            assert np.array_equal(read(h, ns, nb), got)
    finally:
        h.close()


@pytest.mark.parametrize("precision", [L.PRECISION_FP64, L.PRECISION_FP32_ACC64])
def test_split_independent_bodies_match_per_body_oracle(precision):
    e = make_ensemble(sizes=[(3, 40, 5.0), (1, 0, 3.0), (0, 70, 6.0), (5, 120, 7.0), (2, 33, 4.0)], num_seg=6)
    print(f"case split, {len(e.bodies)} independent bodies, precision {precision}", flush=True)
    vs, vb = [], []
    for b in e.bodies:
        s, bl = O.full_mobility(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=6, bs_accumulate=True)
        vs.append(s)
        vb.append(bl)
    ref = np.concatenate(vs + vb)
    h = RpyHandle(e.num_segs, e.num_blobs, seg_dof=6, precision=precision, bs_accumulate=True)
    try:
        h.set_bodies(e.segs_per_body, e.blobs_per_body)
        h.upload(L.SEG_POSITIONS, e.x_segs)
        h.upload(L.BLOB_POSITIONS, e.x_blobs)
        h.upload(L.SEG_FORCES, e.f_segs)
        h.upload(L.BLOB_FORCES, e.f_blobs)
        h.set_split_shape(precision, 256, 2, 256)
        h.evaluate()
        got = read(h, e.num_segs, e.num_blobs)
        assert rel(got, ref) <= TOL[precision]
    finally:
        h.close()


def test_split_argument_rules():
    ns, nb = 64, 64
    xs, xb, fs, fb = cluster(ns, nb, 5.0, 3)
    h = handle(ns, nb, xs, xb, fs, fb, 6, L.PRECISION_FP64)
    p = L.PRECISION_FP64
    try:
        for bad in ((128, 1, 128), (64, 1, 64), (512, 3, 512), (512, 1, 256), (1024, 1, 1024)):
            with pytest.raises(RuntimeError):
                h.set_split_shape(p, *bad)
        with pytest.raises(RuntimeError):
            h.set_tile_shape(p, 512, 1, 512)           # 512 threads only with the split
        h.set_split_shape(p, 256, 1, 256)
        with pytest.raises(RuntimeError):
            h.set_persistent(p, 2)                      # a chosen split shape has no persistent form
        h.set_tile_shape(p, 128, 1, 128)
        h.set_persistent(p, 2)
        with pytest.raises(RuntimeError):
            h.set_split_shape(p, 256, 1, 256)
        h.set_persistent(p, 0)
        h.set_split_shape(p, 512, 2, 512)
        assert h.tile_shape(p) == {"block": 512, "per_thread": 2, "chunk": 512, "persistent": 0, "split": 1}
        h.evaluate()
        got = read(h, ns, nb)
        rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=6, bs_accumulate=True)
        assert rel(got, np.concatenate([rs, rb])) <= TOL[p]
    finally:
        h.close()
    # a default split (or symmetric) shape gives way to the plain default for a persistent launch
    big = RpyHandle(12800, 9000, seg_dof=3, precision=p, bs_accumulate=True)
    try:
        assert big.symmetric(p)
        big.set_persistent(p, 2)
        assert big.tile_shape(p) == {"block": 256, "per_thread": 2, "chunk": 256, "persistent": 2, "split": 0}
    finally:
        big.close()


@pytest.mark.parametrize("seg_dof", [3, 6])
def test_default_shapes(seg_dof):
    big = RpyHandle(12800, 9000, seg_dof=seg_dof, precision=L.PRECISION_FP32_ACC64, bs_accumulate=True)
    small = RpyHandle(320, 2000, seg_dof=seg_dof, precision=L.PRECISION_FP64, bs_accumulate=True)
    try:
        f32, f64 = big.tile_shape(L.PRECISION_FP32_ACC64), big.tile_shape(L.PRECISION_FP64)
        print("12,800 + 9,000:", f32, f64, flush=True)
        if seg_dof == 3:
            # the symmetric form (tests/test_symmetric_gpu.py) is the default of
            # a large handle with 3-DOF segments; the split its alternative
            for p in (L.PRECISION_FP32_ACC64, L.PRECISION_FP64):
                assert big.symmetric(p)
                assert big.tile_shape(p) == {"block": 256, "per_thread": 1, "chunk": 0, "persistent": 0, "split": 0}
                big.set_symmetric(p, False)
            f32, f64 = big.tile_shape(L.PRECISION_FP32_ACC64), big.tile_shape(L.PRECISION_FP64)
        assert not big.symmetric(L.PRECISION_FP32_ACC64) and not big.symmetric(L.PRECISION_FP64)
        assert f32 == {"block": 256, "per_thread": 1, "chunk": 768, "persistent": 0, "split": 1}
        assert f64 == {"block": 256, "per_thread": 2, "chunk": 512, "persistent": 0, "split": 1}
        s64 = small.tile_shape(L.PRECISION_FP64)
        print("320 + 2,000:", s64, flush=True)
        assert s64 == {"block": 64, "per_thread": 1, "chunk": 32, "persistent": 0, "split": 0}
    finally:
        big.close()
        small.close()
