"""Independent bodies in one launch (block-diagonal grand mobility): every
body's velocities match the FP64 oracle run on that body alone; the
persistent launch, the source split, every autotuned shape and hipGraph
replays give the same numbers; per-body completion times come back for every
body."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.ensemble import make_ensemble
from timfilsimmod_amd.mobility import RpyHandle

pytestmark = pytest.mark.gpu

SIZES = [(3, 40, 5.0), (1, 0, 3.0), (0, 70, 6.0), (5, 120, 7.0), (2, 33, 4.0)]


def per_body_reference(e, seg_dof):
    vs, vb = [], []
    for b in e.bodies:
        s, bl = O.full_mobility(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=seg_dof, bs_accumulate=True)
        vs.append(s)
        vb.append(bl)
    return np.concatenate(vs), np.concatenate(vb)


def handle_for(e, seg_dof, precision=L.PRECISION_FP64):
    h = RpyHandle(e.num_segs, e.num_blobs, seg_dof=seg_dof, precision=precision, bs_accumulate=True)
    h.set_bodies(e.segs_per_body, e.blobs_per_body)
    h.upload(L.SEG_POSITIONS, e.x_segs)
    h.upload(L.BLOB_POSITIONS, e.x_blobs)
    h.upload(L.SEG_FORCES, e.f_segs)
    h.upload(L.BLOB_FORCES, e.f_blobs)
    return h


def read(h, e):
    vs, vb = np.zeros(6 * e.num_segs), np.zeros(3 * e.num_blobs)
    h.download(L.SEG_VELOCITIES, vs)
    h.download(L.BLOB_VELOCITIES, vb)
    h.sync()
    return np.concatenate([vs, vb])


@pytest.mark.parametrize("seg_dof", [6, 3])
@pytest.mark.parametrize("persistent", [0, 2])
def test_batched_bodies_match_per_body_oracle(seg_dof, persistent):
    e = make_ensemble(sizes=SIZES, num_seg=6)
    ref = np.concatenate(per_body_reference(e, seg_dof))
    h = handle_for(e, seg_dof)
    try:
        h.set_tile_shape(0, 64, 1, 64)
        h.set_persistent(0, persistent)
        h.body_timing(True)
        h.evaluate()
        got = read(h, e)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        if persistent:
            t = h.body_times(len(SIZES))
            assert np.all(np.isfinite(t)) and np.all(t >= 0)
    finally:
        h.close()


def test_autotune_keeps_results_and_picks_a_valid_shape():
    e = make_ensemble(sizes=SIZES, num_seg=8)
    ref = np.concatenate(per_body_reference(e, 6))
    h = handle_for(e, 6)
    try:
        for prec, tol in ((L.PRECISION_FP64, 1e-12), (L.PRECISION_FP32_ACC64, 1e-6)):
            best = h.autotune(prec, reps=2)
            shape = h.tile_shape(prec)
            print(prec, shape, best)
            assert shape["block"] in (64, 128, 256, 512) and shape["per_thread"] in (1, 2) and best > 0
            assert shape["split"] == 1 or shape["block"] != 512
            h.set_precision(prec)
            h.evaluate()
            got = read(h, e)
            assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
    finally:
        h.close()


@pytest.mark.parametrize("persistent", [0, 2, "split"])
def test_graph_replays_match_eager_every_time(persistent):
    e = make_ensemble(sizes=SIZES, num_seg=6)
    h = handle_for(e, 6)
    try:
        if persistent == "split":
            h.set_split_shape(0, 256, 2, 256)
        else:
            h.set_tile_shape(0, 256, 2, 256)
            h.set_persistent(0, persistent)
        h.evaluate()
        eager = read(h, e)
        g = h.graph(L.BLOCK_ALL, 2)
        rng = np.random.default_rng(0)
        for k in range(4):
            # eager work between replays: overwrite the outputs with noise
            h.upload(L.SEG_VELOCITIES, rng.normal(size=6 * e.num_segs))
            h.upload(L.BLOB_VELOCITIES, rng.normal(size=3 * e.num_blobs))
            h.sync()
            g.launch()
            assert np.array_equal(read(h, e), eager), k
        g.close()
    finally:
        h.close()
