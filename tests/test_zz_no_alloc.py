"""tfs_rpy_apply_noalloc, the matvec for callers that cannot allocate (its
work list, partial sums, near-pair lists and grid in one file-scope device
array), against tfs_rpy_apply on the same inputs: two calls in a row, the
second with a smaller input, FP64 tiles to 1e-12 and FP32 tiles to 1e-6
relative, every write rule; sizes beyond its limit and host pointers are
rejected with a clear error."""
import ctypes

import numpy as np
import pytest
import torch

from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.mobility import RpyHandle

pytestmark = pytest.mark.gpu


def desc_for(ns, nb, seg_dof, precision, bs_acc, memory=L.MEMORY_DEVICE):
    d = L.RpyDesc()
    d.num_segs, d.num_blobs, d.seg_dof, d.precision = ns, nb, seg_dof, precision
    d.bs_mode = L.BS_ACCUMULATE if bs_acc else L.BS_OVERWRITE
    d.memory = memory
    d.seg_radius = d.blob_radius = d.mu = 1.0
    d.stream = None
    return d


def noalloc(body, blocks, seg_dof, precision, bs_acc, v_init):
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    xs, xb, fs, fb = (dev(a) for a in (body.x_segs, body.x_blobs, body.f_segs, body.f_blobs))
    vs, vb = dev(v_init[0]), dev(v_init[1])
    d = desc_for(body.num_segs, body.num_blobs, seg_dof, precision, bs_acc)
    L.check(L.lib().tfs_rpy_apply_noalloc(ctypes.byref(d), blocks, L.ptr(xs), L.ptr(xb), L.ptr(fs), L.ptr(fb),
                                          L.ptr(vs), L.ptr(vb)))
    torch.cuda.synchronize()
    return np.concatenate([vs.cpu().numpy(), vb.cpu().numpy()])


def with_handle(body, blocks, seg_dof, precision, bs_acc, v_init):
    h = RpyHandle(body.num_segs, body.num_blobs, seg_dof=seg_dof, precision=precision, bs_accumulate=bs_acc)
    try:
        h.upload(L.SEG_POSITIONS, body.x_segs)
        h.upload(L.BLOB_POSITIONS, body.x_blobs)
        vs, vb = v_init[0].copy(), v_init[1].copy()
        h.apply(blocks, body.f_segs, body.f_blobs, vs, vb)
        h.sync()
        return np.concatenate([vs, vb])
    finally:
        h.close()


@pytest.mark.parametrize("precision", [L.PRECISION_FP64, L.PRECISION_FP32_ACC64])
@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (3, True)])
@pytest.mark.parametrize("blocks", [L.BLOCK_ALL, L.BLOCK_SB | L.BLOCK_BS, L.BLOCK_BB])
def test_noalloc_matches_apply_twice_second_smaller(precision, seg_dof, bs_acc, blocks):
    tol = 1e-12 if precision == L.PRECISION_FP64 else 1e-6
    rng = np.random.default_rng(11)
    for body in (G.make_body(40, 20, 3000, 15.0), G.make_body(8, 10, 500, 8.0)):
        ns, nb = body.num_segs, body.num_blobs
        print(f"case noalloc precision {precision} seg_dof {seg_dof} blocks {blocks}: {ns} segments, {nb} blobs",
              flush=True)
        v_init = (rng.normal(size=6 * ns), rng.normal(size=3 * nb))   # read by accumulating blocks
        got = noalloc(body, blocks, seg_dof, precision, bs_acc, v_init)
        ref = with_handle(body, blocks, seg_dof, precision, bs_acc, v_init)
        ref64 = with_handle(body, blocks, seg_dof, L.PRECISION_FP64, bs_acc, v_init)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        err64 = np.linalg.norm(got - ref64) / np.linalg.norm(ref64)
        print(f"  relative difference to tfs_rpy_apply {err:.2e} (to its FP64 result {err64:.2e})", flush=True)
        assert err <= tol and err64 <= tol


def test_noalloc_rejects_what_it_cannot_hold():
    lib = L.lib()
    big = desc_for(30000, 3000, 6, L.PRECISION_FP64, False)
    rc = lib.tfs_rpy_apply_noalloc(ctypes.byref(big), L.BLOCK_ALL, None, None, None, None, None, None)
    assert rc == L.TFS_ERR_INVALID_ARGUMENT and b"at most 32768 particles" in lib.tfs_last_error()
    host = desc_for(10, 10, 6, L.PRECISION_FP64, False, memory=L.MEMORY_HOST)
    rc = lib.tfs_rpy_apply_noalloc(ctypes.byref(host), L.BLOCK_ALL, None, None, None, None, None, None)
    assert rc == L.TFS_ERR_INVALID_ARGUMENT and b"device pointers" in lib.tfs_last_error()
