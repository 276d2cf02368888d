"""GPU barrier forces (tfs_rpy_barrier_forces, cuda_functions.cu:1240-1436)
against the FP64 oracle: random clusters with overlapping filaments and blobs
(host and device memory, with and without the plane wall), the production
body, Newton's third law at full size, and the Python mirror's
apply_interparticle_forces."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.mobility import RpyHandle, RpyMobilitySolver

pytestmark = pytest.mark.gpu

S = 396.0


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cluster(seed, nfil, nseg, nb, box, lift=0.0):
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, box, (nfil, 3))
    dirs = rng.standard_normal((nfil, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    xs = base[:, None, :] + 1.9 * np.arange(nseg)[None, :, None] * dirs[:, None, :]
    xb = rng.uniform(0, box, (nb, 3))
    xs[..., 2] += lift
    xb[:, 2] += lift
    return xs.reshape(-1), xb.reshape(-1)


def gpu_barrier(xs, xb, f0, *, nseg, wall=False, dl=2.2, a_seg=1.0, a_blob=1.0, memory=L.MEMORY_HOST,
                with_blobs=True):
    ns, nb = xs.size // 3, xb.size // 3
    h = RpyHandle(ns, nb, seg_dof=6, memory=memory, seg_radius=a_seg, blob_radius=a_blob)
    try:
        if memory == L.MEMORY_HOST:
            h.upload(L.SEG_POSITIONS, xs)
            if nb:
                h.upload(L.BLOB_POSITIONS, xb)
            fs, rep = f0.copy(), np.zeros(3 * nb)
            h.barrier_forces(fs, rep if with_blobs and nb else None, strength=S, seg_per_fil=nseg, wall=wall, dl=dl)
            return fs, rep
        dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
        h.upload(L.SEG_POSITIONS, dev(xs))
        if nb:
            h.upload(L.BLOB_POSITIONS, dev(xb))
        fs, rep = dev(f0), torch.zeros(3 * nb, dtype=torch.float64, device="cuda")
        h.barrier_forces(fs, rep if with_blobs and nb else None, strength=S, seg_per_fil=nseg, wall=wall, dl=dl)
        h.sync()
        return fs.cpu().numpy(), rep.cpu().numpy()
    finally:
        h.close()


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
@pytest.mark.parametrize("wall,a_blob", [(False, 1.0), (True, 1.0), (False, 1.4)])
def test_cluster_matches_oracle(memory, wall, a_blob):
    print(f"case cluster memory={memory} wall={wall} a_blob={a_blob}", flush=True)
    nseg = 8
    xs, xb = cluster(3, 40, nseg, 400, 24.0, lift=0.3 if wall else 0.0)
    ns = xs.size // 3
    f0 = np.random.default_rng(4).standard_normal(6 * ns)
    fs, rep = gpu_barrier(xs, xb, f0, nseg=nseg, wall=wall, a_blob=a_blob, memory=memory)
    fs_o, rep_o = O.barrier_forces(xs, xb, f0, nseg=nseg, strength=S, a_blob=a_blob, wall=wall, dl=2.2)
    assert np.abs(fs_o - f0).max() > S                     # many pairs interact
    np.testing.assert_allclose(fs, fs_o, rtol=0, atol=1e-12 * S)
    np.testing.assert_allclose(rep, rep_o, rtol=0, atol=1e-12 * S)
    assert np.all(fs.reshape(-1, 6)[:, 3:] == f0.reshape(-1, 6)[:, 3:])


def test_segments_only_and_null_blob_output():
    xs, _ = cluster(5, 30, 10, 0, 15.0)
    ns = xs.size // 3
    fs, _ = gpu_barrier(xs, np.zeros(0), np.zeros(6 * ns), nseg=10)
    fs_o, _ = O.barrier_forces(xs, np.zeros(0), np.zeros(6 * ns), nseg=10, strength=S)
    np.testing.assert_allclose(fs, fs_o, rtol=0, atol=1e-12 * S)
    xs, xb = cluster(6, 10, 10, 100, 12.0)
    fs1, _ = gpu_barrier(xs, xb, np.zeros(60 * 10), nseg=10, with_blobs=False)
    fs2, _ = gpu_barrier(xs, xb, np.zeros(60 * 10), nseg=10)
    assert np.array_equal(fs1, fs2)


def test_production_body_and_newton_third_law():
    print("case production body 12800 segs + 9000 blobs", flush=True)
    body = G.make_body(640, 20, 9000, 50.0, sigma=0.3)
    xs, xb = body.x_segs, body.x_blobs
    ns, nb = xs.size // 3, xb.size // 3
    fs, rep = gpu_barrier(xs, xb, np.zeros(6 * ns), nseg=20)
    fs_o, rep_o = O.barrier_forces(xs, xb, np.zeros(6 * ns), nseg=20, strength=S)
    assert np.count_nonzero(fs_o) > 100
    np.testing.assert_allclose(fs, fs_o, rtol=0, atol=1e-12 * S)
    np.testing.assert_allclose(rep, rep_o, rtol=0, atol=1e-12 * S)
    total = fs.reshape(ns, 6)[:, :3].sum(0) + rep.reshape(nb, 3).sum(0)
    assert np.linalg.norm(total) <= 1e-10 * S


def test_mobility_solver_mirror_adds_barrier_to_device_forces():
    xs, xb = cluster(7, 12, 6, 60, 10.0)
    ns, nb = xs.size // 3, xb.size // 3
    f = np.random.default_rng(8).standard_normal(6 * ns)
    ms = RpyMobilitySolver(ns, nb, seg_dof=6)
    ms.initialise()
    try:
        ms.set_barrier(strength=S, seg_per_fil=6)
        ms.x_segs_host[:] = xs
        ms.x_blobs_host[:] = xb
        ms.f_segs_host[:] = f
        ms.f_blobs_host[:] = 0.0
        ms.copy_segment_positions_to_device()
        ms.copy_segment_forces_to_device()
        ms.copy_blob_positions_to_device()
        ms.copy_blob_forces_to_device()
        ms.apply_interparticle_forces()
        ms.copy_interparticle_blob_forces_to_host()
        ms.evaluate_full_mobility()
        ms.wait_for_device()
        assert np.array_equal(ms.f_segs_host, f)          # the host keeps the model's forces
        fs_o, rep_o = O.barrier_forces(xs, xb, f, nseg=6, strength=S)
        np.testing.assert_allclose(ms.f_blobs_repulsion_host, rep_o, rtol=0, atol=1e-12 * S)
        vs_o, vb_o = O.full_mobility(xs, xb, fs_o, np.zeros(3 * nb))
        assert rel(np.asarray(ms.v_segs_host), vs_o) <= 1e-12
        assert rel(np.asarray(ms.v_blobs_host), vb_o) <= 1e-12
    finally:
        ms.finalise()
