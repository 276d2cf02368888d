"""The float (*_f32) entry points of the C-ABI, for the reference's Real =
float build (config.hpp:36,52), against the FP64 oracle run on the same
float-rounded inputs: the outputs must be the FP64 results rounded once to
float (within half a float ulp, 2^-24 relative, plus the FP64 path's own
1e-12), in host and device memory."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.flow import STOKESLET_ROTLET, FlowDesc
from timfilsimmod_amd.mobility import RpyHandle
from timfilsimmod_amd.solver import (BLOB_DIRECT, SOLVE_FP64, ElasticVelocities, SolveStats, Solver, beat_setup,
                                     solver_desc)

pytestmark = pytest.mark.gpu

ULP = 2.0 ** -24


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def rounded(a):
    """The FP64 value a float array holds."""
    return f32(a).astype(np.float64)


def buf(a, memory):
    """A float32 array in the requested memory space."""
    a = f32(a)
    return torch.from_numpy(a.copy()).cuda() if memory == L.MEMORY_DEVICE else a.copy()


def host(a):
    return a.cpu().numpy().astype(np.float64) if hasattr(a, "cpu") else np.asarray(a, dtype=np.float64)


def assert_rounded_once(got, ref, extra=1e-12):
    got, ref = host(got), np.asarray(ref, dtype=np.float64)
    bound = ULP * np.abs(ref) + extra * np.abs(ref).max()
    bad = np.abs(got - ref) > bound
    assert not bad.any(), (np.abs(got - ref)[bad][:5], bound[bad][:5])


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
@pytest.mark.parametrize("seg_dof", [6, 3])
def test_rpy_f32_upload_apply_download(memory, seg_dof):
    print(f"case rpy f32, memory {memory}, seg_dof {seg_dof}", flush=True)
    b = G.make_body(8, 10, 400, 9.0)
    ns, nb = b.num_segs, b.num_blobs
    xs, xb, fs, fb = (rounded(a) for a in (b.x_segs, b.x_blobs, b.f_segs, b.f_blobs))
    h = RpyHandle(ns, nb, seg_dof=seg_dof, precision=L.PRECISION_FP64, bs_accumulate=seg_dof == 3, memory=memory)
    lib = L.lib()
    try:
        arrs = {k: buf(a, memory) for k, a in (("xs", xs), ("xb", xb), ("fs", fs), ("fb", fb))}
        L.check(lib.tfs_rpy_upload_f32(h.h, L.SEG_POSITIONS, L.ptr(arrs["xs"])))
        L.check(lib.tfs_rpy_upload_f32(h.h, L.BLOB_POSITIONS, L.ptr(arrs["xb"])))
        L.check(lib.tfs_rpy_upload_f32(h.h, L.SEG_FORCES, L.ptr(arrs["fs"])))
        L.check(lib.tfs_rpy_upload_f32(h.h, L.BLOB_FORCES, L.ptr(arrs["fb"])))
        L.check(lib.tfs_rpy_evaluate(h.h, L.BLOCK_ALL))
        vs, vb = buf(np.zeros(6 * ns), memory), buf(np.zeros(3 * nb), memory)
        L.check(lib.tfs_rpy_download_f32(h.h, L.SEG_VELOCITIES, L.ptr(vs)))
        L.check(lib.tfs_rpy_download_f32(h.h, L.BLOB_VELOCITIES, L.ptr(vb)))
        h.sync()
        rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=seg_dof == 3)
        assert_rounded_once(vs, rs)
        assert_rounded_once(vb, rb)
        # apply with an accumulating block: v_segs (given) += M_sb f_blobs
        v0 = np.random.default_rng(2).normal(size=6 * ns)
        vs2 = buf(v0, memory)
        L.check(lib.tfs_rpy_apply_f32(h.h, L.BLOCK_SB, None, L.ptr(arrs["fb"]), L.ptr(vs2), None))
        h.sync()
        ref = O.block("sb", xs, xb, fs, fb, rounded(v0), seg_dof=seg_dof)
        assert_rounded_once(vs2, ref)
        # the handle's FP32 tiles behind the same float arrays
        h.set_precision(L.PRECISION_FP32_ACC64)
        vs3, vb3 = buf(np.zeros(6 * ns), memory), buf(np.zeros(3 * nb), memory)
        L.check(lib.tfs_rpy_apply_f32(h.h, L.BLOCK_ALL, L.ptr(arrs["fs"]), L.ptr(arrs["fb"]), L.ptr(vs3), L.ptr(vb3)))
        h.sync()
        got = np.concatenate([host(vs3), host(vb3)])
        ref = np.concatenate([rs, rb])
        assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)
    finally:
        h.close()


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_rpy_f32_barrier_forces(memory):
    b = G.make_body(6, 10, 300, 7.0)
    ns, nb = b.num_segs, b.num_blobs
    # pull the filaments in so that segments and blobs touch
    xs = rounded(b.x_segs.reshape(-1, 3) * 0.97).reshape(-1)
    xb, fs = rounded(b.x_blobs), rounded(b.f_segs)
    h = RpyHandle(ns, nb, seg_dof=6, precision=L.PRECISION_FP64, memory=memory)
    try:
        keep = [buf(xs, memory), buf(xb, memory)]   # alive until the asynchronous uploads ran
        L.check(L.lib().tfs_rpy_upload_f32(h.h, L.SEG_POSITIONS, L.ptr(keep[0])))
        L.check(L.lib().tfs_rpy_upload_f32(h.h, L.BLOB_POSITIONS, L.ptr(keep[1])))
        d = L.BarrierDesc(300.0, 10, 0, 2.2)
        f_io, f_rep = buf(fs, memory), buf(np.zeros(3 * nb), memory)
        L.check(L.lib().tfs_rpy_barrier_forces_f32(h.h, ctypes.byref(d), L.ptr(f_io), L.ptr(f_rep)))
        h.sync()
        ref_s, ref_b = O.barrier_forces(xs, xb, fs, nseg=10, strength=300.0)
        assert np.abs(ref_s - fs).max() > 1.0 and np.abs(ref_b).max() > 1.0   # some pairs touch
        assert_rounded_once(f_io, ref_s, extra=1e-12)
        assert_rounded_once(f_rep, ref_b, extra=1e-12)
    finally:
        h.close()


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_solver_f32_entry_points(memory):
    nfil, nseg, nblob = 6, 12, 150
    body = G.make_body(nfil, nseg, nblob, 7.0)
    su = beat_setup(body)
    geo = SO.beat_geometry(su, nseg, B.BEAT_A, B.BEAT_B)
    xs, xb, kp, kt = (rounded(a) for a in (geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"]))
    s = Solver(solver_desc(nfil, nseg, nblob, tol=1e-10, restart=80, max_cycles=20))
    lib = L.lib()
    try:
        geo_bufs = [buf(a, memory) for a in (xs, xb, kp, kt)]   # alive until the calls ran
        L.check(lib.tfs_solver_set_geometry_f32(s.h, *[L.ptr(a) for a in geo_bufs], memory))
        A = SO.saddle_matrix(xs, xb, kp, kt, nfil, nseg)
        v = rounded(np.random.default_rng(4).normal(size=s.n))
        vin, out = buf(v, memory), buf(np.zeros(s.n), memory)
        L.check(lib.tfs_solver_system_mult_f32(s.h, L.ptr(vin), L.ptr(out), L.PRECISION_FP64, memory))
        assert_rounded_once(out, A @ v, extra=1e-12)
        pout = buf(np.zeros(s.n), memory)
        L.check(lib.tfs_solver_precondition_f32(s.h, L.ptr(vin), L.ptr(pout), memory))
        assert_rounded_once(pout, SO.precondition_ref(v, kp, kt, nfil, nseg, nblob), extra=1e-12)
        rhs = v
        x = buf(np.zeros(s.n), memory)
        st = SolveStats()
        L.check(lib.tfs_solver_solve_f32(s.h, L.ptr(vin), L.ptr(x), memory, ctypes.byref(st)))
        exact = np.linalg.solve(A, rhs)
        assert st.converged == 1
        # the FP64 solution (to 1e-10) rounded once to float
        assert_rounded_once(x, exact, extra=1e-8)
        rel = ctypes.c_double()
        L.check(lib.tfs_solver_residual_f32(s.h, L.ptr(vin), L.ptr(x), L.PRECISION_FP64, memory, ctypes.byref(rel)))
        xr = host(x)
        assert abs(rel.value - np.linalg.norm(rhs - A @ xr) / np.linalg.norm(rhs)) <= 1e-9
    finally:
        s.close()


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_flow_f32_entry_points(memory):
    b = G.make_body(6, 10, 300, 8.0)
    xs, fs, xb, fb = (rounded(a) for a in (b.x_segs, b.f_segs, b.x_blobs, b.f_blobs))
    pts = rounded(np.random.default_rng(6).uniform(-30, 30, size=(2000, 3)).reshape(-1))
    d = FlowDesc(1, b.num_segs, b.num_blobs, STOKESLET_ROTLET, L.PRECISION_FP64, 0, 1.0, 1.0, 1.0, memory, None)
    h = ctypes.c_void_p()
    lib = L.lib()
    L.check(lib.tfs_flow_create(ctypes.byref(d), ctypes.byref(h)))
    try:
        srcs = [buf(a, memory) for a in (xs, fs, xb, fb)]   # alive until the calls ran
        L.check(lib.tfs_flow_set_sources_f32(h, *[L.ptr(a) for a in srcs], None))
        p, v = buf(pts, memory), buf(np.zeros(pts.size), memory)
        L.check(lib.tfs_flow_evaluate_f32(h, pts.size // 3, L.ptr(p), L.ptr(v)))
        L.check(lib.tfs_flow_sync(h))
        ref = O.flow_velocities(pts, xs, fs, xb, fb, np.zeros(3))
        assert_rounded_once(v, ref, extra=1e-12)
    finally:
        lib.tfs_flow_destroy(h)


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_velocities_f32_matches_fp64_entry_point(memory):
    """tfs_velocities_compute_f32 against tfs_velocities_compute on the same
    float-rounded inputs (direct blob solve): outputs rounded once."""
    b = G.make_body(4, 10, 300, 8.0)
    ns, nb = b.num_segs, b.num_blobs
    xs, fs, xb = (rounded(a) for a in (b.x_segs, b.f_segs, b.x_blobs))
    xbod, vbod = rounded(b.centre), rounded([0.1, -0.2, 0.05, 0.01, 0.0, -0.02])
    ref_v, ref_f = np.zeros(6 * ns), np.zeros(3 * nb)
    ev64 = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=50, tol=1e-12)
    ev64.set_preconditioner(BLOB_DIRECT)
    try:
        ref_v, ref_f, _ = ev64.compute(xs, fs, xb, xbod, vbod)
    finally:
        ev64.close()
    ev = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=50, tol=1e-12, memory=memory)
    ev.set_preconditioner(BLOB_DIRECT)
    try:
        vs, fbl = buf(np.zeros(6 * ns), memory), buf(np.zeros(3 * nb), memory)
        st = SolveStats()
        ins = [buf(a, memory) for a in (xs, fs, xb, xbod, vbod)]   # alive until the call ran
        L.check(L.lib().tfs_velocities_compute_f32(ev.h, *[L.ptr(a) for a in ins], L.ptr(vs), L.ptr(fbl),
                                                   ctypes.byref(st)))
        torch.cuda.synchronize()
        assert_rounded_once(vs, ref_v, extra=1e-11)
        assert_rounded_once(fbl, ref_f, extra=1e-11)
    finally:
        ev.close()
