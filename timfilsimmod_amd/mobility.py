"""Host-side mirror of the reference's RPY mobility back-end over the C-ABI.

``RpyMobilitySolver`` has the method names and host-array members of
``rpy_mobility_solver`` (src/cilia/mobility/rpy_mobility_solver.cu:1-337, the
virtual surface of src/cilia/mobility/mobility_solver.hpp:26-63): callers fill
``x_segs_host`` / ``f_segs_host`` / ``x_blobs_host`` / ``f_blobs_host`` in place,
call ``copy_*_to_device``, ``evaluate_*_mobility``, ``copy_*_to_host`` and
``wait_for_device``, and read ``v_segs_host`` / ``v_blobs_host``.  Host arrays
are pinned (the reference's cudaHostAlloc) and exposed as numpy views.

Every call goes to libtfs_rpy.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class PinnedArray:
    """A pinned host allocation viewed as a float64 numpy array."""

    def __init__(self, n: int):
        self._p = ctypes.c_void_p()
        L.check(L.lib().tfs_host_alloc(max(n, 1) * 8, ctypes.byref(self._p)))
        buf = (ctypes.c_double * max(n, 1)).from_address(self._p.value)
        self.array = np.frombuffer(buf, dtype=np.float64, count=n)
        self.array[:] = 0.0

    def free(self):
        if self._p:
            L.lib().tfs_host_free(self._p)
            self._p = ctypes.c_void_p()


def profile_read(h):
    ms = (ctypes.c_double * 2)()
    n = (ctypes.c_int32 * 2)()
    L.check(L.lib().tfs_rpy_profile_read(h, ms, n))
    return {L.PRECISION_FP64: (ms[0], n[0]), L.PRECISION_FP32_ACC64: (ms[1], n[1])}


class RpyHandle:
    """Thin owner of one ``tfs_rpy`` handle."""

    def __init__(self, num_segs: int, num_blobs: int, *, seg_dof: int = 6, precision: int = 0,
                 bs_accumulate: bool = False, seg_radius: float = 1.0, blob_radius: float = 1.0,
                 mu: float = 1.0, memory: int = L.MEMORY_HOST, stream: int | None = None):
        d = L.RpyDesc()
        d.num_segs = num_segs
        d.num_blobs = num_blobs
        d.seg_dof = seg_dof
        d.precision = precision
        d.bs_mode = L.BS_ACCUMULATE if bs_accumulate else L.BS_OVERWRITE
        d.memory = memory
        d.seg_radius = seg_radius
        d.blob_radius = blob_radius
        d.mu = mu
        d.stream = stream
        self.desc = d
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_rpy_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tfs_rpy_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, what: int, src):
        L.check(L.lib().tfs_rpy_upload(self.h, what, L.ptr(src)))

    def download(self, what: int, dst):
        L.check(L.lib().tfs_rpy_download(self.h, what, L.ptr(dst)))

    def evaluate(self, blocks: int = L.BLOCK_ALL):
        L.check(L.lib().tfs_rpy_evaluate(self.h, blocks))

    def apply(self, blocks, f_segs, f_blobs, v_segs, v_blobs):
        L.check(L.lib().tfs_rpy_apply(self.h, blocks, L.ptr(f_segs), L.ptr(f_blobs), L.ptr(v_segs),
                                      L.ptr(v_blobs)))

    def sync(self):
        L.check(L.lib().tfs_rpy_sync(self.h))

    def set_precision(self, precision: int):
        L.check(L.lib().tfs_rpy_set_precision(self.h, precision))

    def set_domain(self, domain: int):
        """DOMAIN_UNBOUNDED or DOMAIN_HALF_SPACE (no-slip wall at z = 0)."""
        L.check(L.lib().tfs_rpy_set_domain(self.h, domain))

    def barrier_forces(self, f_segs, f_blobs_repulsion=None, *, strength: float, seg_per_fil: int,
                       wall: bool = False, dl: float = 2.2):
        """Adds the barrier forces of the current positions to f_segs (slots
        0-2) and writes the blobs' into f_blobs_repulsion (arrays in the
        handle's memory space)."""
        d = L.BarrierDesc(float(strength), int(seg_per_fil), int(bool(wall)), float(dl))
        L.check(L.lib().tfs_rpy_barrier_forces(self.h, ctypes.byref(d), L.ptr(f_segs), L.ptr(f_blobs_repulsion)))

    def set_tile_shape(self, precision: int, block: int, per_thread: int, chunk: int):
        L.check(L.lib().tfs_rpy_set_tile_shape(self.h, precision, block, per_thread, chunk))

    def profile(self, enable: bool = True):
        L.check(L.lib().tfs_rpy_profile(self.h, int(enable)))

    def profile_read(self):
        """{precision: (summed pair-kernel ms, launches)} since the last read."""
        return profile_read(self.h)

    def set_bodies(self, segs_per_body, blobs_per_body):
        """Independent bodies stored one after another (block-diagonal mobility)."""
        s = np.ascontiguousarray(segs_per_body, dtype=np.int32)
        b = np.ascontiguousarray(blobs_per_body, dtype=np.int32)
        L.check(L.lib().tfs_rpy_set_bodies(self.h, len(s), s.ctypes.data, b.ctypes.data))
        self.num_bodies = len(s)

    def set_persistent(self, precision: int, workgroups_per_cu: int):
        L.check(L.lib().tfs_rpy_set_persistent(self.h, precision, workgroups_per_cu))

    def set_split_shape(self, precision: int, block: int, per_thread: int, chunk: int):
        """Launch shape whose waves share one tile of targets and split its sources."""
        L.check(L.lib().tfs_rpy_set_split_shape(self.h, precision, block, per_thread, chunk))

    def set_symmetric(self, precision: int, enable: bool = True):
        """Each unordered pair evaluated once for both particles (3-DOF segments)."""
        L.check(L.lib().tfs_rpy_set_symmetric(self.h, precision, int(enable)))

    def symmetric(self, precision: int) -> bool:
        out = ctypes.c_int32()
        L.check(L.lib().tfs_rpy_get_symmetric(self.h, precision, ctypes.byref(out)))
        return bool(out.value)

    def tile_shape(self, precision: int):
        out = (ctypes.c_int32 * 4)()
        L.check(L.lib().tfs_rpy_get_tile_shape(self.h, precision, out))
        split = ctypes.c_int32()
        L.check(L.lib().tfs_rpy_get_source_split(self.h, precision, ctypes.byref(split)))
        return {"block": out[0], "per_thread": out[1], "chunk": out[2], "persistent": out[3],
                "split": split.value}

    def body_timing(self, enable: bool = True):
        L.check(L.lib().tfs_rpy_body_timing(self.h, int(enable)))

    def body_times(self, num_bodies: int):
        out = (ctypes.c_double * num_bodies)()
        L.check(L.lib().tfs_rpy_body_times(self.h, out))
        return np.array(out[:])

    def autotune(self, precision: int, blocks: int = L.BLOCK_ALL, reps: int = 3) -> float:
        best = ctypes.c_double()
        L.check(L.lib().tfs_rpy_autotune(self.h, precision, blocks, reps, ctypes.byref(best)))
        return best.value

    def graph(self, blocks: int = L.BLOCK_ALL, evaluations: int = 1) -> "Graph":
        return Graph(self, blocks, evaluations)

    def device_buffer(self, what: int) -> int:
        p = ctypes.c_void_p()
        L.check(L.lib().tfs_rpy_device_buffer(self.h, what, ctypes.byref(p)))
        return p.value


class Graph:
    """A captured sequence of evaluations, replayed on the handle's stream."""

    def __init__(self, handle: RpyHandle, blocks: int, evaluations: int):
        g = ctypes.c_void_p()
        L.check(L.lib().tfs_rpy_graph_create(handle.h, blocks, evaluations, ctypes.byref(g)))
        self.g = g

    def launch(self):
        L.check(L.lib().tfs_graph_launch(self.g))

    def close(self):
        if self.g:
            L.lib().tfs_graph_destroy(self.g)
            self.g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RpyMobilitySolver:
    """rpy_mobility_solver over the gfx950 library (one GPU, no target split)."""

    def __init__(self, num_segs: int, num_blobs: int, *, seg_dof: int = 6, precision: str = "fp64",
                 bs_accumulate: bool | None = None, seg_radius: float = 1.0, blob_radius: float = 1.0,
                 mu: float = 1.0):
        self.NS = num_segs
        self.NB = num_blobs
        self.seg_dof = seg_dof
        # The reference accumulates in Mbs exactly in prescribed-cilia builds
        # (seg_dof 3) or Broyden-for-everything builds (cuda_functions.cu:739-751).
        self.bs_accumulate = (seg_dof == 3) if bs_accumulate is None else bs_accumulate
        self.precision = {"fp64": L.PRECISION_FP64, "fp32": L.PRECISION_FP32_ACC64}[precision]
        self.radii = (seg_radius, blob_radius, mu)
        self._host = []
        self.handle = None

    # -- allocation (rpy_mobility_solver.cu:14-117) --------------------------
    def allocate_host_memory(self):
        def pinned(n):
            a = PinnedArray(n)
            self._host.append(a)
            return a.array
        self.v_segs_host = pinned(6 * self.NS)
        self.v_blobs_host = pinned(3 * self.NB)
        self.x_segs_host = pinned(3 * self.NS)
        self.x_blobs_host = pinned(3 * self.NB)
        self.f_segs_host = pinned(6 * self.NS)
        self.f_blobs_host = pinned(3 * self.NB)
        self.f_blobs_repulsion_host = pinned(3 * self.NB)

    def allocate_device_memory(self):
        s, b, mu = self.radii
        self.handle = RpyHandle(self.NS, self.NB, seg_dof=self.seg_dof, precision=self.precision,
                                bs_accumulate=self.bs_accumulate, seg_radius=s, blob_radius=b, mu=mu)

    def free_host_memory(self):
        for a in self._host:
            a.free()
        self._host = []

    def free_device_memory(self):
        if self.handle is not None:
            self.handle.close()
            self.handle = None

    def initialise(self):
        """mobility_solver::initialise (mobility_solver.cpp:28-40): allocate, zero blob forces."""
        self.allocate_host_memory()
        self.allocate_device_memory()
        self.f_blobs_host[:] = 0.0

    def finalise(self):
        self.free_host_memory()
        self.free_device_memory()

    # -- copies (rpy_mobility_solver.cu:119-205) -------------------------------
    def copy_segment_positions_to_device(self):
        self.handle.upload(L.SEG_POSITIONS, self.x_segs_host)

    def copy_segment_forces_to_device(self):
        self.handle.upload(L.SEG_FORCES, self.f_segs_host)

    def copy_blob_positions_to_device(self):
        self.handle.upload(L.BLOB_POSITIONS, self.x_blobs_host)

    def copy_blob_forces_to_device(self):
        self.handle.upload(L.BLOB_FORCES, self.f_blobs_host)

    def copy_segment_velocities_to_host(self):
        self.handle.download(L.SEG_VELOCITIES, self.v_segs_host)

    def copy_blob_velocities_to_host(self):
        self.handle.download(L.BLOB_VELOCITIES, self.v_blobs_host)

    def set_barrier(self, *, strength: float, seg_per_fil: int, wall: bool = False, dl: float = 2.2):
        """Barrier-force parameters of a non-prescribed build (REPULSIVE_FORCE_FACTOR *
        END_FORCE_MAGNITUDE, NSEG, INFINITE_PLANE_WALL, DL); without them
        apply_interparticle_forces adds nothing, as in a prescribed-cilia build."""
        self.barrier = dict(strength=strength, seg_per_fil=seg_per_fil, wall=wall, dl=dl)

    def copy_interparticle_blob_forces_to_host(self):
        rep = getattr(self, "_blob_repulsion", None)
        self.f_blobs_repulsion_host[:] = 0.0 if rep is None else rep

    def apply_interparticle_forces(self):
        """rpy_mobility_solver.cu:207-232: the device segment forces gain the
        barrier forces (the host array keeps the model's), and the blob
        repulsion is computed for copy_interparticle_blob_forces_to_host."""
        nb = len(self.f_blobs_repulsion_host)
        self._blob_repulsion = np.zeros(nb)
        if getattr(self, "barrier", None) is None:
            return
        staged = np.array(self.f_segs_host, dtype=np.float64)
        self.handle.barrier_forces(staged, self._blob_repulsion if nb else None, **self.barrier)

    def wait_for_device(self):
        self.handle.sync()

    # -- evaluation (rpy_mobility_solver.cu:258-337) ----------------------------
    def evaluate_segment_segment_mobility(self):
        self.handle.evaluate(L.BLOCK_SS)

    def evaluate_segment_blob_mobility(self):
        self.handle.evaluate(L.BLOCK_SB)

    def evaluate_blob_blob_mobility(self):
        self.handle.evaluate(L.BLOCK_BB)

    def evaluate_blob_segment_mobility(self):
        self.handle.evaluate(L.BLOCK_BS)

    def evaluate_full_mobility(self):
        """Mss(=), Msb(+=), D2H v_segs; Mbb(=), Mbs(+=|=), D2H v_blobs."""
        self.handle.evaluate(L.BLOCK_ALL)
        self.copy_segment_velocities_to_host()
        self.copy_blob_velocities_to_host()


def matvec(x_segs, x_blobs, f_segs, f_blobs, *, seg_dof=6, precision="fp64", bs_accumulate=False,
           blocks=L.BLOCK_ALL, v_segs=None, v_blobs=None, seg_radius=1.0, blob_radius=1.0, mu=1.0,
           tile_shape=None, domain=L.DOMAIN_UNBOUNDED):
    """One-shot matvec on host numpy arrays: returns (v_segs[6*Ns], v_blobs[3*Nb])."""
    xs = np.ascontiguousarray(x_segs, dtype=np.float64).reshape(-1)
    xb = np.ascontiguousarray(x_blobs, dtype=np.float64).reshape(-1)
    ns, nb = xs.size // 3, xb.size // 3
    fs = np.ascontiguousarray(f_segs, dtype=np.float64).reshape(-1)
    fb = np.ascontiguousarray(f_blobs, dtype=np.float64).reshape(-1)
    vs = np.zeros(6 * ns) if v_segs is None else np.array(v_segs, dtype=np.float64).reshape(-1)
    vb = np.zeros(3 * nb) if v_blobs is None else np.array(v_blobs, dtype=np.float64).reshape(-1)
    prec = {"fp64": L.PRECISION_FP64, "fp32": L.PRECISION_FP32_ACC64}[precision]
    h = RpyHandle(ns, nb, seg_dof=seg_dof, precision=prec, bs_accumulate=bs_accumulate,
                  seg_radius=seg_radius, blob_radius=blob_radius, mu=mu)
    try:
        if tile_shape is not None:
            h.set_tile_shape(prec, *tile_shape)
        if domain != L.DOMAIN_UNBOUNDED:
            h.set_domain(domain)
        if ns:
            h.upload(L.SEG_POSITIONS, xs)
        if nb:
            h.upload(L.BLOB_POSITIONS, xb)
        h.apply(blocks, fs if ns else None, fb if nb else None, vs if ns else None, vb if nb else None)
        h.sync()
    finally:
        h.close()
    return vs, vb
