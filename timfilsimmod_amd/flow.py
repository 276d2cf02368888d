"""Flow-field evaluator over the C-ABI of include/tfs_flow.h.

``FlowFieldEvaluator`` mirrors flow_field_evaluator
(src/flow_field/flow_field_evaluator.hpp:8-63): ``set_sources`` is
read_positions_and_forces (flow_field_evaluator.cu:401-544) given arrays
instead of save files, ``velocities`` is velocities() (:546-579) at caller
points.  Every evaluation runs in libtfs_rpy.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

STOKESLET_ROTLET = 0
RPY = 1


class FlowDesc(ctypes.Structure):
    _fields_ = [
        ("num_bodies", ctypes.c_int32), ("segs_per_body", ctypes.c_int32),
        ("blobs_per_body", ctypes.c_int32), ("mode", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("remove_net_force", ctypes.c_int32), ("seg_radius", ctypes.c_double),
        ("blob_radius", ctypes.c_double), ("mu", ctypes.c_double), ("memory", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
    ]


_vp = ctypes.c_void_p
L.extra_signatures({
    "tfs_flow_create": (ctypes.c_int, [ctypes.POINTER(FlowDesc), ctypes.POINTER(_vp)]),
    "tfs_flow_destroy": (ctypes.c_int, [_vp]),
    "tfs_flow_set_sources": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tfs_flow_evaluate": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp]),
    "tfs_flow_set_sources_f32": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tfs_flow_evaluate_f32": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp]),
    "tfs_flow_sync": (ctypes.c_int, [_vp]),
    "tfs_flow_set_block": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "tfs_flow_last_kernel_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
})


class FlowFieldEvaluator:
    def __init__(self, num_bodies, segs_per_body, blobs_per_body, *, mode=STOKESLET_ROTLET,
                 precision="fp64", remove_net_force=False, seg_radius=1.0, blob_radius=1.0, mu=1.0,
                 memory=L.MEMORY_HOST, stream=None):
        d = FlowDesc()
        d.num_bodies, d.segs_per_body, d.blobs_per_body = num_bodies, segs_per_body, blobs_per_body
        d.mode = mode
        d.precision = {"fp64": L.PRECISION_FP64, "fp32": L.PRECISION_FP32_ACC64}[precision]
        d.remove_net_force = int(bool(remove_net_force))
        d.seg_radius, d.blob_radius, d.mu = seg_radius, blob_radius, mu
        d.memory, d.stream = memory, stream
        self.desc = d
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_flow_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            L.lib().tfs_flow_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sources(self, x_segs, f_segs, x_blobs, f_blobs, x_bodies=None):
        arrs = [None if a is None else (a if hasattr(a, "data_ptr") else
                                        np.ascontiguousarray(a, dtype=np.float64))
                for a in (x_segs, f_segs, x_blobs, f_blobs, x_bodies)]
        self._keep = arrs
        L.check(L.lib().tfs_flow_set_sources(self.h, *[L.ptr(a) if a is not None and
                                                          (hasattr(a, "data_ptr") or a.size) else None
                                                          for a in arrs]))

    def set_block(self, block: int):
        L.check(L.lib().tfs_flow_set_block(self.h, block))

    def velocities(self, x_points, out=None):
        """Host mode: numpy in, numpy out.  Device mode: torch tensors (async)."""
        if self.desc.memory == L.MEMORY_HOST:
            x = np.ascontiguousarray(x_points, dtype=np.float64).reshape(-1)
            v = np.zeros_like(x) if out is None else out
            L.check(L.lib().tfs_flow_evaluate(self.h, x.size // 3, L.ptr(x), L.ptr(v)))
            return v
        n = x_points.numel() // 3
        L.check(L.lib().tfs_flow_evaluate(self.h, n, L.ptr(x_points), L.ptr(out)))
        return out

    def sync(self):
        L.check(L.lib().tfs_flow_sync(self.h))

    def last_kernel_ms(self) -> float:
        ms = ctypes.c_double()
        L.check(L.lib().tfs_flow_last_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value
