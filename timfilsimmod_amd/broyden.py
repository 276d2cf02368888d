"""Device-resident Broyden solver over the C-ABI of include/tfs_broyden.h.

``BroydenSolver`` mirrors broyden_solver (src/cilia/broyden_solver.hpp:18-43):
``find_update`` and ``end_of_iter`` follow broyden_solver.cpp:26-67 and :69-129,
``compute_errors`` is mobility_solver::compute_errors
(src/cilia/mobility/mobility_solver.cpp:2661-2946) evaluated on the device.
The approximate Jacobian (or its inverse) is given as block-diagonal dense
blocks, e.g. one 6*NSEG block per filament and a 6x6 block per body.
Every operation runs in libtfs_rpy.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class BroydenDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("max_iter", ctypes.c_int32), ("confidence", ctypes.c_double),
        ("num_blocks", ctypes.c_int32), ("block_sizes", ctypes.POINTER(ctypes.c_int32)),
        ("memory", ctypes.c_int32), ("stream", ctypes.c_void_p),
    ]


class BroydenState(ctypes.Structure):
    _fields_ = [
        ("num_swim", ctypes.c_int32), ("num_fil", ctypes.c_int32), ("num_seg", ctypes.c_int32),
        ("bdf2", ctypes.c_int32), ("dt", ctypes.c_double), ("dl", ctypes.c_double), ("tol", ctypes.c_double),
        ("seg_x", ctypes.c_void_p), ("seg_xm1", ctypes.c_void_p), ("seg_xm2", ctypes.c_void_p),
        ("seg_u", ctypes.c_void_p), ("seg_um1", ctypes.c_void_p), ("v_segs", ctypes.c_void_p),
        ("body_x", ctypes.c_void_p), ("body_xm1", ctypes.c_void_p), ("body_xm2", ctypes.c_void_p),
        ("body_u", ctypes.c_void_p), ("body_um1", ctypes.c_void_p), ("v_bodies", ctypes.c_void_p),
        ("blobs_per_swim", ctypes.c_int32), ("x_blobs", ctypes.c_void_p), ("v_blobs", ctypes.c_void_p),
        ("free_bodies", ctypes.c_int32), ("f_segs", ctypes.c_void_p), ("f_blobs", ctypes.c_void_p),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
L.extra_signatures({
    "tfs_broyden_create": (ctypes.c_int, [ctypes.POINTER(BroydenDesc), ctypes.POINTER(_vp)]),
    "tfs_broyden_destroy": (ctypes.c_int, [_vp]),
    "tfs_broyden_set_jacobian_inverse": (ctypes.c_int, [_vp, _vp]),
    "tfs_broyden_set_jacobian": (ctypes.c_int, [_vp, _vp]),
    "tfs_broyden_set_body_coupling": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp]),
    "tfs_broyden_set_body_coupling_blobs": (ctypes.c_int, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp,
                                                           ctypes.c_double, ctypes.c_double]),
    "tfs_broyden_start": (ctypes.c_int, [_vp, _vp]),
    "tfs_broyden_find_update": (ctypes.c_int, [_vp, _vp]),
    "tfs_broyden_end_of_iter": (ctypes.c_int, [_vp, _vp, _i32]),
    "tfs_broyden_status": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_double)]),
    "tfs_broyden_compute_errors": (ctypes.c_int, [_vp, ctypes.POINTER(BroydenState), _i32,
                                                  ctypes.POINTER(_i32)]),
    "tfs_broyden_read_error": (ctypes.c_int, [_vp, _i32, _vp]),
})

# JACOBIAN_CONFIDENCE_FACTOR, MAX_BROYDEN_ITER, TOL, NUM_EULER_STEPS (config.hpp:299-302, 326)
JACOBIAN_CONFIDENCE_FACTOR = 0.1
MAX_BROYDEN_ITER = 400
TOL = 1e-4
NUM_EULER_STEPS = 1


def _f64(a):
    return None if a is None else (a if hasattr(a, "data_ptr") else np.ascontiguousarray(a, dtype=np.float64))


class BroydenSolver:
    """Limited-memory Broyden engine; vectors live in ``memory`` (host numpy
    arrays or device torch tensors)."""

    def __init__(self, block_sizes, *, max_iter=MAX_BROYDEN_ITER, confidence=JACOBIAN_CONFIDENCE_FACTOR,
                 memory=L.MEMORY_HOST, stream=None):
        sizes = np.ascontiguousarray(block_sizes, dtype=np.int32)
        self.n = int(sizes.sum())
        d = BroydenDesc()
        d.n, d.max_iter, d.confidence = self.n, max_iter, confidence
        d.num_blocks = len(sizes)
        d.block_sizes = sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.memory, d.stream = memory, stream
        self.memory = memory
        self.block_sizes = sizes
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_broyden_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tfs_broyden_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _out(self, n):
        if self.memory == L.MEMORY_HOST:
            return np.empty(n, dtype=np.float64)
        import torch
        return torch.empty(n, dtype=torch.float64, device="cuda")

    def set_jacobian_inverse(self, blocks):
        """J0^{-1} as a list of dense blocks (or their row-major concatenation)."""
        flat = _concat(blocks, self.memory)
        self._jac = flat   # alive until the stream-ordered copy has run
        L.check(L.lib().tfs_broyden_set_jacobian_inverse(self.h, L.ptr(flat)))

    def set_jacobian(self, blocks):
        """J0 as dense blocks; inverted on the device."""
        flat = _concat(blocks, self.memory)
        L.check(L.lib().tfs_broyden_set_jacobian(self.h, L.ptr(flat)))

    def set_body_coupling(self, num_swim, fil_per_swim, A, B, S0):
        """Free bodies: J0^{-1} through each body's Schur complement
        (swimmer::jacobian_inv_mult, swimmer.cpp:1012-1066), after set_jacobian;
        A, B, S0 as FilamentModel.body_coupling writes them."""
        a, b, s = _f64(A), _f64(B), _f64(S0)
        L.check(L.lib().tfs_broyden_set_body_coupling(self.h, int(num_swim), int(fil_per_swim), L.ptr(a), L.ptr(b),
                                                      L.ptr(s)))

    def set_body_coupling_blobs(self, num_swim, fil_per_swim, blobs_per_swim, A, B, S0, blob_disp, blob_mob_fac,
                                dt_fac):
        """Free bodies in the Broyden-for-everything form (swimmer.cpp:931-1010),
        after set_jacobian; A, B, S0, blob_disp as
        FilamentModel.body_coupling_blobs writes them."""
        a, b, s, d = _f64(A), _f64(B), _f64(S0), _f64(blob_disp)
        L.check(L.lib().tfs_broyden_set_body_coupling_blobs(self.h, int(num_swim), int(fil_per_swim),
                                                            int(blobs_per_swim), L.ptr(a), L.ptr(b), L.ptr(s),
                                                            L.ptr(d), float(blob_mob_fac), float(dt_fac)))

    def start(self, error=None):
        e = _f64(error)
        L.check(L.lib().tfs_broyden_start(self.h, L.ptr(e)))

    def find_update(self):
        out = self._out(self.n)
        L.check(L.lib().tfs_broyden_find_update(self.h, L.ptr(out)))
        return out

    def end_of_iter(self, new_error, error_is_too_large):
        e = _f64(new_error)
        L.check(L.lib().tfs_broyden_end_of_iter(self.h, L.ptr(e), int(bool(error_is_too_large))))

    @property
    def iter(self) -> int:
        it = ctypes.c_int32()
        L.check(L.lib().tfs_broyden_status(self.h, ctypes.byref(it), None))
        return it.value

    def error_norm(self) -> float:
        v = ctypes.c_double()
        L.check(L.lib().tfs_broyden_status(self.h, None, ctypes.byref(v)))
        return v.value

    def compute_errors(self, state: dict, slot: int, *, num_swim, num_fil, num_seg, bdf2, dt, dl, tol=TOL,
                       blobs_per_swim=0, free_bodies=False):
        """Residuals into slot 0 (error) or 1 (new_error); returns error_is_too_large.
        ``state`` maps the tfs_broyden_state array names to arrays; blobs_per_swim > 0
        is the Broyden-for-everything form (x_blobs, v_blobs in ``state``), with
        free_bodies its free-body rows (f_segs, f_blobs in ``state``)."""
        st = BroydenState()
        st.num_swim, st.num_fil, st.num_seg, st.bdf2 = num_swim, num_fil, num_seg, int(bool(bdf2))
        st.dt, st.dl, st.tol = dt, dl, tol
        st.blobs_per_swim = int(blobs_per_swim)
        st.free_bodies = int(bool(free_bodies))
        keep = []
        for name, _ in BroydenState._fields_[7:]:
            if name in ("blobs_per_swim", "free_bodies"):
                continue
            a = _f64(state.get(name))
            keep.append(a)
            setattr(st, name, L.ptr(a) if a is not None and (hasattr(a, "data_ptr") or a.size) else None)
        flag = ctypes.c_int32()
        L.check(L.lib().tfs_broyden_compute_errors(self.h, ctypes.byref(st), slot, ctypes.byref(flag)))
        return bool(flag.value)

    def read_error(self, slot=0):
        out = self._out(self.n)
        L.check(L.lib().tfs_broyden_read_error(self.h, slot, L.ptr(out)))
        return out


def _concat(blocks, memory):
    if hasattr(blocks, "data_ptr"):
        return blocks.contiguous()
    if isinstance(blocks, np.ndarray) and blocks.ndim == 1:
        return np.ascontiguousarray(blocks, dtype=np.float64)
    flat = np.concatenate([np.ascontiguousarray(b, dtype=np.float64).reshape(-1) for b in blocks])
    if memory == L.MEMORY_DEVICE:
        import torch
        return torch.from_numpy(flat).cuda()
    return flat
