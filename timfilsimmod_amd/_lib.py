"""ctypes binding of the C-ABI in include/tfs_rpy.h (libtfs_rpy.so, built in-tree).

There is no fallback: if the HIP library is missing or cannot be loaded, every
entry point raises.  Build it with ``python -m timfilsimmod_amd.build``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtfs_rpy.so")

TFS_OK = 0
TFS_ERR_INVALID_ARGUMENT = 1
TFS_ERR_HIP = 2
TFS_ERR_NON_FINITE = 3
TFS_ERR_NOT_CONVERGED = 4
TFS_ERR_NO_DEVICE = 5
TFS_ERR_OUT_OF_MEMORY = 6
TFS_ERR_IO = 7
TFS_ERR_FORMAT = 8

PRECISION_FP64 = 0
PRECISION_FP32_ACC64 = 1
DOMAIN_UNBOUNDED = 0
DOMAIN_HALF_SPACE = 1
BS_OVERWRITE = 0
BS_ACCUMULATE = 1
BLOCK_SS, BLOCK_SB, BLOCK_BB, BLOCK_BS, BLOCK_ALL = 1, 2, 4, 8, 15
MEMORY_HOST, MEMORY_DEVICE = 0, 1
SEG_POSITIONS, SEG_FORCES, BLOB_POSITIONS, BLOB_FORCES, SEG_VELOCITIES, BLOB_VELOCITIES = range(6)


class TfsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[tfs error {code}] {msg}")
        self.code = code


class RpyDesc(ctypes.Structure):
    _fields_ = [
        ("num_segs", ctypes.c_int32),
        ("num_blobs", ctypes.c_int32),
        ("seg_dof", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("bs_mode", ctypes.c_int32),
        ("memory", ctypes.c_int32),
        ("seg_radius", ctypes.c_double),
        ("blob_radius", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("stream", ctypes.c_void_p),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32

class BarrierDesc(ctypes.Structure):
    _fields_ = [("strength", ctypes.c_double), ("seg_per_fil", ctypes.c_int32), ("wall", ctypes.c_int32),
                ("dl", ctypes.c_double)]


# name -> (restype, argtypes); the exported surface of include/*.h.
SIGNATURES = {
    "tfs_rpy_create": (ctypes.c_int, [ctypes.POINTER(RpyDesc), ctypes.POINTER(_vp)]),
    "tfs_rpy_destroy": (ctypes.c_int, [_vp]),
    "tfs_rpy_upload": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_rpy_download": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_rpy_evaluate": (ctypes.c_int, [_vp, _u32]),
    "tfs_rpy_apply": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "tfs_rpy_sync": (ctypes.c_int, [_vp]),
    "tfs_rpy_set_precision": (ctypes.c_int, [_vp, _i32]),
    "tfs_rpy_set_domain": (ctypes.c_int, [_vp, _i32]),
    "tfs_rpy_barrier_forces": (ctypes.c_int, [_vp, ctypes.POINTER(BarrierDesc), _vp, _vp]),
    "tfs_rpy_set_tile_shape": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32]),
    "tfs_rpy_device_buffer": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp)]),
    "tfs_rpy_profile": (ctypes.c_int, [_vp, _i32]),
    "tfs_rpy_profile_read": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i32)]),
    "tfs_solver_rpy": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "tfs_rpy_positions_changed": (ctypes.c_int, [_vp]),
    "tfs_rpy_set_persistent": (ctypes.c_int, [_vp, _i32, _i32]),
    "tfs_rpy_get_tile_shape": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32)]),
    "tfs_rpy_set_split_shape": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32]),
    "tfs_rpy_get_source_split": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32)]),
    "tfs_rpy_set_symmetric": (ctypes.c_int, [_vp, _i32, _i32]),
    "tfs_rpy_get_symmetric": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32)]),
    "tfs_rpy_body_timing": (ctypes.c_int, [_vp, _i32]),
    "tfs_rpy_body_times": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "tfs_rpy_graph_create": (ctypes.c_int, [_vp, _u32, _i32, ctypes.POINTER(_vp)]),
    "tfs_graph_launch": (ctypes.c_int, [_vp]),
    "tfs_graph_destroy": (ctypes.c_int, [_vp]),
    "tfs_rpy_autotune": (ctypes.c_int, [_vp, _i32, _u32, _i32, ctypes.POINTER(ctypes.c_double)]),
    "tfs_rpy_set_bodies": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "tfs_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "tfs_host_free": (ctypes.c_int, [_vp]),
    "tfs_rpy_upload_f32": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_rpy_download_f32": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_rpy_apply_f32": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "tfs_rpy_barrier_forces_f32": (ctypes.c_int, [_vp, ctypes.POINTER(BarrierDesc), _vp, _vp]),
    "tfs_rpy_apply_noalloc": (ctypes.c_int, [ctypes.POINTER(RpyDesc), _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tfs_last_error": (ctypes.c_char_p, []),
    "tfs_version": (ctypes.c_char_p, []),
}

_lib = None


def extra_signatures(table: dict) -> None:
    """Register further entry points (used by the solver and flow modules)."""
    SIGNATURES.update(table)
    if _lib is not None:
        _declare(_lib, table)


def _declare(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _share_hip_runtime_with_torch():
    """Load torch (when installed) before the library.

    torch ships its own libamdhip64 (soname libamdhip64.so.7, the soname the
    library links against).  Loaded first, it is the HIP runtime the library
    binds to; loaded after the library, torch would bring up a second HIP
    runtime in the process, which then sees no GPU.  One runtime per process
    is also what lets torch tensors and streams be passed to the library."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the gfx950 library first "
                "(python -m timfilsimmod_amd.build)")
        _share_hip_runtime_with_torch()
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle, SIGNATURES)
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != TFS_OK:
        msg = lib().tfs_last_error()
        raise TfsError(rc, msg.decode() if msg else "")


def ptr(a) -> int | None:
    """Address of a numpy array (host) or torch tensor (device); None passes NULL."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
