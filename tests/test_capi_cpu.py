"""The C-ABI library loads without a GPU, exports every entry point that
include/*.h declares, and rejects bad descriptors / reports a missing device
with the documented status codes (no compute calls here)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from timfilsimmod_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^[A-Za-z_][\w \*]*?\b(tfs_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "tfs_rpy_create" in names and "tfs_rpy_apply" in names
    assert len(names) >= 14


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (tfs_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_ctypes_table_matches_headers():
    # Every declared function is bound (submodules register theirs on import).
    import timfilsimmod_amd  # noqa: F401
    assert set(declared_functions()) <= set(L.SIGNATURES)
    L.lib()


def _desc(**kw):
    d = L.RpyDesc()
    d.num_segs, d.num_blobs, d.seg_dof, d.precision, d.bs_mode, d.memory = 10, 10, 6, 0, 0, 0
    d.seg_radius = d.blob_radius = d.mu = 1.0
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("field,value", [("seg_dof", 4), ("precision", 7), ("bs_mode", 3),
                                         ("memory", 9), ("num_segs", -1), ("mu", 0.0),
                                         ("seg_radius", -1.0)])
def test_create_rejects_bad_descriptor(field, value):
    h = ctypes.c_void_p()
    rc = L.lib().tfs_rpy_create(ctypes.byref(_desc(**{field: value})), ctypes.byref(h))
    assert rc == L.TFS_ERR_INVALID_ARGUMENT
    assert L.lib().tfs_last_error()
    assert not h.value


def test_null_handles_are_errors_not_crashes():
    lib = L.lib()
    assert lib.tfs_rpy_sync(None) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_rpy_evaluate(None, 15) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_rpy_destroy(None) == L.TFS_OK
    assert lib.tfs_beat_set_warm_start(None, 1) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_solver_set_preconditioner(None, 1) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_rpy_set_symmetric(None, 1, 1) == L.TFS_ERR_INVALID_ARGUMENT
    out = ctypes.c_int32(7)
    assert lib.tfs_rpy_get_symmetric(None, 1, ctypes.byref(out)) == L.TFS_ERR_INVALID_ARGUMENT
    assert out.value == 7


def test_create_reports_missing_device_on_cpu_host():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = L.lib().tfs_rpy_create(ctypes.byref(_desc()), ctypes.byref(h))
    assert rc == L.TFS_ERR_NO_DEVICE


def test_broyden_and_velocities_reject_bad_descriptors():
    import numpy as np
    from timfilsimmod_amd.broyden import BroydenDesc
    from timfilsimmod_amd.solver import VelocitiesDesc
    lib = L.lib()
    sizes = np.array([4, 5], dtype=np.int32)
    d = BroydenDesc()
    d.n, d.max_iter, d.confidence, d.num_blocks = 10, 5, 0.1, 2   # sizes sum to 9, not 10
    d.block_sizes = sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    h = ctypes.c_void_p()
    assert lib.tfs_broyden_create(ctypes.byref(d), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    d.n, d.confidence = 9, 0.0
    assert lib.tfs_broyden_create(ctypes.byref(d), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    assert not h.value
    assert lib.tfs_broyden_destroy(None) == L.TFS_OK
    assert lib.tfs_broyden_find_update(None, None) == L.TFS_ERR_INVALID_ARGUMENT
    v = VelocitiesDesc()
    v.num_swim, v.segs_per_swim, v.blobs_per_swim = 1, 10, 0   # a body needs blobs
    assert lib.tfs_velocities_create(ctypes.byref(v), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    v.blobs_per_swim, v.mode = 10, 7
    assert lib.tfs_velocities_create(ctypes.byref(v), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    assert not h.value
    assert lib.tfs_velocities_destroy(None) == L.TFS_OK


def test_newer_entry_points_reject_null_handles_and_bad_descriptors():
    import timfilsimmod_amd  # noqa: F401  (registers every signature)
    from timfilsimmod_amd.filament import FilamentDesc
    lib = L.lib()
    assert lib.tfs_rpy_barrier_forces(None, None, None, None) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_solver_set_bodies(None, None, 0) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_beat_body(None, None, None, None) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_beat_restore(None, None, None, None) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_beat_velocities(None, None, None) == L.TFS_ERR_INVALID_ARGUMENT
    assert lib.tfs_velocities_set_preconditioner(None, 2) == L.TFS_ERR_INVALID_ARGUMENT
    for fn, args in (("tfs_filament_initial_guess", (None, 0)), ("tfs_filament_forces", (None, 0, None)),
                     ("tfs_filament_jacobian", (None, 0, None, None)), ("tfs_filament_update", (None, None)),
                     ("tfs_filament_sync", (None,)), ("tfs_filament_read", (None, 0, None)),
                     ("tfs_elastic_step", (None, 0, None)), ("tfs_io_read_record", (None, 1, 1, None, None)),
                     ("tfs_io_peek_count", (None, None)), ("tfs_io_rewind", (None,)),
                     ("tfs_filament_set_body", (None, None, None, None, 0, None)),
                     ("tfs_filament_blob_positions", (None, None)),
                     ("tfs_filament_body_coupling", (None, 0, None, 0.0, None, None, None)),
                     ("tfs_broyden_set_body_coupling", (None, 1, 1, None, None, None)),
                     ("tfs_velocities_compute_free", (None,) * 10 + (None,)),
                     ("tfs_velocities_set_body_frame", (None, None)),
                     ("tfs_velocities_set_blob_motion", (None, 0)),
                     ("tfs_elastic_backup_size", (None, None)), ("tfs_elastic_backup", (None, None, None)),
                     ("tfs_elastic_restore", (None, None))):
        assert getattr(lib, fn)(*args) == L.TFS_ERR_INVALID_ARGUMENT, fn
    for fn in ("tfs_filament_destroy", "tfs_elastic_destroy", "tfs_io_close"):
        assert getattr(lib, fn)(None) == L.TFS_OK, fn
    h = ctypes.c_void_p()
    d = FilamentDesc()
    d.num_swim, d.fil_per_swim, d.num_seg = 1, 2, 1          # a filament needs two segments
    d.dl = d.mu = d.radius = d.dt = 1.0
    assert lib.tfs_filament_create(ctypes.byref(d), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    d.num_seg, d.dt = 4, 0.0
    assert lib.tfs_filament_create(ctypes.byref(d), ctypes.byref(h)) == L.TFS_ERR_INVALID_ARGUMENT
    assert not h.value
