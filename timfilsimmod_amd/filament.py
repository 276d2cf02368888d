"""Cosserat filament model on the device, over include/tfs_filament.h.

``FilamentModel`` holds every filament of every swimmer and mirrors the
filament methods the elastic implicit timestep calls (filament.cpp):
``initial_guess`` (:1083-1385), ``forces`` (internal_forces_and_torques,
:1467-1575), ``jacobian`` (the J0 blocks of invert_approx_jacobian,
:1761-1894, inverted by broyden.BroydenSolver.set_jacobian) and ``update``
(:1896-1930).  All of it runs in libtfs_rpy.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

X, XM1, XM2, U, UM1, Q, MULT = range(7)
BODY_X, BODY_XM1, BODY_XM2, BODY_U, BODY_UM1, BODY_Q = range(7, 13)


class FilamentDesc(ctypes.Structure):
    _fields_ = [
        ("num_swim", ctypes.c_int32), ("fil_per_swim", ctypes.c_int32), ("num_seg", ctypes.c_int32),
        ("dl", ctypes.c_double), ("kb", ctypes.c_double), ("kt", ctypes.c_double), ("mu", ctypes.c_double),
        ("radius", ctypes.c_double), ("dt", ctypes.c_double), ("end_force", ctypes.c_double),
        ("strain_twist", ctypes.c_double * 3), ("memory", ctypes.c_int32), ("stream", ctypes.c_void_p),
    ]


_vp, _i32 = ctypes.c_void_p, ctypes.c_int32
L.extra_signatures({
    "tfs_filament_create": (ctypes.c_int, [ctypes.POINTER(FilamentDesc), ctypes.POINTER(_vp)]),
    "tfs_filament_destroy": (ctypes.c_int, [_vp]),
    "tfs_filament_set_state": (ctypes.c_int, [_vp, _vp, _vp]),
    "tfs_filament_set_perturbation": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "tfs_filament_initial_guess": (ctypes.c_int, [_vp, _i32]),
    "tfs_filament_forces": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_filament_jacobian": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "tfs_filament_update": (ctypes.c_int, [_vp, _vp]),
    "tfs_filament_sync": (ctypes.c_int, [_vp]),
    "tfs_elastic_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_vp)]),
    "tfs_elastic_destroy": (ctypes.c_int, [_vp]),
    "tfs_elastic_step": (ctypes.c_int, [_vp, _i32, ctypes.c_void_p]),
    "tfs_elastic_set_graphs": (ctypes.c_int, [_vp, _i32]),
    "tfs_elastic_iteration_counts": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "tfs_filament_device_array": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp)]),
    "tfs_filament_read": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_filament_set_body": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _vp]),
    "tfs_filament_blob_positions": (ctypes.c_int, [_vp, _vp]),
    "tfs_filament_body_coupling": (ctypes.c_int, [_vp, _i32, _vp, ctypes.c_double, _vp, _vp, _vp]),
    "tfs_elastic_backup_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "tfs_elastic_backup": (ctypes.c_int, [_vp, _vp, _vp]),
    "tfs_elastic_restore": (ctypes.c_int, [_vp, _vp]),
    "tfs_filament_set_base": (ctypes.c_int, [_vp, _i32, _vp]),
    "tfs_filament_body_coupling_blobs": (ctypes.c_int, [_vp, _i32, _vp, ctypes.c_double, _vp, _vp, _vp, _vp]),
})

BASE_CLAMPED = 0   # CILIA_TYPE 0 (INSTABILITY_CILIA), and 2 (CONSTANT_BASE_ROTATION) with values
BASE_DRIVEN = 1    # CILIA_TYPE 1 (GEOMETRIC_CILIA): free base segment with a driving torque


class FilamentModel:
    def __init__(self, num_swim, fil_per_swim, num_seg, *, dl=2.2, kb=1800.0, kt=1800.0, mu=1.0, radius=1.0,
                 dt=1.0, end_force=0.0, strain_twist=(0.0, 0.0, 0.0), memory=L.MEMORY_HOST, stream=None):
        d = FilamentDesc()
        d.num_swim, d.fil_per_swim, d.num_seg = num_swim, fil_per_swim, num_seg
        d.dl, d.kb, d.kt, d.mu, d.radius, d.dt, d.end_force = dl, kb, kt, mu, radius, dt, end_force
        for i in range(3):
            d.strain_twist[i] = strain_twist[i]
        d.memory, d.stream = memory, stream
        self.desc = d
        self.nfil = num_swim * fil_per_swim
        self.ns = self.nfil * num_seg
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_filament_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tfs_filament_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def update_size(self):
        d = self.desc
        return d.num_swim * (d.fil_per_swim * 6 * d.num_seg + 6)

    def block_sizes(self, blobs_per_swim=0):
        """Sizes of the J0 blocks in the Broyden layout (filaments, then the body,
        then, in the Broyden-for-everything form, one 3 x 3 block per blob)."""
        d = self.desc
        return ([6 * d.num_seg] * d.fil_per_swim + [6] + [3] * int(blobs_per_swim)) * d.num_swim

    @staticmethod
    def _arr(a):
        """Device tensors pass through; anything else becomes a float64 host array."""
        return a if hasattr(a, "data_ptr") else np.ascontiguousarray(a, dtype=np.float64)

    def set_state(self, bases, q):
        b, qq = self._arr(bases), self._arr(q)
        L.check(L.lib().tfs_filament_set_state(self.h, L.ptr(b), L.ptr(qq)))

    def set_perturbation(self, seg_a, seg_b, forces):
        self._pert = self._arr(forces)
        L.check(L.lib().tfs_filament_set_perturbation(self.h, int(seg_a), int(seg_b), L.ptr(self._pert)))

    def set_base(self, mode, values=None):
        """tfs_filament_set_base: BASE_CLAMPED with values = per filament the
        base's turn per step [3] (None: held still), or BASE_DRIVEN with values
        = per filament the lab-frame torque on the base segment [3]."""
        self._base = None if values is None else self._arr(values)
        L.check(L.lib().tfs_filament_set_base(self.h, int(mode), L.ptr(self._base) if self._base is not None
                                              else None))

    def set_body(self, x_bodies, q_bodies, fil_refs, blobs_per_swim=0, blob_refs=None):
        """Free bodies (swimmer.cpp:433-467, 1072-1098): each swimmer's body
        starts at x_bodies [3], q_bodies [4] at rest; filament f is anchored at
        X + R(q) fil_refs[f]; blob_refs [3 per blob] for blob_positions."""
        self._body = [self._arr(a) for a in (x_bodies, q_bodies, fil_refs)]
        self._blob_refs = None if blob_refs is None else self._arr(blob_refs)
        self.nblob = self.desc.num_swim * int(blobs_per_swim)
        L.check(L.lib().tfs_filament_set_body(self.h, *[L.ptr(a) for a in self._body], int(blobs_per_swim),
                                              L.ptr(self._blob_refs)))

    def blob_positions(self, out=None):
        out = np.zeros(3 * self.nblob) if out is None else out
        L.check(L.lib().tfs_filament_blob_positions(self.h, L.ptr(out)))
        return out

    def body_coupling(self, nt, body_mobility, sphere_radius, A=None, B=None, S0=None):
        """J0's body coupling (swimmer.cpp:688-880) after jacobian() at the same
        state: per filament A (6N x 6) and B (6 x 6N), per swimmer S0 (6 x 6)."""
        n6, nf, ns = 6 * self.desc.num_seg, self.nfil, self.desc.num_swim
        A = np.zeros(nf * 6 * n6) if A is None else A
        B = np.zeros(nf * 6 * n6) if B is None else B
        S0 = np.zeros(36 * ns) if S0 is None else S0
        nm = self._arr(body_mobility)
        L.check(L.lib().tfs_filament_body_coupling(self.h, int(nt), L.ptr(nm), float(sphere_radius), L.ptr(A),
                                                   L.ptr(B), L.ptr(S0)))
        return A, B, S0

    def body_coupling_blobs(self, nt, f_blobs, blob_radius=1.0, A=None, B=None, S0=None, blob_disp=None):
        """J0's body coupling in the Broyden-for-everything form with free bodies
        (swimmer.cpp:565-676) after jacobian() at the same state: per filament A
        (6N x 6) and B (6 x 6N: the force and torque balance's dependence), per
        swimmer S0 (6 x 6, with the blob terms), the blobs' displacements."""
        n6, nf, ns = 6 * self.desc.num_seg, self.nfil, self.desc.num_swim
        nb3 = np.asarray(f_blobs).size if not hasattr(f_blobs, "data_ptr") else f_blobs.numel()
        A = np.zeros(nf * 6 * n6) if A is None else A
        B = np.zeros(nf * 6 * n6) if B is None else B
        S0 = np.zeros(36 * ns) if S0 is None else S0
        blob_disp = np.zeros(nb3) if blob_disp is None else blob_disp
        fb = self._arr(f_blobs)
        L.check(L.lib().tfs_filament_body_coupling_blobs(self.h, int(nt), L.ptr(fb), float(blob_radius), L.ptr(A),
                                                         L.ptr(B), L.ptr(S0), L.ptr(blob_disp)))
        return A, B, S0, blob_disp

    def initial_guess(self, nt):
        L.check(L.lib().tfs_filament_initial_guess(self.h, int(nt)))

    def forces(self, nt, out=None):
        out = np.zeros(6 * self.ns) if out is None else out
        L.check(L.lib().tfs_filament_forces(self.h, int(nt), L.ptr(out)))
        return out

    def jacobian(self, nt, f_segs, blocks):
        L.check(L.lib().tfs_filament_jacobian(self.h, int(nt), L.ptr(f_segs), L.ptr(blocks)))
        return blocks

    def update(self, du):
        d = self._arr(du)
        L.check(L.lib().tfs_filament_update(self.h, L.ptr(d)))

    def sync(self):
        L.check(L.lib().tfs_filament_sync(self.h))

    def device_array(self, which) -> int:
        p = ctypes.c_void_p()
        L.check(L.lib().tfs_filament_device_array(self.h, which, ctypes.byref(p)))
        return p.value

    def read(self, which):
        n = {X: 3, XM1: 3, XM2: 3, U: 3, UM1: 3, Q: 4}.get(which)
        size = n * self.ns if n else 3 * (self.desc.num_seg + 1) * self.nfil
        if which >= BODY_X:
            size = (4 if which == BODY_Q else 3) * self.desc.num_swim
        out = np.zeros(size)
        L.check(L.lib().tfs_filament_read(self.h, which, L.ptr(out)))
        return out


class ElasticDesc(ctypes.Structure):
    _fields_ = [("filaments", ctypes.c_void_p), ("velocities", ctypes.c_void_p), ("broyden", ctypes.c_void_p),
                ("x_blobs", ctypes.c_void_p), ("x_bodies", ctypes.c_void_p), ("v_bodies", ctypes.c_void_p),
                ("f_blobs", ctypes.c_void_p), ("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("body_mobility", ctypes.c_void_p), ("sphere_radius", ctypes.c_double),
                ("broyden_blobs", ctypes.c_int32), ("blob_radius", ctypes.c_double)]


class ElasticStats(ctypes.Structure):
    _fields_ = [("broyden_iterations", ctypes.c_int32), ("gmres_iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("velocity_solves", ctypes.c_int32), ("error_norm", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ElasticStep:
    """One elastic implicit timestep on the device (tfs_elastic_step): the
    time-loop body of cilia_sim_main.cu:353-502 for bodies held still, or free
    bodies when the filament model has them (set_body; then x_blobs, x_bodies,
    v_bodies may be None and body_mobility is a device tensor of 36 per
    swimmer).  The handles must use device memory; x_blobs, x_bodies,
    v_bodies, f_blobs are device tensors (f_blobs is the blob-force warm start
    and result)."""

    def __init__(self, filaments, velocities, broyden, x_blobs, x_bodies, v_bodies, f_blobs, *, tol=1e-4,
                 max_iter=400, body_mobility=None, sphere_radius=0.0, broyden_blobs=0, blob_radius=1.0):
        """broyden_blobs = NBLOB: the Broyden-for-everything form (blob forces in
        f_blobs are Broyden unknowns; velocities may be None)."""
        self._keep = (filaments, velocities, broyden, x_blobs, x_bodies, v_bodies, f_blobs, body_mobility)
        d = ElasticDesc(filaments.h.value, velocities.h.value if velocities is not None else None, broyden.h.value,
                        L.ptr(x_blobs), L.ptr(x_bodies), L.ptr(v_bodies), L.ptr(f_blobs), float(tol), int(max_iter),
                        L.ptr(body_mobility), float(sphere_radius), int(broyden_blobs), float(blob_radius))
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_elastic_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tfs_elastic_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def backup(self, held_bodies=None):
        """The state in the reference's backup layout (tfs_elastic_backup):
        a host array; held_bodies (x, q per swimmer) for bodies held still."""
        n = ctypes.c_int64()
        L.check(L.lib().tfs_elastic_backup_size(self.h, ctypes.byref(n)))
        out = np.zeros(n.value)
        hb = None if held_bodies is None else np.ascontiguousarray(held_bodies, dtype=np.float64)
        L.check(L.lib().tfs_elastic_backup(self.h, L.ptr(hb), L.ptr(out)))
        return out

    def restore(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        n = ctypes.c_int64()
        L.check(L.lib().tfs_elastic_backup_size(self.h, ctypes.byref(n)))
        if v.size != n.value:
            raise ValueError(f"backup holds {v.size} values, this step needs {n.value}")
        L.check(L.lib().tfs_elastic_restore(self.h, L.ptr(v)))

    def set_graphs(self, enable: bool):
        """Captured Broyden iterations on (default) or off (tfs_elastic_set_graphs)."""
        L.check(L.lib().tfs_elastic_set_graphs(self.h, int(bool(enable))))

    def iteration_counts(self):
        """(iterations replayed as graphs, iterations launched one by one)."""
        g, e = ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().tfs_elastic_iteration_counts(self.h, ctypes.byref(g), ctypes.byref(e)))
        return g.value, e.value

    def step(self, nt, *, allow_unconverged=False):
        st = ElasticStats()
        rc = L.lib().tfs_elastic_step(self.h, int(nt), ctypes.byref(st))
        if not (allow_unconverged and rc == L.TFS_ERR_NOT_CONVERGED):
            L.check(rc)
        return st.as_dict()
