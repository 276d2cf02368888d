"""The per-timestep solve over the C-ABI of include/tfs_solver.h.

``Solver`` mirrors the generic layer of the reference's ``mobility_solver``
for prescribed cilia (src/cilia/mobility/mobility_solver.cpp):
``system_matrix_mult`` (system_matrix_mult_new, :1806-2090),
``apply_preconditioner`` (apply_preconditioner_new, :1040-1513) and
``solve_linear_system`` (:2092-2570), on device vectors.  ``BeatSimulation``
is the time loop of a prescribed-beat body (cilia_sim_main.cu:355-390,
filament::initial_guess, assemble_rhs, compute_velocities).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import beat as B

SOLVE_FP64 = 0
SOLVE_MIXED = 1
PRECOND_REFERENCE = 0   # D = I/(6 pi mu a), the reference's preconditioner
PRECOND_BLOCKS = 1      # D = filament self-mobility blocks + blob patches
BLOB_DIRECT = 2         # tfs_velocities: dense inverse of the blob mobility


class SolverDesc(ctypes.Structure):
    _fields_ = [
        ("num_fil", ctypes.c_int32), ("num_seg", ctypes.c_int32), ("num_blobs", ctypes.c_int32),
        ("dynamic_phase", ctypes.c_int32), ("dynamic_angle", ctypes.c_int32), ("mode", ctypes.c_int32),
        ("seg_radius", ctypes.c_double), ("blob_radius", ctypes.c_double), ("mu", ctypes.c_double),
        ("restart", ctypes.c_int32), ("max_cycles", ctypes.c_int32), ("tol", ctypes.c_double),
        ("inner_tol", ctypes.c_double), ("max_refinements", ctypes.c_int32), ("stream", ctypes.c_void_p),
        ("num_swim", ctypes.c_int32), ("free_bodies", ctypes.c_int32),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("cycles", ctypes.c_int32), ("refinements", ctypes.c_int32),
        ("fp64_matvecs", ctypes.c_int32), ("fp32_matvecs", ctypes.c_int32), ("converged", ctypes.c_int32),
        ("relres", ctypes.c_double), ("initial_relres", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = ctypes.c_void_p


class BeatDesc(ctypes.Structure):
    _fields_ = [
        ("solver", SolverDesc), ("dt", ctypes.c_double), ("fil_length", ctypes.c_double),
        ("frames", _dp), ("bases", _dp), ("x_blobs", _dp), ("s_values", _dp), ("omega0", _dp),
        ("phase0", _dp), ("angle0", _dp), ("beat_A", _dp), ("beat_B", _dp),
        ("gen_phase_forces", _dp), ("n_gen_phase", ctypes.c_int32),
        ("gen_angle_forces", _dp), ("n_gen_angle", ctypes.c_int32),
        ("gen_force_factor", ctypes.c_double), ("torsion_factor", ctypes.c_double),
        ("body_x", ctypes.c_double * 3), ("body_v", ctypes.c_double * 6),
    ]


_i32 = ctypes.c_int32
L.extra_signatures({
    "tfs_solver_create": (ctypes.c_int, [ctypes.POINTER(SolverDesc), ctypes.POINTER(_dp)]),
    "tfs_solver_destroy": (ctypes.c_int, [_dp]),
    "tfs_solver_size": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_int64)]),
    "tfs_solver_set_geometry": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _i32]),
    "tfs_solver_system_mult": (ctypes.c_int, [_dp, _dp, _dp, _i32, _i32]),
    "tfs_solver_precondition": (ctypes.c_int, [_dp, _dp, _dp, _i32]),
    "tfs_solver_solve": (ctypes.c_int, [_dp, _dp, _dp, _i32, ctypes.POINTER(SolveStats)]),
    "tfs_solver_residual": (ctypes.c_int, [_dp, _dp, _dp, _i32, _i32, ctypes.POINTER(ctypes.c_double)]),
    "tfs_beat_create": (ctypes.c_int, [ctypes.POINTER(BeatDesc), ctypes.POINTER(_dp)]),
    "tfs_beat_destroy": (ctypes.c_int, [_dp]),
    "tfs_beat_step": (ctypes.c_int, [_dp, ctypes.POINTER(SolveStats)]),
    "tfs_beat_state": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp]),
    "tfs_beat_solver": (ctypes.c_int, [_dp, ctypes.POINTER(_dp)]),
    "tfs_beat_set_mode": (ctypes.c_int, [_dp, _i32]),
    "tfs_beat_set_warm_start": (ctypes.c_int, [_dp, _i32]),
    "tfs_beat_restore": (ctypes.c_int, [_dp, _dp, _dp, _dp]),
    "tfs_solver_set_bodies": (ctypes.c_int, [_dp, _dp, _i32]),
    "tfs_beat_body": (ctypes.c_int, [_dp, _dp, _dp, _dp]),
    "tfs_velocities_set_preconditioner": (ctypes.c_int, [_dp, _i32]),
    "tfs_beat_velocities": (ctypes.c_int, [_dp, _dp, _dp]),
    "tfs_solver_set_preconditioner": (ctypes.c_int, [_dp, _i32]),
    "tfs_solver_set_geometry_f32": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _i32]),
    "tfs_solver_set_bodies_f32": (ctypes.c_int, [_dp, _dp, _i32]),
    "tfs_solver_system_mult_f32": (ctypes.c_int, [_dp, _dp, _dp, _i32, _i32]),
    "tfs_solver_precondition_f32": (ctypes.c_int, [_dp, _dp, _dp, _i32]),
    "tfs_solver_solve_f32": (ctypes.c_int, [_dp, _dp, _dp, _i32, ctypes.POINTER(SolveStats)]),
    "tfs_solver_residual_f32": (ctypes.c_int, [_dp, _dp, _dp, _i32, _i32, ctypes.POINTER(ctypes.c_double)]),
    "tfs_solver_set_independent_bodies": (ctypes.c_int, [_dp, _i32, _dp, _dp]),
    "tfs_solver_set_graphs": (ctypes.c_int, [_dp, _i32]),
    "tfs_solver_body_stats": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp]),
    "tfs_solver_cycle_counts": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
})


class VelocitiesDesc(ctypes.Structure):
    _fields_ = [
        ("num_swim", ctypes.c_int32), ("segs_per_swim", ctypes.c_int32), ("blobs_per_swim", ctypes.c_int32),
        ("mode", ctypes.c_int32), ("seg_radius", ctypes.c_double), ("blob_radius", ctypes.c_double),
        ("mu", ctypes.c_double), ("restart", ctypes.c_int32), ("max_cycles", ctypes.c_int32),
        ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double), ("max_refinements", ctypes.c_int32),
        ("memory", ctypes.c_int32), ("stream", ctypes.c_void_p), ("free_bodies", ctypes.c_int32),
    ]


L.extra_signatures({
    "tfs_velocities_compute_free": (ctypes.c_int, [_dp] * 10 + [ctypes.POINTER(SolveStats)]),
    "tfs_velocities_set_body_frame": (ctypes.c_int, [_dp, _dp]),
    "tfs_velocities_set_blob_motion": (ctypes.c_int, [_dp, ctypes.c_int32]),
    "tfs_velocities_create": (ctypes.c_int, [ctypes.POINTER(VelocitiesDesc), ctypes.POINTER(_dp)]),
    "tfs_velocities_destroy": (ctypes.c_int, [_dp]),
    "tfs_velocities_compute": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(SolveStats)]),
    "tfs_velocities_compute_f32": (ctypes.c_int, [_dp] * 8 + [ctypes.POINTER(SolveStats)]),
    "tfs_velocities_compute_free_f32": (ctypes.c_int, [_dp] * 10 + [ctypes.POINTER(SolveStats)]),
})


def solver_desc(num_fil, num_seg, num_blobs, *, dynamic_phase=True, dynamic_angle=True, mode=SOLVE_FP64,
                seg_radius=1.0, blob_radius=1.0, mu=1.0, restart=50, max_cycles=20, tol=1e-8,
                inner_tol=1e-3, max_refinements=10, stream=None, num_swim=1, free_bodies=False) -> SolverDesc:
    d = SolverDesc()
    d.num_fil, d.num_seg, d.num_blobs = num_fil, num_seg, num_blobs
    d.dynamic_phase, d.dynamic_angle, d.mode = int(dynamic_phase), int(dynamic_angle), mode
    d.seg_radius, d.blob_radius, d.mu = seg_radius, blob_radius, mu
    d.restart, d.max_cycles, d.tol = restart, max_cycles, tol
    d.inner_tol, d.max_refinements = inner_tol, max_refinements
    d.stream = stream
    d.num_swim, d.free_bodies = num_swim, int(bool(free_bodies))
    return d


class Solver:
    """Saddle-point system of one prescribed-cilia body on the GPU."""

    def __init__(self, desc: SolverDesc):
        self.desc = desc
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_solver_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        L.check(L.lib().tfs_solver_size(self.h, ctypes.byref(n)))
        self.n = n.value
        self._nbody = 1

    def close(self):
        if self.h:
            L.lib().tfs_solver_destroy(self.h)
            self.h = ctypes.c_void_p()

    def set_preconditioner(self, kind: int):
        """PRECOND_REFERENCE (default) or PRECOND_BLOCKS for later solves."""
        L.check(L.lib().tfs_solver_set_preconditioner(self.h, kind))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bodies(self, x_bodies):
        """Body centres of a free-body system (3 per swimmer)."""
        xb = np.ascontiguousarray(x_bodies, dtype=np.float64)
        L.check(L.lib().tfs_solver_set_bodies(self.h, L.ptr(xb), L.MEMORY_HOST))

    def set_geometry(self, x_segs=None, x_blobs=None, vel_dir_phase=None, vel_dir_angle=None):
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.float64)
             for v in (x_segs, x_blobs, vel_dir_phase, vel_dir_angle)]
        L.check(L.lib().tfs_solver_set_geometry(self.h, *[L.ptr(v) for v in a], L.MEMORY_HOST))

    def system_mult(self, v, precision=L.PRECISION_FP64):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.n)
        L.check(L.lib().tfs_solver_system_mult(self.h, L.ptr(v), L.ptr(out), precision, L.MEMORY_HOST))
        return out

    def precondition(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.n)
        L.check(L.lib().tfs_solver_precondition(self.h, L.ptr(v), L.ptr(out), L.MEMORY_HOST))
        return out

    def solve(self, rhs, x0=None, *, allow_unconverged=False):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        st = SolveStats()
        rc = L.lib().tfs_solver_solve(self.h, L.ptr(rhs), L.ptr(x), L.MEMORY_HOST, ctypes.byref(st))
        if not (allow_unconverged and rc == L.TFS_ERR_NOT_CONVERGED):
            L.check(rc)
        return x, st.as_dict()

    def residual(self, rhs, x, precision=L.PRECISION_FP64):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = ctypes.c_double()
        L.check(L.lib().tfs_solver_residual(self.h, L.ptr(rhs), L.ptr(x), precision, L.MEMORY_HOST,
                                            ctypes.byref(out)))
        return out.value

    def set_independent_bodies(self, fil_per_body, blobs_per_body):
        """Solve the bodies held one after another as independent systems in
        lock-step (tfs_solver_set_independent_bodies)."""
        self._nbody = set_independent_bodies(self.h, fil_per_body, blobs_per_body)

    def set_graphs(self, enable: bool):
        L.check(L.lib().tfs_solver_set_graphs(self.h, int(bool(enable))))

    def body_stats(self):
        return body_stats(self.h, self._nbody)


def set_independent_bodies(handle, fil_per_body, blobs_per_body):
    """tfs_solver_set_independent_bodies on a solver handle."""
    fil = np.ascontiguousarray(fil_per_body, dtype=np.int32)
    blobs = np.ascontiguousarray(blobs_per_body, dtype=np.int32)
    if fil.size != blobs.size:
        raise ValueError("set_independent_bodies: one filament and one blob count per body")
    L.check(L.lib().tfs_solver_set_independent_bodies(handle, fil.size, L.ptr(fil), L.ptr(blobs)))
    return fil.size


def body_stats(handle, nbody: int) -> dict:
    """Per-body results of the last solve (tfs_solver_body_stats)."""
    it, cv = np.zeros(nbody, np.int32), np.zeros(nbody, np.int32)
    rr, ms = np.zeros(nbody), np.zeros(nbody)
    L.check(L.lib().tfs_solver_body_stats(handle, L.ptr(it), L.ptr(cv), L.ptr(rr), L.ptr(ms)))
    gc, ec = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().tfs_solver_cycle_counts(handle, ctypes.byref(gc), ctypes.byref(ec)))
    return {"iterations": it, "converged": cv.astype(bool), "relres": rr, "solve_ms": ms,
            "graph_cycles": gc.value, "eager_cycles": ec.value}


class ElasticVelocities:
    """mobility_solver::compute_velocities for elastic filaments on bodies with
    prescribed velocity (mobility_solver.cpp:2598-2655): segment velocities of
    given segment forces/torques plus the blob forces that keep each body
    surface in rigid motion.  Host arrays; every stage runs on the GPU."""

    def __init__(self, num_swim, segs_per_swim, blobs_per_swim, *, mode=SOLVE_FP64, seg_radius=1.0,
                 blob_radius=1.0, mu=1.0, restart=30, max_cycles=50, tol=1e-10, inner_tol=1e-3,
                 max_refinements=10, stream=None, memory=L.MEMORY_HOST, free_bodies=False):
        d = VelocitiesDesc()
        d.free_bodies = 1 if free_bodies else 0
        d.num_swim, d.segs_per_swim, d.blobs_per_swim, d.mode = num_swim, segs_per_swim, blobs_per_swim, mode
        d.seg_radius, d.blob_radius, d.mu = seg_radius, blob_radius, mu
        d.restart, d.max_cycles, d.tol = restart, max_cycles, tol
        d.inner_tol, d.max_refinements = inner_tol, max_refinements
        d.memory, d.stream = memory, stream
        self.desc = d
        self.ns = num_swim * segs_per_swim
        self.nb = num_swim * blobs_per_swim
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_velocities_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.f_blobs = np.zeros(3 * self.nb)
        self.v_bodies = np.zeros(6 * num_swim)
        self.num_swim = num_swim

    def close(self):
        if self.h:
            L.lib().tfs_velocities_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_preconditioner(self, kind: int):
        """PRECOND_REFERENCE (default) or PRECOND_BLOCKS for the blob solve."""
        L.check(L.lib().tfs_velocities_set_preconditioner(self.h, kind))

    def compute(self, x_segs, f_segs, x_blobs, x_bodies, v_bodies, *, allow_unconverged=False):
        """Returns (v_segs [6 per segment], f_blobs [3 per blob], solve stats);
        the blob forces warm-start the next call.  Host memory only (a
        device-memory handle is driven through the C-ABI, e.g. by
        filament.ElasticStep)."""
        if self.desc.memory != L.MEMORY_HOST:
            raise ValueError("compute() takes host arrays; this handle was made for device memory")
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x_segs, f_segs, x_blobs, x_bodies, v_bodies)]
        v_segs = np.zeros(6 * self.ns)
        st = SolveStats()
        rc = L.lib().tfs_velocities_compute(self.h, *[L.ptr(v) if v.size else None for v in a],
                                            L.ptr(v_segs) if v_segs.size else None, L.ptr(self.f_blobs),
                                            ctypes.byref(st))
        if not (allow_unconverged and rc == L.TFS_ERR_NOT_CONVERGED):
            L.check(rc)
        return v_segs, self.f_blobs.copy(), st.as_dict()

    def set_body_frame(self, blob_refs):
        """Body-frame blob positions for the direct free-body solve (one body)."""
        self._refs = np.ascontiguousarray(blob_refs, dtype=np.float64)
        L.check(L.lib().tfs_velocities_set_body_frame(self.h, L.ptr(self._refs)))

    def compute_free(self, x_segs, f_segs, x_blobs, x_bodies, f_bodies=None, q_bodies=None, *,
                     allow_unconverged=False):
        """Free bodies (a handle made with free_bodies=True): returns (v_segs,
        f_blobs, v_bodies [V, Omega per body], solve stats); the blob forces and
        body velocities warm-start the next call.  f_bodies: external force and
        torque per body, or None; q_bodies: the orientation (4 per body), needed
        by the direct solve.  Host memory only."""
        if self.desc.memory != L.MEMORY_HOST:
            raise ValueError("compute_free() takes host arrays; this handle was made for device memory")
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x_segs, f_segs, x_blobs, x_bodies)]
        fb = None if f_bodies is None else np.ascontiguousarray(f_bodies, dtype=np.float64)
        qb = None if q_bodies is None else np.ascontiguousarray(q_bodies, dtype=np.float64)
        v_segs = np.zeros(6 * self.ns)
        st = SolveStats()
        rc = L.lib().tfs_velocities_compute_free(self.h, *[L.ptr(v) if v.size else None for v in a], L.ptr(fb),
                                                 L.ptr(qb), L.ptr(v_segs) if v_segs.size else None,
                                                 L.ptr(self.f_blobs),
                                                 L.ptr(self.v_bodies), ctypes.byref(st))
        if not (allow_unconverged and rc == L.TFS_ERR_NOT_CONVERGED):
            L.check(rc)
        return v_segs, self.f_blobs.copy(), self.v_bodies.copy(), st.as_dict()

    def body_mobility(self, x_blobs, x_bodies):
        """make_body_reference_matrices (mobility_solver.cpp:2950-3047): per body
        the 6 x 6 mobility, column k the velocities V, Omega under a unit force
        (k < 3) or torque on the body with no filament forces, by six free-body
        solves at the given pose.  With several bodies each column loads every
        body at once, so the matrices include the bodies' mutual flow (the
        reference moves the other bodies far away first)."""
        # the segments carry no force; placed apart and far away they add no
        # near pairs (their positions do not enter the result)
        xb = np.asarray(x_blobs, dtype=np.float64).reshape(-1, 3)
        far = np.abs(xb).max() + 1e4 if xb.size else 1e4
        xs = np.zeros((self.ns, 3))
        xs[:, 0] = far + 10.0 * np.arange(self.ns)
        xs = xs.reshape(-1)
        fs = np.zeros(6 * self.ns)
        out = np.zeros((self.num_swim, 6, 6))
        for k in range(6):
            fe = np.zeros((self.num_swim, 6))
            fe[:, k] = 1.0
            self.f_blobs[:] = 0.0
            self.v_bodies[:] = 0.0
            _, _, v, _ = self.compute_free(xs, fs, x_blobs, x_bodies, fe.reshape(-1))
            out[:, :, k] = v.reshape(self.num_swim, 6)
        self.f_blobs[:] = 0.0
        self.v_bodies[:] = 0.0
        return out


@dataclass
class BeatSetup:
    """Host-side inputs of a prescribed-beat body (kept alive for the C call)."""

    frames: np.ndarray
    bases: np.ndarray
    x_blobs: np.ndarray
    s_values: np.ndarray
    omega0: np.ndarray
    phase0: np.ndarray
    angle0: np.ndarray
    gen_phase: np.ndarray
    gen_angle: np.ndarray
    fil_length: float


def beat_setup(body, *, phase_seed: int = 3, period: float = 1.0) -> BeatSetup:
    """Inputs for a geometry.Body: F&B beat along each filament's polar
    direction, natural frequency 2*pi/period, seeded uniform initial phases."""
    from . import geometry as G
    nf, nseg = body.num_fil, body.num_seg
    fil_length = body.dl * (nseg - 1)
    q_phase, q_angle = B.generalised_force_tables(nseg, fil_length, a=body.a)
    return BeatSetup(
        frames=B.frames(body.fil_normal),
        bases=np.ascontiguousarray(body.fil_base.reshape(-1)),
        x_blobs=np.ascontiguousarray(body.x_blobs),
        s_values=B.uniform_s(nseg),
        omega0=np.full(nf, 2.0 * np.pi / period),
        phase0=2.0 * np.pi * G.uniform_open(phase_seed, nf),
        angle0=np.zeros(nf),
        gen_phase=q_phase, gen_angle=q_angle, fil_length=fil_length)


class BeatSimulation:
    """Prescribed-beat time loop of one body, fully on the GPU."""

    def __init__(self, body, *, mode=SOLVE_FP64, dt=1.0 / 300.0, restart=50, tol=1e-8, inner_tol=1e-3,
                 max_cycles=20, max_refinements=10, dynamic_phase=True, dynamic_angle=True,
                 gen_force_factor=1.0, torsion_factor=0.046, stream=None, setup: BeatSetup | None = None,
                 warm_start: int = 0, precond: int = PRECOND_REFERENCE, free_body: bool = False):
        self.setup = setup or beat_setup(body)
        su = self.setup
        d = BeatDesc()
        d.solver = solver_desc(body.num_fil, body.num_seg, body.num_blobs, dynamic_phase=dynamic_phase,
                               dynamic_angle=dynamic_angle, mode=mode, seg_radius=body.a,
                               blob_radius=body.a, restart=restart, max_cycles=max_cycles, tol=tol,
                               inner_tol=inner_tol, max_refinements=max_refinements, stream=stream,
                               free_bodies=free_body)
        d.dt, d.fil_length = dt, su.fil_length
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (su.frames, su.bases, su.x_blobs, su.s_values, su.omega0, su.phase0, su.angle0,
                       B.BEAT_A.reshape(-1), B.BEAT_B.reshape(-1), su.gen_phase, su.gen_angle)]
        (d.frames, d.bases, d.x_blobs, d.s_values, d.omega0, d.phase0, d.angle0, d.beat_A, d.beat_B,
         d.gen_phase_forces, d.gen_angle_forces) = [a.ctypes.data for a in self._keep]
        d.n_gen_phase, d.n_gen_angle = len(su.gen_phase), len(su.gen_angle)
        d.gen_force_factor, d.torsion_factor = gen_force_factor, torsion_factor
        for i in range(3):
            d.body_x[i] = float(body.centre[i])
        self.desc = d
        self.num_fil, self.num_seg = body.num_fil, body.num_seg
        h = ctypes.c_void_p()
        L.check(L.lib().tfs_beat_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        sp = ctypes.c_void_p()
        L.check(L.lib().tfs_beat_solver(self.h, ctypes.byref(sp)))
        n = ctypes.c_int64()
        L.check(L.lib().tfs_solver_size(sp, ctypes.byref(n)))
        self.n = n.value
        if warm_start:
            self.set_warm_start(warm_start)
        if precond != PRECOND_REFERENCE:
            L.check(L.lib().tfs_solver_set_preconditioner(sp, precond))

    def close(self):
        if self.h:
            L.lib().tfs_beat_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, mode: int):
        L.check(L.lib().tfs_beat_set_mode(self.h, mode))

    def set_warm_start(self, order: int):
        """Initial guess of later solves: 0 the previous solution (the
        reference's), 1 / 2 its linear / quadratic extrapolation in time."""
        L.check(L.lib().tfs_beat_set_warm_start(self.h, order))

    def rpy_handle(self) -> ctypes.c_void_p:
        """The tfs_rpy handle the driver's solver evaluates with."""
        sp, rp = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().tfs_beat_solver(self.h, ctypes.byref(sp)))
        L.check(L.lib().tfs_solver_rpy(sp, ctypes.byref(rp)))
        return rp

    def step(self):
        st = SolveStats()
        L.check(L.lib().tfs_beat_step(self.h, ctypes.byref(st)))
        return st.as_dict()

    def state(self):
        nf, ns = self.num_fil, self.num_fil * self.num_seg
        ph, an, xs, sol = np.zeros(nf), np.zeros(nf), np.zeros(3 * ns), np.zeros(self.n)
        L.check(L.lib().tfs_beat_state(self.h, L.ptr(ph), L.ptr(an), L.ptr(xs), L.ptr(sol)))
        return {"phase": ph, "angle": an, "x_segs": xs, "solution": sol}

    def body(self):
        """(centre [3], orientation quaternion [4] (w, x, y, z), velocity [6])."""
        x, q, v = np.zeros(3), np.zeros(4), np.zeros(6)
        L.check(L.lib().tfs_beat_body(self.h, L.ptr(x), L.ptr(q), L.ptr(v)))
        return x, q, v

    def restore(self, phase, angle=None, solution=None):
        """Continue from saved state: the next step solves at these phases
        (no update first) from ``solution`` as its initial guess."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (phase, angle, solution)]
        if arrs[0].size != self.num_fil or (arrs[1] is not None and arrs[1].size != self.num_fil) or \
                (arrs[2] is not None and arrs[2].size != self.n):
            raise ValueError("restore: phase/angle need num_fil values, solution the system size")
        L.check(L.lib().tfs_beat_restore(self.h, *[L.ptr(a) if a is not None else None for a in arrs]))

    def velocities(self):
        """(v_segs [6 Ns], v_blobs [3 Nb]): M lambda of the last solution in
        FP64, rotational slots zero."""
        ns, nb = self.num_fil * self.num_seg, self.desc.solver.num_blobs
        vs, vb = np.zeros(6 * ns), np.zeros(3 * nb)
        L.check(L.lib().tfs_beat_velocities(self.h, L.ptr(vs), L.ptr(vb) if nb else None))
        return vs, vb


@dataclass
class _Concatenated:
    """The Body fields BeatSimulation reads, for bodies held one after another."""

    num_fil: int
    num_seg: int
    num_blobs: int
    a: float
    centre: np.ndarray


class EnsembleBeat(BeatSimulation):
    """Prescribed-beat time loop of independent bodies of uneven size on one
    handle: the bodies' arrays one after another, every timestep's solve one
    lock-step GMRES per body (tfs_solver_set_independent_bodies).  Each body
    follows exactly the trajectory its own BeatSimulation would (same setup,
    same solver settings).  Bodies are held still (no prescribed motion)."""

    def __init__(self, bodies, *, setups=None, graphs: bool = True, **kw):
        if not bodies:
            raise ValueError("EnsembleBeat: no bodies")
        nseg, a, dl = bodies[0].num_seg, bodies[0].a, bodies[0].dl
        if any(b.num_seg != nseg or b.a != a or b.dl != dl for b in bodies):
            raise ValueError("EnsembleBeat: every body needs the same segment count, radius and spacing")
        setups = setups or [beat_setup(b, phase_seed=3 + i) for i, b in enumerate(bodies)]
        cat = lambda k: np.ascontiguousarray(np.concatenate([np.asarray(getattr(s, k)).reshape(-1)
                                                             for s in setups]))
        s0 = setups[0]
        su = BeatSetup(frames=cat("frames"), bases=cat("bases"), x_blobs=cat("x_blobs"), s_values=s0.s_values,
                       omega0=cat("omega0"), phase0=cat("phase0"), angle0=cat("angle0"),
                       gen_phase=s0.gen_phase, gen_angle=s0.gen_angle, fil_length=s0.fil_length)
        body = _Concatenated(num_fil=sum(b.num_fil for b in bodies), num_seg=nseg,
                             num_blobs=sum(b.num_blobs for b in bodies), a=a, centre=np.zeros(3))
        super().__init__(body, setup=su, **kw)
        self.fil_per_body = np.array([b.num_fil for b in bodies], dtype=np.int32)
        self.blobs_per_body = np.array([b.num_blobs for b in bodies], dtype=np.int32)
        self.num_bodies = len(bodies)
        sp = self._solver()
        set_independent_bodies(sp, self.fil_per_body, self.blobs_per_body)
        if not graphs:
            L.check(L.lib().tfs_solver_set_graphs(sp, 0))

    def _solver(self):
        sp = ctypes.c_void_p()
        L.check(L.lib().tfs_beat_solver(self.h, ctypes.byref(sp)))
        return sp

    def body_stats(self) -> dict:
        return body_stats(self._solver(), self.num_bodies)

    def split(self, state: dict) -> list:
        """Per-body pieces of state(): phase, angle, x_segs and solution (the
        solution in the one-body layout [lambda_s | lambda_b | phase | angle])."""
        nf, nb, nseg = self.fil_per_body, self.blobs_per_body, self.num_seg
        fo = np.concatenate([[0], np.cumsum(nf)])
        bo = np.concatenate([[0], np.cumsum(nb)])
        sol = state["solution"]
        ns3 = 3 * int(fo[-1]) * nseg
        nfil = int(fo[-1])
        off_phase = ns3 + 3 * int(bo[-1])
        dyn_p, dyn_a = self.desc.solver.dynamic_phase, self.desc.solver.dynamic_angle
        off_angle = off_phase + (nfil if dyn_p else 0)
        out = []
        for i in range(self.num_bodies):
            f0, f1, b0, b1 = int(fo[i]), int(fo[i + 1]), int(bo[i]), int(bo[i + 1])
            parts = [sol[3 * f0 * nseg:3 * f1 * nseg], sol[ns3 + 3 * b0:ns3 + 3 * b1]]
            if dyn_p:
                parts.append(sol[off_phase + f0:off_phase + f1])
            if dyn_a:
                parts.append(sol[off_angle + f0:off_angle + f1])
            out.append({"phase": state["phase"][f0:f1], "angle": state["angle"][f0:f1],
                        "x_segs": state["x_segs"][3 * f0 * nseg:3 * f1 * nseg], "solution": np.concatenate(parts)})
        return out
