"""Simulation data files in the reference's formats, over include/tfs_io.h.

The reference saves a run as text records, one per line (cilia_sim_main.cu:
168-210 and 572-670, mobility_solver.cpp:3051-3173, swimmer.cpp:1110-1148,
rigid_body.cpp:531-578, filament.cpp:1932-2020); its flow-field program reads
them back (flow_field_main.cpp, flow_field_evaluator.cu:315-544).  Writing,
parsing and the state -> position reconstruction all run in libtfs_rpy.so.

``SimulationFiles`` names the files of one run; ``BeatRecorder`` writes a
``solver.BeatSimulation``'s saves and backups the way a prescribed-cilia run
of the reference does, and ``resume_beat`` restarts a simulation from them;
``flow_field_run`` is the flow-field program over those files on the GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L

IO_END = 100
APPEND, WITH_STEP, DEFAULT_FLOAT = 1, 2, 4


class IoParam(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("value", ctypes.c_double), ("is_integer", ctypes.c_int32)]


_vp = ctypes.c_void_p
_i32, _i64 = ctypes.c_int32, ctypes.c_int64
L.extra_signatures({
    "tfs_io_write_par": (ctypes.c_int, [ctypes.c_char_p, _i32, ctypes.POINTER(IoParam)]),
    "tfs_io_read_par": (ctypes.c_int, [ctypes.c_char_p, _i32, ctypes.POINTER(IoParam), ctypes.POINTER(_i32)]),
    "tfs_io_write_record": (ctypes.c_int, [ctypes.c_char_p, _i32, _i64, _i64, _vp]),
    "tfs_io_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "tfs_io_close": (ctypes.c_int, [_vp]),
    "tfs_io_read_record": (ctypes.c_int, [_vp, _i32, _i64, ctypes.POINTER(_i64), _vp]),
    "tfs_io_peek_count": (ctypes.c_int, [_vp, ctypes.POINTER(_i64)]),
    "tfs_io_rewind": (ctypes.c_int, [_vp]),
    "tfs_io_count_records": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_i64)]),
    "tfs_io_sources_from_states": (ctypes.c_int, [_i32, _i32, _i32, _i32, ctypes.c_double, _i32, _vp, _vp,
                                                  _vp, _vp, _vp, _vp, _vp]),
    "tfs_io_flow_run": (ctypes.c_int, [ctypes.c_char_p, _i32, _vp, ctypes.POINTER(_i64)]),
})


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


# ---- .par -----------------------------------------------------------------
def write_par(path, params) -> None:
    """``params``: (NAME, value) pairs in file order; Python ints (and bools)
    are written as integers, floats in the default format."""
    items = list(params.items()) if isinstance(params, dict) else list(params)
    arr = (IoParam * max(len(items), 1))()
    for i, (name, value) in enumerate(items):
        b = name.encode()
        if not b or len(b) >= 64:
            raise ValueError(f"parameter name {name!r} must have 1..63 bytes")
        arr[i].name = b
        arr[i].value = float(value)
        arr[i].is_integer = int(isinstance(value, (int, np.integer)))
    L.check(L.lib().tfs_io_write_par(_path(path), len(items), arr))


def read_par(path) -> dict:
    """NAME -> value (int where the file holds an integer token), in file order."""
    cap = 256
    arr = (IoParam * cap)()
    n = _i32()
    L.check(L.lib().tfs_io_read_par(_path(path), cap, arr, ctypes.byref(n)))
    if n.value > cap:
        raise ValueError(f"{path}: more than {cap} parameters")
    return {arr[i].name.decode(): (int(arr[i].value) if arr[i].is_integer else arr[i].value)
            for i in range(n.value)}


# ---- records ----------------------------------------------------------------
def write_record(path, values, *, step=None, append=True, default_float=False) -> None:
    v = _f64(values)
    flags = (APPEND if append else 0) | (WITH_STEP if step is not None else 0) | \
        (DEFAULT_FLOAT if default_float else 0)
    L.check(L.lib().tfs_io_write_record(_path(path), flags, int(step or 0), v.size, v.ctypes.data))


def count_records(path) -> int:
    n = _i64()
    L.check(L.lib().tfs_io_count_records(_path(path), ctypes.byref(n)))
    return n.value


class RecordReader:
    """Sequential records of one file; ``read`` returns (step, values) or
    None at the end of the file."""

    def __init__(self, path):
        self.path = os.fspath(path)
        h = _vp()
        L.check(L.lib().tfs_io_open(_path(path), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tfs_io_close(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read(self, count: int, *, with_step: bool = True):
        out = np.empty(int(count))
        step = _i64()
        rc = L.lib().tfs_io_read_record(self.h, int(with_step), out.size, ctypes.byref(step), out.ctypes.data)
        if rc == IO_END:
            return None
        L.check(rc)
        return (step.value if with_step else None), out

    def peek_count(self):
        n = _i64()
        rc = L.lib().tfs_io_peek_count(self.h, ctypes.byref(n))
        if rc == IO_END:
            return None
        L.check(rc)
        return n.value

    def rewind(self):
        L.check(L.lib().tfs_io_rewind(self.h))

    def records(self, count: int, *, with_step: bool = True):
        while True:
            r = self.read(count, with_step=with_step)
            if r is None:
                return
            yield r


def read_records(path, count: int, *, with_step: bool = True):
    """All records of a file: (steps [n] or None, values [n, count])."""
    with RecordReader(path) as r:
        recs = list(r.records(count, with_step=with_step))
    vals = np.stack([v for _, v in recs]) if recs else np.zeros((0, count))
    steps = np.array([s for s, _ in recs], dtype=np.int64) if with_step else None
    return steps, vals


# ---- reconstruction -----------------------------------------------------------
def sources_from_states(body_states, seg_states, fil_refs, blob_refs, *, num_fil, num_seg, dl,
                        quaternion_based=True):
    """(x_segs, x_blobs, x_bodies) of one save (flow_field_evaluator.cu:401-498)."""
    bs = _f64(body_states)
    if bs.size % 7:
        raise ValueError("body states hold 7 values per body")
    nbod = bs.size // 7
    ss, fr = _f64(seg_states), _f64(fil_refs)
    br = _f64(blob_refs) if blob_refs is not None else np.zeros(0)
    nblob = br.size // 3
    per = 4 if quaternion_based else 3
    if ss.size != per * nbod * num_fil * num_seg or fr.size < 3 * num_fil:
        raise ValueError("segment states or filament references do not match the sizes")
    xs = np.zeros(3 * nbod * num_fil * num_seg)
    xb = np.zeros(3 * nbod * nblob)
    xbod = np.zeros(3 * nbod)
    L.check(L.lib().tfs_io_sources_from_states(
        nbod, num_fil, num_seg, nblob, float(dl), int(bool(quaternion_based)), bs.ctypes.data, ss.ctypes.data,
        fr.ctypes.data, br.ctypes.data if nblob else None, xs.ctypes.data, xb.ctypes.data if nblob else None,
        xbod.ctypes.data))
    return xs, xb, xbod


# ---- flow-field program -------------------------------------------------------
def write_flow_parameters(path, sim_name, nflow, *, quaternion_based=True, periodic=False, cycles=1,
                          remove_net_force=False, advance_tracers=False) -> None:
    """The flow-field program's parameter file (flow_field_main.cpp:13-20, 112-151)."""
    keys = [("SIMULATION_NAME", sim_name), ("NFLOW", int(nflow)),
            ("QUATERNION_BASED_DATA", int(bool(quaternion_based))),
            ("TREAT_DATA_AS_PERIODIC", int(bool(periodic))), ("NUMBER_OF_CYCLES_FOR_PERIODIC_DATA", int(cycles)),
            ("REMOVE_NET_FORCE_STOKESLET", int(bool(remove_net_force))),
            ("ADVANCE_TRACER_POSITIONS", int(bool(advance_tracers)))]
    with open(path, "w") as f:
        for k, v in keys:
            f.write(f"{k} {v}\n")


def flow_field_run(input_path, *, precision="fp64", x_points=None) -> int:
    """Runs the flow-field program on the GPU; returns the saves evaluated."""
    prec = {"fp64": L.PRECISION_FP64, "fp32": L.PRECISION_FP32_ACC64}[precision]
    xp = _f64(x_points) if x_points is not None else None
    n = _i64()
    L.check(L.lib().tfs_io_flow_run(_path(input_path), prec, xp.ctypes.data if xp is not None else None,
                                    ctypes.byref(n)))
    return n.value


# ---- one run's files ------------------------------------------------------------
class SimulationFiles:
    """The files of one run, named as the reference names them
    (cilia_sim_main.cu:93-103, swimmer.cpp:1114, rigid_body.cpp:535)."""

    SUFFIXES = {
        "par": ".par", "backup": ".backup", "body_states": "_body_states.dat",
        "seg_states": "_seg_states.dat", "body_vels": "_body_vels.dat", "seg_vels": "_seg_vels.dat",
        "blob_forces": "_blob_forces.dat", "seg_forces": "_seg_forces.dat",
        "fil_references": "_fil_references.dat", "blob_references": "_blob_references.dat",
        "filament_phases": "_filament_phases.dat", "filament_angles": "_filament_shape_rotation_angles.dat",
        "true_states": "_true_states.dat", "flow_locations": "_flow_locations.dat",
        "flow_velocities": "_flow_velocities.dat",
    }

    def __init__(self, name):
        self.name = os.fspath(name)

    def path(self, kind: str) -> str:
        return self.name + self.SUFFIXES[kind]

    def write_references(self, fil_refs, blob_refs=None) -> None:
        """Reference positions in the body frame, default format, one line
        each (swimmer.cpp:1110-1125, rigid_body.cpp:531-548)."""
        write_record(self.path("fil_references"), fil_refs, append=False, default_float=True)
        if blob_refs is not None and np.size(blob_refs):
            write_record(self.path("blob_references"), blob_refs, append=False, default_float=True)

    def read_references(self, num_fil: int, num_blobs: int):
        with RecordReader(self.path("fil_references")) as r:
            _, fr = r.read(3 * num_fil, with_step=False)
        br = np.zeros(0)
        if num_blobs:
            with RecordReader(self.path("blob_references")) as r:
                _, br = r.read(3 * num_blobs, with_step=False)
        return fr, br


class BeatRecorder:
    """Writes a prescribed-beat simulation (one body, held still) the way the
    reference's prescribed-cilia run writes its files: the .par and reference
    positions once, then per save the state, force and velocity records and
    the backup (cilia_sim_main.cu:572-670, mobility_solver.cpp:3051-3173)."""

    def __init__(self, sim, name, *, dt, total_time_steps, plot_frequency=1, steps_per_period=300,
                 tol=1e-4, kb=1800.0, kt=1800.0, mu=1.0, period=1.0):
        self.sim, self.files = sim, SimulationFiles(name)
        self.period = float(period)
        su = sim.setup
        self.nfil, self.nseg = sim.num_fil, sim.num_seg
        self.nblob = sim.desc.solver.num_blobs
        self.centre = np.array(sim.desc.body_x[:3], dtype=np.float64)
        self.body_v = np.array(sim.desc.body_v[:6], dtype=np.float64)
        self.omega0 = np.asarray(su.omega0, dtype=np.float64)
        self.dynamic_angle = bool(sim.desc.solver.dynamic_angle)
        a = float(sim.desc.solver.seg_radius)
        dl = su.fil_length / (self.nseg - 1)
        write_par(self.files.path("par"), [
            ("NFIL", self.nfil), ("NSEG", self.nseg), ("NSWIM", 1), ("NBLOB", self.nblob), ("MU", float(mu)),
            ("RSEG", a), ("DL", float(dl)), ("FIL_LENGTH", float(su.fil_length)),
            ("RBLOB", float(sim.desc.solver.blob_radius)), ("KB", float(kb)), ("KT", float(kt)), ("TOL", float(tol)),
            ("TOTAL_TIME_STEPS", int(total_time_steps)), ("DT", float(dt)),
            ("PLOT_FREQUENCY_IN_STEPS", int(plot_frequency)), ("STEPS_PER_PERIOD", int(steps_per_period)),
            ("PRESCRIBED_CILIA", 1),
            ("TORSIONAL_SPRING_MAGNITUDE_FACTOR", float(sim.desc.torsion_factor))])
        fil_refs = su.bases.reshape(-1, 3) - self.centre
        blob_refs = su.x_blobs.reshape(-1, 3) - self.centre if self.nblob else None
        self.files.write_references(fil_refs.reshape(-1), None if blob_refs is None else blob_refs.reshape(-1))

    def _generalised_velocities(self, sol):
        nf = self.nfil
        off = 3 * (self.nfil * self.nseg + self.nblob)
        dyn_phase = bool(self.sim.desc.solver.dynamic_phase)
        phase_dot = sol[off:off + nf] if dyn_phase else self.omega0.copy()
        angle_dot = sol[off + nf * dyn_phase:off + nf * dyn_phase + nf] if self.dynamic_angle else np.zeros(nf)
        return phase_dot, angle_dot

    def save(self, step: int) -> None:
        """One save after the solve of `step` (cilia_sim_main.cu:572-670)."""
        f, st = self.files, self.sim.state()
        ns = self.nfil * self.nseg
        sol = st["solution"]
        body_state = np.concatenate([self.centre, [1.0, 0.0, 0.0, 0.0]])
        write_record(f.path("seg_states"), st["x_segs"], step=step)
        write_record(f.path("body_states"), body_state, step=step)
        v_segs, v_blobs = self.sim.velocities()
        seg_forces = np.zeros((ns, 6))
        seg_forces[:, :3] = sol[:3 * ns].reshape(ns, 3)
        write_record(f.path("body_vels"), self.body_v)
        write_record(f.path("blob_forces"), sol[3 * ns:3 * (ns + self.nblob)], step=step)
        write_record(f.path("seg_vels"), v_segs, step=step)
        write_record(f.path("seg_forces"), seg_forces, step=step)
        self.backup(step, st)
        write_record(f.path("filament_phases"), st["phase"], step=step)
        true_state = [self.period, *st["phase"]]
        if self.dynamic_angle:
            write_record(f.path("filament_angles"), st["angle"], step=step)
            true_state += list(st["angle"])
        write_record(f.path("true_states"), true_state, step=step)

    def backup(self, step: int, st=None) -> None:
        """The .backup of the prescribed build (filament.cpp:1952-1983 per
        filament, then rigid_body.cpp:556-566), overwritten."""
        st = st or self.sim.state()
        phase_dot, angle_dot = self._generalised_velocities(st["solution"])
        q = np.array([1.0, 0.0, 0.0, 0.0])
        fil = np.empty((self.nfil, 13))
        fil[:, 0], fil[:, 1], fil[:, 2] = st["phase"], phase_dot, self.omega0
        fil[:, 3:7], fil[:, 7:11] = q, q
        fil[:, 11] = st["angle"] if self.dynamic_angle else 0.0
        fil[:, 12] = angle_dot if self.dynamic_angle else 0.0
        body = np.concatenate([q, self.centre, self.centre, self.centre, np.zeros(6)])
        write_record(self.files.path("backup"), np.concatenate([fil.reshape(-1), body]), step=step, append=False)


def read_beat_backup(path, num_fil: int, num_swim: int = 1):
    """Fields of a prescribed-cilia .backup (cilia_sim_main.cu:118-160:
    13 values per filament, then 19 per body, per swimmer)."""
    per = 19 + 13 * num_fil
    with RecordReader(path) as r:
        rec = r.read(per * num_swim)
    if rec is None:
        raise ValueError(f"{path}: empty backup")
    step, v = rec
    v = v.reshape(num_swim, per)
    fil = v[:, :13 * num_fil].reshape(num_swim, num_fil, 13)
    body = v[:, 13 * num_fil:]
    return {"step": step, "phase": fil[..., 0], "phase_dot": fil[..., 1], "omega0": fil[..., 2],
            "body_qm1": fil[..., 3:7], "body_q": fil[..., 7:11], "angle": fil[..., 11], "angle_dot": fil[..., 12],
            "body_orientation": body[:, 0:4], "body_x": body[:, 4:7], "body_xm1": body[:, 7:10],
            "body_xm2": body[:, 10:13], "body_u": body[:, 13:16], "body_um1": body[:, 16:19]}


def resume_beat(sim, name) -> int:
    """Restarts ``sim`` from the backup of run ``name`` and the forces of the
    same save (READ_INITIAL_CONDITIONS_FROM_BACKUP): the next step re-solves
    the saved step from the saved solution.  Returns the saved step."""
    files = SimulationFiles(name)
    nf, nseg = sim.num_fil, sim.num_seg
    nblob = sim.desc.solver.num_blobs
    b = read_beat_backup(files.path("backup"), nf)
    step = b["step"]
    ns = nf * nseg

    def record_at(kind, count):
        with RecordReader(files.path(kind)) as r:
            for s, v in r.records(count):
                if s == step:
                    return v
        raise ValueError(f"{files.path(kind)} holds no save of step {step}")

    sol = np.zeros(sim.n)
    sol[:3 * ns] = record_at("seg_forces", 6 * ns).reshape(ns, 6)[:, :3].reshape(-1)
    if nblob:
        sol[3 * ns:3 * (ns + nblob)] = record_at("blob_forces", 3 * nblob)
    off = 3 * (ns + nblob)
    dyn_phase = bool(sim.desc.solver.dynamic_phase)
    if dyn_phase:
        sol[off:off + nf] = b["phase_dot"][0]
    if sim.desc.solver.dynamic_angle:
        sol[off + nf * dyn_phase:off + nf * dyn_phase + nf] = b["angle_dot"][0]
    sim.restore(b["phase"][0], b["angle"][0], sol)
    return step


def write_elastic_backup(path, step: int, elastic, held_bodies=None) -> None:
    """The elastic run's backup file (cilia_sim_main.cu:611-625): one record,
    the step then every swimmer's state (filament.ElasticStep.backup),
    overwritten each time."""
    write_record(path, elastic.backup(held_bodies), step=step, append=False)


def resume_elastic(path, elastic) -> int:
    """Reads a backup written by write_elastic_backup (or by the reference, for
    the same sizes) into the step's state; returns the step it was taken at,
    so the run continues with nt = step + 1 (cilia_sim_main.cu:140-166)."""
    with RecordReader(path) as r:
        count = r.peek_count()
        if count is None:
            raise ValueError(f"{path}: empty backup")
        step, values = r.read(count - 1)
    elastic.restore(values)
    return int(step)
