"""The reference's file-driven runs on the GPU (include/tfs_io.h):

* the flow-field program over saved state/force files (flow_field_main.cpp,
  flow_field_evaluator.cu:315-579) against the FP64 oracle at sources that
  numpy and scipy rebuild from the same files (parsed with numpy), with the
  reference's tracer seeding, tracer advection and periodic cycles;
* a prescribed-beat run saved in the reference's files (seg_vels = M lambda
  against the oracle) and resumed from its backup."""
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from timfilsimmod_amd import datafiles as D
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import BeatSimulation

pytestmark = pytest.mark.gpu

NBOD, NFIL, NSEG, NBLOB, RSEG = 2, 6, 5, 40, 1.0
DT, PLOT, TOTAL = 0.01, 2, 5          # 1 + (5-1)/2 = 3 saves


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _quats(rng, n, spread):
    # rotations near the one taking e_x to e_z, so filaments rise from their bases
    base = Rotation.from_rotvec([0.0, -np.pi / 2, 0.0])
    r = Rotation.from_rotvec(spread * rng.standard_normal((n, 3))) * base
    q = r.as_quat()                       # (x, y, z, w)
    return np.roll(q, 1, axis=1)          # (w, x, y, z) as the files hold them


def write_run(name, *, quaternion_based=True, nblob=NBLOB, saves=3, seed=0):
    rng = np.random.default_rng(seed)
    files = D.SimulationFiles(name)
    D.write_par(files.path("par"), [
        ("NFIL", NFIL), ("NSEG", NSEG), ("NSWIM", NBOD), ("NBLOB", nblob), ("MU", 1.0), ("RSEG", RSEG),
        ("DL", 2.2 * RSEG), ("FIL_LENGTH", 2.2 * RSEG * (NSEG - 1)), ("RBLOB", 1.0), ("KB", 1800.0),
        ("KT", 1800.0), ("TOL", 1e-4), ("TOTAL_TIME_STEPS", TOTAL), ("DT", DT),
        ("PLOT_FREQUENCY_IN_STEPS", PLOT), ("STEPS_PER_PERIOD", 300), ("PRESCRIBED_CILIA", 0)])
    fil_refs = 6.0 * G.fibonacci_sphere(NFIL, 1.0)
    blob_refs = G.fibonacci_sphere(nblob, 5.0) if nblob else None
    files.write_references(fil_refs.reshape(-1), None if blob_refs is None else blob_refs.reshape(-1))
    for k in range(saves):
        step = PLOT * k
        X = np.array([[15.0, 0.0, 16.0], [-10.0, 60.0, 12.0]]) + rng.standard_normal((NBOD, 3))
        qb = np.roll(Rotation.random(NBOD, random_state=seed + k).as_quat(), 1, axis=1)
        if nblob == 0:     # half-space: keep every source above the wall z = 0
            qb = np.tile([1.0, 0.0, 0.0, 0.0], (NBOD, 1))
        D.write_record(files.path("body_states"), np.concatenate([X, qb], axis=1), step=step)
        if quaternion_based:
            seg = _quats(rng, NBOD * NFIL * NSEG, 0.3)
        else:
            seg = rng.uniform(-20, 20, (NBOD * NFIL * NSEG, 3)) + [0.0, 0.0, 30.0]
        D.write_record(files.path("seg_states"), seg, step=step)
        D.write_record(files.path("seg_forces"), rng.standard_normal(6 * NBOD * NFIL * NSEG), step=step)
        if nblob:
            D.write_record(files.path("blob_forces"), rng.standard_normal(3 * NBOD * nblob), step=step)
    return files


def expected_sources(files, k, quaternion_based, nblob):
    """Sources of save k rebuilt with numpy/scipy from the text files."""
    load = lambda kind: np.atleast_2d(np.loadtxt(files.path(kind)))
    body = load("body_states")[k, 1:].reshape(NBOD, 7)
    seg = load("seg_states")[k, 1:]
    fs = load("seg_forces")[k, 1:]
    fil_refs = np.loadtxt(files.path("fil_references")).reshape(NFIL, 3)
    X, q = body[:, :3], body[:, 3:]
    R = Rotation.from_quat(np.roll(q, -1, axis=1)).as_matrix()
    if quaternion_based:
        qs = seg.reshape(NBOD, NFIL, NSEG, 4)
        t = Rotation.from_quat(np.roll(qs.reshape(-1, 4), -1, axis=1)).as_matrix()[:, :, 0]
        t = t.reshape(NBOD, NFIL, NSEG, 3)
        base = X[:, None, :] + np.einsum("bij,fj->bfi", R, fil_refs)
        dl = 2.2 * RSEG
        steps = 0.5 * dl * (t[:, :, :-1] + t[:, :, 1:])
        xs = base[:, :, None, :] + np.concatenate([np.zeros((NBOD, NFIL, 1, 3)), np.cumsum(steps, axis=2)], axis=2)
    else:
        xs = seg
    xb, fb = np.zeros(0), np.zeros(0)
    if nblob:
        blob_refs = np.loadtxt(files.path("blob_references")).reshape(nblob, 3)
        xb = X[:, None, :] + np.einsum("bij,kj->bki", R, blob_refs)
        fb = load("blob_forces")[k, 1:]
    return xs.reshape(-1), fs, xb.reshape(-1), fb, X.reshape(-1)


def reference_seeding(n):
    k = np.arange(n, dtype=np.float64)
    row = np.floor(k / 551.0)
    return np.stack([0.0 * k, -2.2032e+03 + 8.0 * (k - 551.0 * row), 8.0 * (1.0 + row)], axis=1).reshape(-1)


@pytest.mark.parametrize("quaternion_based,nblob,remove,advance,cycles,precision", [
    (True, NBLOB, False, False, 1, "fp64"),
    (True, NBLOB, True, True, 2, "fp64"),
    (False, NBLOB, False, True, 1, "fp64"),
    (False, 0, False, False, 1, "fp64"),
    (True, NBLOB, True, True, 1, "fp32"),
])
def test_flow_run_from_files_matches_oracle(tmp_path, quaternion_based, nblob, remove, advance, cycles, precision):
    print(f"case q={quaternion_based} nblob={nblob} remove={remove} advance={advance} cycles={cycles} "
          f"{precision}", flush=True)
    nflow = 1200
    files = write_run(str(tmp_path / "run"), quaternion_based=quaternion_based, nblob=nblob)
    inp = tmp_path / "flow_field_simulation_parameters.input"
    D.write_flow_parameters(inp, "run", nflow, quaternion_based=quaternion_based, periodic=cycles > 1,
                            cycles=cycles, remove_net_force=remove, advance_tracers=advance)
    assert D.flow_field_run(inp, precision=precision) == 3 * cycles
    loc = np.atleast_2d(np.loadtxt(files.path("flow_locations")))
    vel = np.atleast_2d(np.loadtxt(files.path("flow_velocities")))
    assert loc.shape == vel.shape == (3 * cycles, 1 + 3 * nflow)
    assert list(vel[:, 0].astype(int)) == [0, 2, 4] * cycles
    x, v = reference_seeding(nflow), np.zeros(3 * nflow)
    tol = 1e-10 if precision == "fp64" else 2e-6
    for r in range(3 * cycles):
        if advance:
            x = x + DT * PLOT * v
        xs, fs, xb, fb, xbod = expected_sources(files, r % 3, quaternion_based, nblob)
        v = O.flow_velocities(x, xs, fs, xb, fb, xbod, num_bodies=NBOD, remove_net_force=remove)
        np.testing.assert_allclose(loc[r, 1:], x, rtol=1e-10, atol=1e-9)
        assert rel(vel[r, 1:], v) <= tol, (r, rel(vel[r, 1:], v))


def test_flow_run_with_caller_points(tmp_path):
    files = write_run(str(tmp_path / "run"))
    inp = tmp_path / "in.input"
    rng = np.random.default_rng(9)
    pts = rng.uniform(-30, 70, 3 * 500)
    D.write_flow_parameters(inp, "run", 500)
    assert D.flow_field_run(inp, x_points=pts) == 3
    vel = np.loadtxt(files.path("flow_velocities"))
    xs, fs, xb, fb, xbod = expected_sources(files, 1, True, NBLOB)
    assert rel(vel[1, 1:], O.flow_velocities(pts, xs, fs, xb, fb, xbod, num_bodies=NBOD)) <= 1e-10


def test_beat_run_saves_and_resumes(tmp_path):
    print("case beat save/resume 6x20, 200 blobs", flush=True)
    body = G.make_body(6, 20, 200, 8.0)
    kw = dict(tol=1e-11, restart=60, max_cycles=50)
    a = BeatSimulation(body, **kw)
    name = str(tmp_path / "beat")
    rec = D.BeatRecorder(a, name, dt=a.desc.dt, total_time_steps=10)
    for step in range(3):
        a.step()
        rec.save(step)
    files = rec.files
    ns, nb = 6 * 20, 200
    par = D.read_par(files.path("par"))
    assert par["NFIL"] == 6 and par["NSEG"] == 20 and par["NBLOB"] == 200 and par["PRESCRIBED_CILIA"] == 1
    for kind, count in [("seg_states", 3 * ns), ("body_states", 7), ("seg_forces", 6 * ns),
                        ("blob_forces", 3 * nb), ("seg_vels", 6 * ns), ("filament_phases", 6),
                        ("true_states", 1 + 12)]:
        steps, vals = D.read_records(files.path(kind), count)
        assert list(steps) == [0, 1, 2], kind
    _, bv = D.read_records(files.path("body_vels"), 6, with_step=False)
    assert bv.shape == (3, 6)
    # saved velocities are M lambda of the saved forces at the saved positions
    x = np.loadtxt(files.path("seg_states"))[-1, 1:]
    f = np.loadtxt(files.path("seg_forces"))[-1, 1:]
    fb = np.loadtxt(files.path("blob_forces"))[-1, 1:]
    vs_file = np.loadtxt(files.path("seg_vels"))[-1, 1:]
    vs, _ = O.full_mobility(x, body.x_blobs, f, fb, seg_dof=3)
    assert rel(vs_file.reshape(ns, 6)[:, :3], vs.reshape(ns, 6)[:, :3]) <= 1e-9
    assert np.all(vs_file.reshape(ns, 6)[:, 3:] == 0.0)
    # the backup holds the last save's phases
    bk = D.read_beat_backup(files.path("backup"), 6)
    st = a.state()
    assert bk["step"] == 2
    np.testing.assert_allclose(bk["phase"][0], st["phase"], rtol=1e-10)
    # resume: the first step re-solves save 2, the next one continues as the original run
    b = BeatSimulation(body, **kw)
    assert D.resume_beat(b, name) == 2
    st_b = b.step()
    assert st_b["initial_relres"] <= 1e-6
    np.testing.assert_allclose(b.state()["phase"], st["phase"], rtol=1e-10)
    a.step()
    b.step()
    sa, sb = a.state(), b.state()
    np.testing.assert_allclose(sb["phase"], sa["phase"], rtol=1e-9)
    assert rel(sb["solution"], sa["solution"]) <= 1e-7
