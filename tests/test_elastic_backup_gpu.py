"""Checkpoint and resume of the elastic step (tfs_elastic_backup / _restore,
datafiles.write_elastic_backup / resume_elastic) on the GPU:

* the record's length is the reference's data_per_swimmer
  (cilia_sim_main.cu:118-140: 19 + NFIL (9 + 28 NSEG), plus 9 NBLOB in the
  Broyden-for-everything form) and its fields sit where filament::write_backup
  and rigid_body::write_backup put them (segment positions, quaternions,
  multipliers checked against the handle's own arrays);
* a run restored in memory from the backup of step 1 takes step 2 exactly as
  the uninterrupted run (direct blob solve and the Broyden-for-everything form:
  within 1e-12; a free body, whose GMRES warm start is not part of a backup:
  within 1e-8), and through the file, written with the reference's 10
  significant digits, within 1e-6."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from oracle import swimmer_oracle as S
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import datafiles as D
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
from timfilsimmod_amd.solver import BLOB_DIRECT, SOLVE_FP64, ElasticVelocities

pytestmark = pytest.mark.gpu

NFIL, NSEG, NB, R = 4, 20, 400, 10.0


def dev(a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")


def make_run(form):
    """A fresh elastic step of the 4 x 20 + 400 blob case in one of the forms:
    'direct' (body held still, dense blob solve), 'everything' (blob forces as
    Broyden unknowns), 'free' (free body, GMRES blob solve)."""
    body = G.make_body(NFIL, NSEG, NB, R)
    dl, kb = 2.2, 1800.0
    efm = 220.0 * kb / (dl * dl * NSEG * NSEG)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    pert = np.random.default_rng(0).standard_normal((NFIL, 6)) * 0.1 * efm
    model = FM.FilamentModel(1, NFIL, NSEG, dl=dl, kb=kb, kt=kb, dt=0.1, end_force=efm, memory=L.MEMORY_DEVICE)
    model.set_state(dev(body.fil_base.reshape(-1)), dev(np.repeat(q0[:, None, :], NSEG, axis=1).reshape(-1)))
    model.set_perturbation(4, 12, dev(pert.reshape(-1)))
    keep = {"model": model, "f_blobs": dev(np.zeros(3 * NB))}
    ev, kw = None, {}
    if form == "direct":
        ev = ElasticVelocities(1, NFIL * NSEG, NB, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                               memory=L.MEMORY_DEVICE)
        ev.set_preconditioner(BLOB_DIRECT)
        eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
        arrays = (dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)))
    elif form == "everything":
        eng = BroydenSolver(model.block_sizes(blobs_per_swim=NB), memory=L.MEMORY_DEVICE)
        arrays = (dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)))
        kw = dict(broyden_blobs=NB, blob_radius=1.0)
    else:
        model.set_body(dev(body.centre), dev([1.0, 0.0, 0.0, 0.0]),
                       dev(np.array([body.fil_base[i] - body.centre for i in range(NFIL)]).reshape(-1)), NB,
                       dev(body.blob_refs.reshape(-1)))
        ev = ElasticVelocities(1, NFIL * NSEG, NB, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                               memory=L.MEMORY_DEVICE, free_bodies=True)
        eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
        arrays = (None, None, None)
        N_ref = S.body_mobility_reference(O.dense_mobility(np.zeros(0), body.x_blobs), body.blob_refs)
        keep["nmob"] = dev(N_ref.reshape(-1))
        kw = dict(body_mobility=keep["nmob"], sphere_radius=R)
    keep.update(ev=ev, eng=eng, arrays=arrays)
    keep["step"] = FM.ElasticStep(model, ev, eng, *arrays, keep["f_blobs"], tol=1e-8, max_iter=400, **kw)
    return keep


def close(run):
    run["step"].close()
    run["model"].close()
    if run["ev"] is not None:
        run["ev"].close()
    run["eng"].close()


@pytest.mark.parametrize("form,tol", [("direct", 1e-12), ("everything", 1e-12), ("free", 1e-8)])
def test_restored_run_takes_the_same_step(form, tol, tmp_path):
    print(f"case elastic backup and resume, {form}", flush=True)
    a = make_run(form)
    try:
        a["step"].step(0)
        a["step"].step(1)
        snap = a["step"].backup()
        path = str(tmp_path / "run.backup")
        D.write_elastic_backup(path, 1, a["step"])
        # the record's layout
        per = 19 + NFIL * (9 + 28 * NSEG) + (9 * NB if form == "everything" else 0)
        assert snap.size == per
        fil = snap[:NFIL * (9 + 28 * NSEG)].reshape(NFIL, 9 + 28 * NSEG)
        seg = fil[:, :19 * NSEG].reshape(NFIL, NSEG, 19)
        m = a["model"]
        assert np.array_equal(seg[..., 0:4].reshape(-1), m.read(FM.Q))
        assert np.array_equal(seg[..., 4:7].reshape(-1), m.read(FM.X))
        assert np.array_equal(seg[..., 13:16].reshape(-1), m.read(FM.U))
        mult = m.read(FM.MULT).reshape(NFIL, NSEG + 1, 3)
        assert np.array_equal(fil[:, 19 * NSEG:19 * NSEG + 3], mult[:, 0])             # tether
        assert np.array_equal(fil[:, 19 * NSEG + 9:19 * NSEG + 12], mult[:, NSEG])     # clamp
        assert np.array_equal(fil[:, 19 * NSEG + 18:19 * NSEG + 21], mult[:, 1])       # lambda_0
        if form == "everything":
            assert np.array_equal(snap[-9 * NB::3], a["f_blobs"].cpu().numpy())
        st_a = a["step"].step(2)
        x_a = m.read(FM.X)
    finally:
        close(a)
    for source in ("memory", "file"):
        b = make_run(form)
        try:
            if source == "memory":
                b["step"].restore(snap)
            else:
                assert D.resume_elastic(path, b["step"]) == 1
            st_b = b["step"].step(2)
            x_b = b["model"].read(FM.X)
        finally:
            close(b)
        d = float(np.abs(x_a - x_b).max())
        print(source, "iterations", st_a["broyden_iterations"], st_b["broyden_iterations"], "max |dx|", d, flush=True)
        if source == "memory":
            assert d <= tol and st_b["broyden_iterations"] == st_a["broyden_iterations"]
        else:
            assert d <= 1e-6 and abs(st_b["broyden_iterations"] - st_a["broyden_iterations"]) <= 2
