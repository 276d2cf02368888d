"""How tfs_elastic_step leaves the handles it is given.

* The step's error_norm is the norm of the residual of its last evaluation,
  read back from the Broyden engine (tfs_broyden_read_error), whether that
  evaluation updated the engine's error (slot 0) or not (slot 1).
* The velocities handle's blob-motion mode is the caller's again after a step:
  a later tfs_velocities_compute on the same handle with moved blobs solves
  for the new blob geometry, as a fresh handle does (the direct blob solve,
  whose stored inverse would otherwise be reused)."""
import ctypes

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
from timfilsimmod_amd.solver import BLOB_DIRECT, SOLVE_FP64, ElasticVelocities, SolveStats

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")


def small_case():
    body = G.make_body(4, 10, 300, 8.0)
    nfil, nseg = body.num_fil, body.num_seg
    dl, kb, dt = 2.2, 1800.0, 0.1
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([n], [ex])[0].as_quat(), 1) for n in body.fil_normal])
    qs = np.repeat(q0[:, None, :], nseg, axis=1)
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    return body, dict(dl=dl, kb=kb, dt=dt, efm=efm, qs=qs, pert=pert)


def velocities_call(ev, xs, fs, xb, v_segs, f_blobs):
    st = SolveStats()
    z3, z6 = dev(np.zeros(3)), dev(np.zeros(6))
    L.check(L.lib().tfs_velocities_compute(ev.h, L.ptr(xs), L.ptr(fs), L.ptr(xb), L.ptr(z3), L.ptr(z6),
                                           L.ptr(v_segs), L.ptr(f_blobs), ctypes.byref(st)))
    torch.cuda.synchronize()
    return st


def test_error_norm_and_blob_motion_after_elastic_steps():
    body, c = small_case()
    nfil, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    ns = nfil * nseg
    print(f"case elastic step {nfil}x{nseg} + {nb} blobs, direct blob solve", flush=True)
    model = FM.FilamentModel(1, nfil, nseg, dl=c["dl"], kb=c["kb"], kt=c["kb"], dt=c["dt"], end_force=c["efm"],
                             memory=L.MEMORY_DEVICE)
    ev = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10, memory=L.MEMORY_DEVICE)
    ev.set_preconditioner(BLOB_DIRECT)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
    keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * nb))]
    step = None
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(c["qs"].reshape(-1)))
        model.set_perturbation(4, 8, dev(c["pert"].reshape(-1)))
        step = FM.ElasticStep(model, ev, eng, *keep, tol=1e-8, max_iter=400)
        for nt in range(2):
            st = step.step(nt)
            assert st["converged"] == 1 and st["broyden_iterations"] > 1
            # the last evaluation met the tolerance, so its residual is in slot 1
            last = eng.read_error(1)
            norm = float(torch.linalg.norm(last))
            print(f"step {nt}: {st['broyden_iterations']} iterations, error_norm {st['error_norm']:.3e}, "
                  f"||slot 1|| {norm:.3e}, ||slot 0|| {float(torch.linalg.norm(eng.read_error(0))):.3e}", flush=True)
            assert abs(st["error_norm"] - norm) <= 1e-12 * norm
            assert st["error_norm"] < float(torch.linalg.norm(eng.read_error(0)))

        # the caller's handle afterwards: moved (stretched) blobs are solved for anew
        xs, fs = dev(model.read(FM.X)), dev(np.random.default_rng(5).normal(size=6 * ns))
        moved = dev(1.05 * body.x_blobs)
        got_v, got_f = dev(np.zeros(6 * ns)), dev(np.zeros(3 * nb))
        velocities_call(ev, xs, fs, moved, got_v, got_f)
        fresh = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                                  memory=L.MEMORY_DEVICE)
        fresh.set_preconditioner(BLOB_DIRECT)
        try:
            ref_v, ref_f = dev(np.zeros(6 * ns)), dev(np.zeros(3 * nb))
            velocities_call(fresh, xs, fs, moved, ref_v, ref_f)
        finally:
            fresh.close()
        err = float(torch.linalg.norm(got_v - ref_v) / torch.linalg.norm(ref_v))
        print(f"velocities after the elastic step, moved blobs: relative difference to a fresh handle {err:.2e}",
              flush=True)
        assert err <= 1e-10
        assert float(torch.linalg.norm(got_f - ref_f) / torch.linalg.norm(ref_f)) <= 1e-9
    finally:
        if step is not None:
            step.close()
        model.close()
        ev.close()
        eng.close()
