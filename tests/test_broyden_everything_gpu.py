"""The Broyden-for-everything form of the elastic step (USE_BROYDEN_FOR_EVERYTHING,
bodies held still) on the GPU against the oracle (oracle/swimmer_oracle.py
held_still_loop, pinned on the CPU by reaching the trajectory of the form that
solves for the blob forces):

* the residual's blob rows (tfs_broyden_compute_errors with blobs_per_swim)
  against the oracle's compute_errors for two swimmers, <= 1e-14;
* two timesteps of 4 filaments x 20 segments on 400 blobs through
  tfs_elastic_step: the same Broyden iteration counts as the oracle loop
  (within 10 %), positions within 1e-6, blob forces within 1e-5."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import broyden_oracle as BO
from oracle import filament_oracle as F
from oracle import swimmer_oracle as S
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("bdf2", [False, True])
def test_residual_blob_rows_match_oracle(bdf2):
    print(f"case residual with blob rows, 2 swimmers x 2 x 5 segments + 7 blobs, bdf2 {bdf2}", flush=True)
    rng = np.random.default_rng(5)
    nsw, nfil, nseg, nb = 2, 2, 5, 7
    ns = nsw * nfil * nseg
    st = {k: rng.standard_normal(3 * ns) for k in ("seg_x", "seg_xm1", "seg_xm2", "seg_u", "seg_um1")}
    st["v_segs"] = rng.standard_normal(6 * ns)
    for k in ("body_x", "body_xm1", "body_xm2", "body_u", "body_um1"):
        st[k] = rng.standard_normal(3 * nsw) * 1e-3
    st["v_bodies"] = rng.standard_normal(6 * nsw)
    st["x_blobs"] = rng.standard_normal(3 * nsw * nb) * 4
    st["v_blobs"] = rng.standard_normal(3 * nsw * nb)
    n = 6 * (ns + nsw) + 3 * nsw * nb
    eng = BroydenSolver([n // nsw] * nsw)
    try:
        big = eng.compute_errors(st, 0, num_swim=nsw, num_fil=nfil, num_seg=nseg, bdf2=bdf2, dt=0.3, dl=2.2,
                                 tol=1e-8, blobs_per_swim=nb)
        e = eng.read_error(0)
    finally:
        eng.close()
    want, want_big = BO.compute_errors(st, nsw, nfil, nseg, bdf2, 0.3, 2.2, 1e-8, blobs_per_swim=nb)
    print("rel err", rel(e, want), flush=True)
    assert rel(e, want) < 1e-14 and big == want_big


STEPS = 2


@pytest.fixture(scope="module")
def case():
    nfil, nseg, nb, R = 4, 20, 400, 10.0
    body = G.make_body(nfil, nseg, nb, R)
    dl, kb = 2.2, 1800.0
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    p = F.Params(nseg=nseg, dl=dl, kb=kb, kt=kb, dt=0.1, end_force=efm)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (nseg, 1)), nseg, dl) for i in range(nfil)]
    runs = S.held_still_loop(fils, body.x_blobs, p, STEPS, 1e-8, pert, True)
    return dict(body=body, p=p, q0=q0, pert=pert, runs=runs)


def test_broyden_for_everything_step_matches_oracle_loop(case):
    print("case Broyden-for-everything elastic step 4x20 + 400 blobs, tfs_elastic_step", flush=True)
    body, p = case["body"], case["p"]
    nfil, nseg, nb = body.num_fil, p.nseg, body.num_blobs
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    model = FM.FilamentModel(1, nfil, nseg, dl=p.dl, kb=p.kb, kt=p.kt, dt=p.dt, end_force=p.end_force,
                             memory=L.MEMORY_DEVICE)
    eng = BroydenSolver(model.block_sizes(blobs_per_swim=nb), memory=L.MEMORY_DEVICE)
    x_blobs, x_bodies, v_bodies, f_blobs = dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * nb))
    step = None
    iters, xs, fbs = [], [], []
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(np.repeat(case["q0"][:, None, :], nseg, axis=1).reshape(-1)))
        model.set_perturbation(4, 12, dev(case["pert"].reshape(-1)))
        step = FM.ElasticStep(model, None, eng, x_blobs, x_bodies, v_bodies, f_blobs, tol=1e-8, max_iter=400,
                              broyden_blobs=nb, blob_radius=1.0)
        for nt in range(STEPS):
            st = step.step(nt)
            print("step", nt, st, flush=True)
            assert st["converged"]
            iters.append(st["broyden_iterations"])
            xs.append(model.read(FM.X))
            fbs.append(f_blobs.cpu().numpy().copy())
    finally:
        if step is not None:
            step.close()
        model.close()
        eng.close()
    ref = case["runs"]
    print("broyden iterations", iters, "oracle", [r["iters"] for r in ref],
          "max |dx|", [float(np.abs(a - r["x"]).max()) for a, r in zip(xs, ref)],
          "f_blobs rel", [rel(a, r["f_blobs"]) for a, r in zip(fbs, ref)], flush=True)
    for s in range(STEPS):
        r = ref[s]
        assert abs(iters[s] - r["iters"]) <= max(2, r["iters"] // 10)
        assert np.abs(xs[s] - r["x"]).max() <= 1e-6
        assert rel(fbs[s], r["f_blobs"]) <= 1e-5
