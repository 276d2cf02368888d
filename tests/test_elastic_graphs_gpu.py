"""Captured Broyden iterations of tfs_elastic_step (tfs_elastic_set_graphs):
with the three handles on one stream, each step replays its iteration and
its end of iteration as hipGraphs, and the results equal the launch-by-launch
loop's bit for bit (positions, blob forces, Broyden counts, error norm) over
two steps, for a body held still with the dense direct blob solve, a free
body with the body-frame direct solve, and the Broyden-for-everything form.
Handles on streams of their own take the launch-by-launch loop."""
import contextlib

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
from timfilsimmod_amd.solver import BLOB_DIRECT, SOLVE_FP64, ElasticVelocities

pytestmark = pytest.mark.gpu

STEPS = 2


def dev(a):
    return torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")


def run(kind, graphs, shared_stream=True):
    body = G.make_body(4, 10, 300, 8.0)
    nfil, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    dl, kb, dt = 2.2, 1800.0, 0.1
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([n], [ex])[0].as_quat(), 1) for n in body.fil_normal])
    qs = np.repeat(q0[:, None, :], nseg, axis=1)
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    # a stream of our own (torch's default stream would read as NULL: each
    # handle would then create one), torch's work on it too
    ts = torch.cuda.Stream() if shared_stream else None
    stream = ts.cuda_stream if ts is not None else None
    with torch.cuda.stream(ts) if ts is not None else contextlib.nullcontext():
        return _run(kind, graphs, body, stream, nfil, nseg, nb, dl, kb, dt, efm, qs, pert)


def _run(kind, graphs, body, stream, nfil, nseg, nb, dl, kb, dt, efm, qs, pert):
    handles = []
    try:
        model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=kb, kt=kb, dt=dt, end_force=efm, memory=L.MEMORY_DEVICE,
                                 stream=stream)
        handles.append(model)
        model.set_state(dev(body.fil_base.reshape(-1)), dev(qs.reshape(-1)))
        model.set_perturbation(4, 8, dev(pert.reshape(-1)))
        f_blobs = dev(np.zeros(3 * nb))
        if kind == "held":
            ev = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                                   memory=L.MEMORY_DEVICE, stream=stream)
            handles.append(ev)
            ev.set_preconditioner(BLOB_DIRECT)
            eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE, stream=stream)
            handles.append(eng)
            keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6))]
            step = FM.ElasticStep(model, ev, eng, *keep, f_blobs, tol=1e-8, max_iter=400)
        elif kind == "free":
            model.set_body(dev(body.centre), dev([1.0, 0.0, 0.0, 0.0]),
                           dev((body.fil_base - body.centre).reshape(-1)), nb, dev(body.blob_refs.reshape(-1)))
            evh = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-12,
                                    free_bodies=True)
            nmob = dev(evh.body_mobility(body.x_blobs, body.centre)[0].reshape(-1))
            evh.close()
            ev = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                                   memory=L.MEMORY_DEVICE, free_bodies=True, stream=stream)
            handles.append(ev)
            ev.set_preconditioner(BLOB_DIRECT)
            eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE, stream=stream)
            handles.append(eng)
            step = FM.ElasticStep(model, ev, eng, None, None, None, f_blobs, tol=1e-8, max_iter=400,
                                  body_mobility=nmob, sphere_radius=float(body.radius))
        else:
            eng = BroydenSolver(model.block_sizes(blobs_per_swim=nb), memory=L.MEMORY_DEVICE, stream=stream)
            handles.append(eng)
            keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6))]
            step = FM.ElasticStep(model, None, eng, *keep, f_blobs, tol=1e-8, max_iter=400, broyden_blobs=nb,
                                  blob_radius=1.0)
        handles.append(step)
        step.set_graphs(graphs)
        stats, xs, fbs = [], [], []
        for nt in range(STEPS):
            st = step.step(nt)
            torch.cuda.synchronize()
            assert st["converged"] == 1
            stats.append(st)
            xs.append(model.read(FM.X))
            fbs.append(f_blobs.cpu().numpy())
        counts = step.iteration_counts()
        return stats, xs, fbs, counts
    finally:
        for h in reversed(handles):
            h.close()


@pytest.mark.parametrize("kind", ["held", "free", "bfe"])
def test_graph_iterations_equal_launch_by_launch(kind):
    print(f"case elastic step, {kind}, 4 x 10 segments + 300 blobs, graphs on against off", flush=True)
    st_g, x_g, fb_g, c_g = run(kind, True)
    st_e, x_e, fb_e, c_e = run(kind, False)
    print("graphs:", [s["broyden_iterations"] for s in st_g], c_g, " launch by launch:",
          [s["broyden_iterations"] for s in st_e], c_e, flush=True)
    assert c_g[0] == sum(s["broyden_iterations"] for s in st_g) and c_g[1] == 0
    assert c_e[0] == 0 and c_e[1] == sum(s["broyden_iterations"] for s in st_e)
    for a, b in zip(st_g, st_e):
        assert a["broyden_iterations"] == b["broyden_iterations"]
        assert a["velocity_solves"] == b["velocity_solves"]
        assert a["error_norm"] == b["error_norm"]
    for a, b in zip(x_g, x_e):
        assert np.array_equal(a, b)
    for a, b in zip(fb_g, fb_e):
        assert np.array_equal(a, b)


def test_separate_streams_take_the_launch_by_launch_loop():
    print("case elastic step, held still, handles on streams of their own", flush=True)
    st, _, _, counts = run("held", True, shared_stream=False)
    assert counts[0] == 0 and counts[1] == sum(s["broyden_iterations"] for s in st)
