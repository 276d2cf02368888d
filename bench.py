"""Benchmark of the RPY mobility hot path on one MI355X.

Workload (production body): 640 filaments x 20 segments (12,800 segments) and
9,000 surface blobs on a sphere of radius 50a (a = 1, mu = 1), seeded
synthetic geometry, resident in HBM before timing starts.

One step = one prescribed-beat timestep of that body: phase update, beat
kinematics, right-hand side and the warm-started saddle-point solve
(RPY <3,3> mobility + K_phi/K_theta couplings, n = 3*12,800 + 3*9,000 +
2*640 = 66,680) to a relative residual of 1e-8 with GMRES(50) and FP64
Arnoldi.  The initial guess of each solve is the quadratic extrapolation of the
last three solutions (WARM_START); the right preconditioner is the saddle form
with each filament's own mobility block and 16-blob patches (PRECOND).  The headline is the mixed-precision form
(FP32-tile matvecs in the inner GMRES, to INNER_TOL of the outer residual,
inside FP64 iterative refinement); the all-FP64 form runs the same steps from
the same state beside it.

Also timed: the full 6-DOF grand-mobility matvec (dimension 6*12,800 +
3*9,000) in FP64 and FP32-tile/FP64-accumulate forms -- the "RPY matvec GB/s"
part of the metric -- with its relative L2 error against FP64; 200 timesteps
of the production body mixed and FP64 against an FP64 truth solved to 1e-12;
the flow field, the 64-body ensemble (matvec and timestep) and the elastic
Cosserat steps.

Prints one JSON line.  `--gpus` above 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_VALU_TFLOPS = 78.6     # MI355X vector FP64, published
PEAK_FP32_VALU_TFLOPS = 157.3    # MI355X vector FP32, published (packed issue)
PEAK_HBM_GBPS = 8000.0

# Algorithmic flops per pair (canonical far-field count: +,-,* = 1, FMA = 2,
# sqrt/rsqrt/div = 1): tt 36, tr 16, rt 15, rr 21.
FLOPS_TT = 36
FLOPS_SS6, FLOPS_SB6, FLOPS_BS6, FLOPS_BB = 88, 52, 52, 36
DL = 2.2
WARM_START = 2      # tfs_beat_set_warm_start order, both precisions
INNER_TOL = 1e-6    # inner FP32 GMRES: relative reduction of the outer residual
PRECOND = 1         # TFS_PRECOND_BLOCKS (filament self-mobility blocks + blob patches), both precisions


def matvec_flops(ns: int, nb: int, seg_dof: int) -> float:
    if seg_dof == 6:
        return FLOPS_SS6 * ns * ns + FLOPS_SB6 * ns * nb + FLOPS_BS6 * nb * ns + FLOPS_BB * nb * nb
    return float(FLOPS_TT) * (ns + nb) ** 2


def matvec_bytes(ns: int, nb: int, seg_dof: int) -> float:
    """Compulsory FP64 bytes: positions + strengths in, velocities out."""
    per_seg = 3 + 6 + 6 if seg_dof == 6 else 3 + 3 + 3
    return 8.0 * (per_seg * ns + 9 * nb)


def matvec_leg(torch, handle, precision: int, steps: int, warmup: int):
    from timfilsimmod_amd import _lib as L
    handle.set_precision(precision)
    for _ in range(warmup):
        handle.evaluate(L.BLOCK_ALL)
    torch.cuda.synchronize()
    handle.profile(True)
    handle.profile_read()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        handle.evaluate(L.BLOCK_ALL)
    stop.record()
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms, n = handle.profile_read()[precision]
    handle.profile(False)
    return {"ms_per_step": total_ms / steps, "pair_kernel_ms": ms / max(n, 1), "launches": n}


def timestep_leg(torch, body, mode: int, steps: int, warmup: int, stream, warm_start: int = WARM_START,
                 inner_tol: float = INNER_TOL, profiled: int = 5):
    """Warm-up steps, then `steps` timed timesteps; returns timings and state.
    The pair kernel's own time comes from `profiled` further steps with HIP
    events around each of its launches, outside the timed region (the events
    add a few microseconds per launch)."""
    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd.mobility import profile_read
    from timfilsimmod_amd.solver import BeatSimulation
    sim = BeatSimulation(body, mode=mode, tol=1e-8, restart=50, inner_tol=inner_tol, max_cycles=40,
                         max_refinements=20, stream=stream, warm_start=warm_start, precond=PRECOND)
    rpy = sim.rpy_handle()
    for _ in range(warmup):
        sim.step()
    stats = []
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        stats.append(sim.step())
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    state = sim.state()
    # the pair kernel's launches over further steps, each bracketed by events
    L.check(L.lib().tfs_rpy_profile(rpy, 1))
    profile_read(rpy)
    prof_stats = [sim.step() for _ in range(profiled)]
    torch.cuda.synchronize()
    kern = profile_read(rpy)
    L.check(L.lib().tfs_rpy_profile(rpy, 0))
    sim.close()
    return {"ms_per_step": start.elapsed_time(stop) / steps, "wall_s": wall, "stats": stats,
            "kernels": kern, "profiled_stats": prof_stats, "state": state}


def flow_leg(torch, stream):
    """Flow-field velocities of the flow-source body (1,280 x 20 segments +
    9,000 blobs, N(0,1) forces/torques) at a 2048^2 plane through the centre
    and a 256^3 volume grid, Stokeslet+rotlet mode, device-resident points."""
    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.flow import STOKESLET_ROTLET, FlowFieldEvaluator
    body = G.flow_source_body()
    half = body.radius + 1.2 * body.dl * (body.num_seg - 1)
    dev = torch.device("cuda", 0)
    t = torch.linspace(-half, half, 2048, dtype=torch.float64, device=dev)
    gy, gx = torch.meshgrid(t, t, indexing="ij")
    plane = torch.stack([gx.reshape(-1), gy.reshape(-1), torch.zeros_like(gx).reshape(-1)], 1).contiguous()
    t3 = torch.linspace(-half, half, 256, dtype=torch.float64, device=dev)
    gz, gy3, gx3 = torch.meshgrid(t3, t3, t3, indexing="ij")
    vol = torch.stack([gx3.reshape(-1), gy3.reshape(-1), gz.reshape(-1)], 1).contiguous()
    del gz, gy3, gx3, gx, gy
    srcs = [torch.tensor(a, dtype=torch.float64, device=dev) for a in
            (body.x_segs, body.f_segs, body.x_blobs, body.f_blobs, body.centre)]
    flops_per_target = 44.0 * body.num_segs + 28.0 * body.num_blobs
    out = {"sources": {"segments": body.num_segs, "blobs": body.num_blobs},
           "algorithmic_flop_per_target": flops_per_target}
    ref_plane = None
    for prec in ("fp64", "fp32"):
        ev = FlowFieldEvaluator(1, body.num_segs, body.num_blobs, mode=STOKESLET_ROTLET, precision=prec,
                                memory=L.MEMORY_DEVICE, stream=stream)
        ev.set_sources(*srcs)
        for name, pts in (("plane_2048x2048", plane), ("volume_256x256x256", vol)):
            v = torch.empty_like(pts)
            small = pts[:4096].contiguous()
            ev.velocities(small, torch.empty_like(small))      # warm-up
            ev.velocities(pts, v)
            ms = ev.last_kernel_ms()
            n = pts.shape[0]
            entry = {"points": n, "kernel_ms": ms, "points_per_s": n / (ms * 1e-3),
                     "TFLOP_s": n * flops_per_target / (ms * 1e-3) / 1e12}
            if name.startswith("plane"):
                if prec == "fp64":
                    ref_plane = v.clone()
                else:
                    entry["rel_l2_err_vs_fp64"] = float(torch.linalg.norm(v - ref_plane) /
                                                        torch.linalg.norm(ref_plane))
            out[f"{name}_{prec}"] = entry
            del v
        ev.close()
    return out


def ensemble_leg(torch, stream, steps: int = 5):
    """64 independent bodies of uneven size (filaments U[100,1000] x 20 segments,
    blobs U[2000,9000]) in one block-diagonal 6-DOF matvec launch: autotuned
    shape, hipGraph replays; per-body completion p50/p99 (ms from the launch's
    first item) read from the stamps of the last timed replay."""
    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd.ensemble import make_ensemble
    from timfilsimmod_amd.mobility import Graph, RpyHandle
    e = make_ensemble(64)
    flops = 0.0
    for b in e.bodies:
        flops += matvec_flops(b.num_segs, b.num_blobs, 6)
    out = {"bodies": 64, "segments": e.num_segs, "blobs": e.num_blobs,
           "algorithmic_tflop_per_launch": flops / 1e12}
    h = RpyHandle(e.num_segs, e.num_blobs, seg_dof=6, precision=L.PRECISION_FP64, bs_accumulate=True,
                  stream=stream)
    h.set_bodies(e.segs_per_body, e.blobs_per_body)
    for what, arr in ((L.SEG_POSITIONS, e.x_segs), (L.BLOB_POSITIONS, e.x_blobs), (L.SEG_FORCES, e.f_segs),
                      (L.BLOB_FORCES, e.f_blobs)):
        h.upload(what, arr)
    h.sync()
    for prec, name in ((L.PRECISION_FP64, "fp64"), (L.PRECISION_FP32_ACC64, "mixed")):
        tuned_ms = h.autotune(prec, reps=1)
        shape = h.tile_shape(prec)
        h.set_precision(prec)
        h.body_timing(True)     # every replay stores per-item start/finish stamps
        g = Graph(h, L.BLOCK_ALL, 1)
        g.launch()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            g.launch()
        t1.record()
        torch.cuda.synchronize()
        bt = h.body_times(64)   # the last timed replay
        g.close()
        h.body_timing(False)
        ms = t0.elapsed_time(t1) / steps
        out[name] = {"ms_per_launch": ms, "body_matvecs_per_s": 64 * 1000.0 / ms,
                     "TFLOP_s": flops / (ms * 1e-3) / 1e12, "autotuned_shape": shape,
                     "autotune_best_ms": tuned_ms,
                     "per_body_completion_ms_p50": float(np.percentile(bt, 50)),
                     "per_body_completion_ms_p99": float(np.percentile(bt, 99)),
                     "per_body_completion_ms_max": float(bt.max()),
                     "per_body_completion_p99_over_p50": float(np.percentile(bt, 99) / np.percentile(bt, 50))}
    h.close()
    return out


def trajectory_leg(torch, body, stream, steps: int = 200, every: int = 20):
    """The production body over `steps` prescribed-beat timesteps three ways:
    mixed precision and FP64 at the headline tolerance 1e-8, and an FP64
    truth at 1e-12.  Reports each one's largest deviation from the truth over
    states taken every `every` steps (segment positions / DL, solution
    relative L2), and their step rates."""
    from timfilsimmod_amd.solver import SOLVE_FP64, SOLVE_MIXED, BeatSimulation
    runs = {}
    for name, mode, tol in (("truth", SOLVE_FP64, 1e-12), ("fp64", SOLVE_FP64, 1e-8), ("mixed", SOLVE_MIXED, 1e-8)):
        print(f"bench: trajectory leg, {name} ({steps} steps, tol {tol})", file=sys.stderr, flush=True)
        sim = BeatSimulation(body, mode=mode, tol=tol, restart=50, inner_tol=INNER_TOL, max_cycles=60,
                             max_refinements=20, stream=stream, warm_start=WARM_START, precond=PRECOND)
        stats, states = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            stats.append(sim.step())
            if k % every == every - 1:
                states.append(sim.state())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        sim.close()
        runs[name] = (stats, states, wall)
    out = {"steps": steps, "states_compared": steps // every}
    truth = runs["truth"][1]
    for name, (stats, states, wall) in runs.items():
        entry = {"tol": 1e-12 if name == "truth" else 1e-8, "timesteps_per_s_wall": steps / wall,
                 "gmres_iterations_mean": float(np.mean([s["iterations"] for s in stats])),
                 "converged_all": bool(all(s["converged"] for s in stats))}
        if name != "truth":
            entry["x_segs_dev_over_DL_max"] = float(max(np.abs(a["x_segs"] - b["x_segs"]).max()
                                                        for a, b in zip(states, truth)) / DL)
            entry["solution_rel_err_max"] = float(max(np.linalg.norm(a["solution"] - b["solution"]) /
                                                      np.linalg.norm(b["solution"]) for a, b in zip(states, truth)))
        out[name] = entry
    return out


def ensemble_timestep_leg(torch, stream, steps: int = 2, warmup: int = 2):
    """The 64 uneven bodies (filaments U[100,1000] x 20 segments, blobs
    U[2000,9000]) taking prescribed-beat timesteps together on one handle:
    one lock-step GMRES per body in every launch, each GMRES cycle a captured
    hipGraph, converged bodies dropping out of the pair launches.  Same solver
    settings as the single-body timestep.  Reports aggregate body-timesteps/s
    and each body's solve time (device clock, solve start to the body's last
    Arnoldi stop) over the timed steps, mixed and FP64, and the bodies'
    mixed-vs-FP64 trajectory deviation after the same steps."""
    from timfilsimmod_amd.ensemble import make_ensemble
    from timfilsimmod_amd.solver import SOLVE_FP64, SOLVE_MIXED, EnsembleBeat
    e = make_ensemble(64)
    out = {"bodies": 64, "segments": e.num_segs, "blobs": e.num_blobs, "warmup_steps": warmup,
           "timed_steps": steps}
    states = {}
    for mode, name in ((SOLVE_MIXED, "mixed"), (SOLVE_FP64, "fp64")):
        print(f"bench: ensemble timestep leg, {name}", file=sys.stderr, flush=True)
        sim = EnsembleBeat(e.bodies, mode=mode, tol=1e-8, restart=50, inner_tol=INNER_TOL, max_cycles=40,
                           max_refinements=20, stream=stream, warm_start=WARM_START, precond=PRECOND)
        out["system_size"] = sim.n
        for _ in range(warmup):
            sim.step()
        torch.cuda.synchronize()
        stats, solve_ms, iters, relres = [], [], [], []
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            stats.append(sim.step())
            bs = sim.body_stats()
            solve_ms.append(bs["solve_ms"])
            iters.append(bs["iterations"])
            relres.append(bs["relres"].max())
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / steps
        bs = sim.body_stats()
        solve_ms = np.concatenate(solve_ms)
        out[name] = {"ms_per_step": ms, "ensemble_timesteps_per_s": 1000.0 / ms,
                     "body_timesteps_per_s": 64 * 1000.0 / ms,
                     "per_body_solve_ms_p50": float(np.percentile(solve_ms, 50)),
                     "per_body_solve_ms_p99": float(np.percentile(solve_ms, 99)),
                     "per_body_solve_ms_max": float(solve_ms.max()),
                     "gmres_iterations_per_body_mean": float(np.mean(iters)),
                     "gmres_iterations_per_body_max": int(np.max(iters)),
                     "lockstep_iterations_per_step": float(np.mean([s["iterations"] for s in stats])),
                     "fp32_matvecs_per_step": float(np.mean([s["fp32_matvecs"] for s in stats])),
                     "fp64_matvecs_per_step": float(np.mean([s["fp64_matvecs"] for s in stats])),
                     "final_relres_max": float(max(relres)),
                     "converged_all": bool(all(s["converged"] for s in stats)),
                     "graph_cycles": bs["graph_cycles"], "eager_cycles": bs["eager_cycles"]}
        states[name] = sim.split(sim.state())
        sim.close()
        torch.cuda.empty_cache()
    out["speedup_mixed_over_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    out["trajectory_dev_over_DL_max"] = float(max(np.abs(a["x_segs"] - b["x_segs"]).max()
                                                  for a, b in zip(states["mixed"], states["fp64"])) / DL)
    return out


def elastic_leg(torch, stream, steps: int = 3):
    """The small elastic configuration on the device: one rigid sphere of
    radius 20a with 2,000 blobs and 16 filaments x 20 segments of the
    reference's Cosserat model (clamped, follower end force of dimensionless
    strength 220, a first-step kick), each timestep through tfs_elastic_step:
    blob forces from the dense inverse of the blob mobility, Broyden to 1e-8.
    Every handle of a step runs on the bench's stream."""
    from scipy.spatial.transform import Rotation

    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd import filament as FM
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.broyden import BroydenSolver
    from timfilsimmod_amd.solver import BLOB_DIRECT, SOLVE_FP64, ElasticVelocities
    body = G.small_body()
    nfil, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    dl, kb, dt = 2.2, 1800.0, 0.1
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([n], [ex])[0].as_quat(), 1) for n in body.fil_normal])
    qs = np.repeat(q0[:, None, :], nseg, axis=1)
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=kb, kt=kb, dt=dt, end_force=efm, memory=L.MEMORY_DEVICE,
                             stream=stream)
    ev = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                           memory=L.MEMORY_DEVICE, stream=stream)
    ev.set_preconditioner(BLOB_DIRECT)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE, stream=stream)
    keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * nb))]
    model.set_state(dev(body.fil_base.reshape(-1)), dev(qs.reshape(-1)))
    model.set_perturbation(4, 12, dev(pert.reshape(-1)))
    step = FM.ElasticStep(model, ev, eng, *keep, tol=1e-8, max_iter=400)
    per = []
    for nt in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = step.step(nt)
        torch.cuda.synchronize()
        per.append({"nt": nt, "ms": (time.perf_counter() - t0) * 1e3, "broyden_iterations": st["broyden_iterations"],
                    "gmres_iterations": st["gmres_iterations"]})
    step.close()
    model.close()
    ev.close()
    eng.close()
    out = {"workload": "16 filaments x 20 segments (Cosserat, clamped, end force 220) + 2,000 blobs, "
                       "dt 0.1, Broyden to 1e-8, blob forces from the dense inverse of the blob "
                       "mobility (formed once, blocked Gauss-Jordan on the device)", "steps": per,
           "ms_per_step_mean": float(np.mean([p["ms"] for p in per]))}

    def timed(step_obj):
        rows = []
        for nt in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = step_obj.step(nt)
            torch.cuda.synchronize()
            rows.append({"nt": nt, "ms": (time.perf_counter() - t0) * 1e3,
                         "broyden_iterations": st["broyden_iterations"], "gmres_iterations": st["gmres_iterations"]})
        return rows

    # the same body swimming freely: body state on the device, the blob-only
    # solve with body rows from the stored body-frame inverse, J0 with the
    # body's Schur complement
    fil_refs = dev((body.fil_base - body.centre).reshape(-1))
    model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=kb, kt=kb, dt=dt, end_force=efm, memory=L.MEMORY_DEVICE,
                             stream=stream)
    model.set_state(dev(body.fil_base.reshape(-1)), dev(qs.reshape(-1)))
    model.set_body(dev(body.centre), dev([1.0, 0.0, 0.0, 0.0]), fil_refs, nb, dev(body.blob_refs.reshape(-1)))
    model.set_perturbation(4, 12, dev(pert.reshape(-1)))
    evh = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=60, max_cycles=100, tol=1e-12,
                            free_bodies=True)
    nmob = dev(evh.body_mobility(body.x_blobs, body.centre)[0].reshape(-1))
    evh.close()
    ev = ElasticVelocities(1, nfil * nseg, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                           memory=L.MEMORY_DEVICE, free_bodies=True, stream=stream)
    ev.set_preconditioner(BLOB_DIRECT)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE, stream=stream)
    fb = dev(np.zeros(3 * nb))
    step = FM.ElasticStep(model, ev, eng, None, None, None, fb, tol=1e-8, max_iter=400, body_mobility=nmob,
                          sphere_radius=float(body.radius))
    rows = timed(step)
    step.close()
    model.close()
    ev.close()
    eng.close()
    out["free_body"] = {"workload": "the same body swimming freely (blob forces and body velocity from the "
                                    "body-frame dense inverse, J0 with the body's Schur complement)", "steps": rows,
                        "ms_per_step_mean": float(np.mean([r["ms"] for r in rows]))}
    # the body held still with the blob forces as Broyden unknowns
    model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=kb, kt=kb, dt=dt, end_force=efm, memory=L.MEMORY_DEVICE,
                             stream=stream)
    model.set_state(dev(body.fil_base.reshape(-1)), dev(qs.reshape(-1)))
    model.set_perturbation(4, 12, dev(pert.reshape(-1)))
    eng = BroydenSolver(model.block_sizes(blobs_per_swim=nb), memory=L.MEMORY_DEVICE, stream=stream)
    keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * nb))]
    step = FM.ElasticStep(model, None, eng, *keep, tol=1e-8, max_iter=400, broyden_blobs=nb, blob_radius=1.0)
    rows = timed(step)
    step.close()
    model.close()
    eng.close()
    out["broyden_for_everything"] = {"workload": "the body held still, blob forces as Broyden unknowns "
                                                 "(one plain matvec per evaluation)", "steps": rows,
                                     "ms_per_step_mean": float(np.mean([r["ms"] for r in rows]))}
    return out


def elastic_cpu_baseline():
    """The first elastic step of the small body on the host cores: the same
    loop on the FP64 oracle (C block sums with OpenMP, an LU of the blob
    mobility factored once, the numpy Cosserat restatement, the numpy Broyden
    restatement).  Reported, not gating."""
    import scipy.linalg as sla
    from scipy.spatial.transform import Rotation

    from oracle import broyden_oracle as BO
    from oracle import filament_oracle as F
    from oracle import oracle as O
    from timfilsimmod_amd import geometry as G
    body = G.small_body()
    nfil, nseg, nb = body.num_fil, body.num_seg, body.num_blobs
    dl, kb, dt, tol = 2.2, 1800.0, 0.1, 1e-8
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    p = F.Params(nseg=nseg, dl=dl, kb=kb, kt=kb, dt=dt, end_force=efm)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([n], [ex])[0].as_quat(), 1) for n in body.fil_normal])
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    fils = [F.Filament.at_rest(body.fil_base[i], np.tile(q0[i], (nseg, 1)), nseg, dl) for i in range(nfil)]
    t0 = time.perf_counter()
    lu = sla.lu_factor(O.dense_mobility(np.zeros(0), body.x_blobs))
    z3, z6 = np.zeros(3), np.zeros(6)
    cat = lambda name: np.concatenate([getattr(f, name).reshape(-1) for f in fils])

    def evaluate():
        f = np.concatenate([F.internal_forces(fl, p, {4: pert[i, :3], 12: pert[i, 3:]}).reshape(-1)
                            for i, fl in enumerate(fils)])
        x = cat("x")
        vb = O.block("bs", x, body.x_blobs, f, np.zeros(3 * nb), seg_dof=6, bs_accumulate=False)
        lam = sla.lu_solve(lu, -vb)
        v = O.block("sb", x, body.x_blobs, f, lam, O.block("ss", x, body.x_blobs, f, lam, seg_dof=6), seg_dof=6)
        st = {"seg_x": x, "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"), "seg_um1": cat("um1"),
              "v_segs": v, "body_x": z3, "body_xm1": z3, "body_xm2": z3, "body_u": z3, "body_um1": z3,
              "v_bodies": z6}
        e, big = BO.compute_errors(st, 1, nfil, nseg, False, dt, dl, tol)
        return f, e, big

    for fl in fils:
        F.initial_guess(fl, 0, p)
    f, e, big = evaluate()
    inv = [np.linalg.inv(F.approx_jacobian(fl, f.reshape(nfil, nseg, 6)[i], False, p))
           for i, fl in enumerate(fils)] + [np.eye(6)]
    ref = BO.BroydenRef(6 * (nfil * nseg + 1), 400, 0.1, inv)
    ref.start(e)
    while big and ref.iter < 400:
        du = ref.find_update()
        for i, fl in enumerate(fils):
            F.update(fl, du[6 * nseg * i:6 * nseg * (i + 1)], p)
        f, e, big = evaluate()
        ref.end_of_iter(e, big)
    secs = time.perf_counter() - t0
    return {"value": 1.0 / secs, "unit": "timesteps/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"the first elastic step (nt = 0, {ref.iter} Broyden iterations) of the same workload on "
                      f"the oracle: {secs:.2f} s including the blob mobility's LU"}


def copy_leg(torch, nbytes: int = 256 << 20, reps: int = 3):
    """A plain device-to-device copy of `nbytes` (torch's copy kernel), `reps`
    times: the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    dev = torch.device("cuda", 0)
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB_s": 2 * nbytes / (ms * 1e-3) / 1e9}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_rpy_pairs_sym_production_body.md")


def measured_traffic():
    """HBM bytes per launch of the FP32 <3,3> pair kernel (the symmetric form), as recorded from the
    two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; scaled by the copy
    calibration) in the committed profile summary; None if it is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(body, budget_s: float = 15.0):
    """FP64 oracle direct sum on the host cores (reported, not gating): the
    prescribed-beat <3,3> matvec of the production body, timed."""
    from oracle import oracle as O
    O.full_mobility(body.x_segs[:300], body.x_blobs[:300], body.f_segs[:600], body.f_blobs[:300])
    times = []
    t_begin = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        O.full_mobility(body.x_segs, body.x_blobs, body.f_segs, body.f_blobs, seg_dof=3,
                        bs_accumulate=True)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_begin > budget_s or len(times) >= 5:
            break
    return float(np.median(times)), len(times), O.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed timesteps per precision")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--matvec-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", default="timestep,trajectory,matvec,flow,ensemble,ensemble_step,elastic",
                    help="comma list of timestep, trajectory, matvec, flow, ensemble, ensemble_step, elastic, copy "
                         "(copy: a 256 MiB device copy for calibrating HBM byte counters; not part of the "
                         "default run)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    unknown = legs - {"timestep", "trajectory", "matvec", "flow", "ensemble", "ensemble_step", "copy", "elastic"}
    if unknown:
        print(f"bench.py: unknown legs {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: the library runs on one GPU "
              "(multi-GPU target sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.mobility import RpyHandle
    from timfilsimmod_amd.solver import SOLVE_FP64, SOLVE_MIXED

    torch.cuda.set_device(0)
    # A dedicated (default-priority) stream: the library's kernels and torch's
    # timing events go on the same queue.
    tstream = torch.cuda.Stream()
    torch.cuda.set_stream(tstream)
    stream = tstream.cuda_stream
    body = G.production_body()
    ns, nb = body.num_segs, body.num_blobs

    copy_cal = None
    if "copy" in legs:
        copy_cal = copy_leg(torch)

    # ---- timestep legs (headline) ------------------------------------------
    mixed = fp64 = None
    ts_mixed = ts_fp64 = dev = sol_err = nsys = None
    if "timestep" in legs:
        print("bench: timestep leg, mixed precision", file=sys.stderr, flush=True)
        mixed = timestep_leg(torch, body, SOLVE_MIXED, args.steps, args.warmup, stream)
        print("bench: timestep leg, FP64", file=sys.stderr, flush=True)
        fp64 = timestep_leg(torch, body, SOLVE_FP64, args.steps, args.warmup, stream)
        ts_mixed = 1000.0 / mixed["ms_per_step"]
        ts_fp64 = 1000.0 / fp64["ms_per_step"]
        dev = float(np.abs(mixed["state"]["x_segs"] - fp64["state"]["x_segs"]).max() / DL)
        sol_err = float(np.linalg.norm(mixed["state"]["solution"] - fp64["state"]["solution"]) /
                        np.linalg.norm(fp64["state"]["solution"]))
        nsys = mixed["state"]["solution"].size

    def summ(leg):
        st = leg["stats"]
        return {"ms_per_step": leg["ms_per_step"],
                "timesteps_per_s": 1000.0 / leg["ms_per_step"],
                "gmres_iterations_mean": float(np.mean([s["iterations"] for s in st])),
                "fp32_matvecs_per_step": float(np.mean([s["fp32_matvecs"] for s in st])),
                "fp64_matvecs_per_step": float(np.mean([s["fp64_matvecs"] for s in st])),
                "refinements_mean": float(np.mean([s["refinements"] for s in st])),
                "final_relres_max": float(max(s["relres"] for s in st)),
                "converged_all": bool(all(s["converged"] for s in st))}

    # Dominant kernel of the headline: the FP32-tile <3,3> pair kernel, timed
    # over the profiled steps that follow the timed ones.  The summed event
    # time covers every launch, including the few the solver enqueued past
    # convergence (they exit at once); the work is counted only for the
    # matvecs the solver used (its own counters).
    flops3 = matvec_flops(ns, nb, 3)
    k32_ms = k32_n = k64_ms = k64_n = 0
    used32 = used64 = 0
    if mixed is not None:
        k32_ms, k32_n = mixed["kernels"][L.PRECISION_FP32_ACC64]
        k64_ms, k64_n = fp64["kernels"][L.PRECISION_FP64]
        used32 = int(sum(s["fp32_matvecs"] for s in mixed["profiled_stats"]))
        used64 = int(sum(s["fp64_matvecs"] for s in fp64["profiled_stats"]))
    tf32 = flops3 * used32 / (k32_ms * 1e-3) / 1e12 if used32 else 0.0
    tf64 = flops3 * used64 / (k64_ms * 1e-3) / 1e12 if used64 else 0.0

    # ---- 6-DOF matvec legs ---------------------------------------------------
    flops6 = matvec_flops(ns, nb, 6)
    bytes6 = matvec_bytes(ns, nb, 6)
    matvec = None
    if "matvec" in legs:
        print("bench: 6-DOF matvec legs", file=sys.stderr, flush=True)
        h = RpyHandle(ns, nb, seg_dof=6, precision=L.PRECISION_FP64, bs_accumulate=True, stream=stream)
        h.upload(L.SEG_POSITIONS, body.x_segs)
        h.upload(L.BLOB_POSITIONS, body.x_blobs)
        h.upload(L.SEG_FORCES, body.f_segs)
        h.upload(L.BLOB_FORCES, body.f_blobs)
        h.sync()
        mv64 = matvec_leg(torch, h, L.PRECISION_FP64, args.matvec_steps, 3)
        v64s, v64b = np.zeros(6 * ns), np.zeros(3 * nb)
        h.download(L.SEG_VELOCITIES, v64s)
        h.download(L.BLOB_VELOCITIES, v64b)
        h.sync()
        mv32 = matvec_leg(torch, h, L.PRECISION_FP32_ACC64, args.matvec_steps, 3)
        v32s, v32b = np.zeros(6 * ns), np.zeros(3 * nb)
        h.download(L.SEG_VELOCITIES, v32s)
        h.download(L.BLOB_VELOCITIES, v32b)
        h.sync()
        h.close()
        v64 = np.concatenate([v64s, v64b])
        mv_err = float(np.linalg.norm(np.concatenate([v32s, v32b]) - v64) / np.linalg.norm(v64))

        def mv(leg, peak):
            t = leg["pair_kernel_ms"] * 1e-3
            return {"ms": leg["ms_per_step"], "GB_s": bytes6 / (leg["ms_per_step"] * 1e-3) / 1e9,
                    "pair_kernel_ms": leg["pair_kernel_ms"], "TFLOP_s": flops6 / t / 1e12,
                    "valu_frac": flops6 / t / 1e12 / peak}

        matvec = {"workload": "6-DOF grand mobility, dim 6*12800+3*9000",
                  "fp64": mv(mv64, PEAK_FP64_VALU_TFLOPS), "mixed": mv(mv32, PEAK_FP32_VALU_TFLOPS),
                  "mixed_rel_l2_err_vs_fp64": mv_err,
                  "algorithmic_gflop": flops6 / 1e9, "compulsory_bytes": bytes6}

    flow = ens = None
    traj = None
    if "trajectory" in legs:
        traj = trajectory_leg(torch, body, stream)
        torch.cuda.empty_cache()
    if "flow" in legs:
        print("bench: flow-field leg", file=sys.stderr, flush=True)
        flow = flow_leg(torch, stream)
        torch.cuda.empty_cache()
    if "ensemble" in legs:
        print("bench: ensemble leg", file=sys.stderr, flush=True)
        ens = ensemble_leg(torch, stream)
    ens_step = None
    if "ensemble_step" in legs:
        ens_step = ensemble_timestep_leg(torch, stream)
        torch.cuda.empty_cache()
    elastic = None
    if "elastic" in legs:
        print("bench: elastic-step leg", file=sys.stderr, flush=True)
        elastic = elastic_leg(torch, stream)
        if not args.no_cpu_baseline:
            print("bench: elastic-step CPU baseline", file=sys.stderr, flush=True)
            elastic["cpu_baseline"] = elastic_cpu_baseline()

    cpu = None
    if not args.no_cpu_baseline and fp64 is not None:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        per_step = summ(fp64)["fp64_matvecs_per_step"]
        t_cpu, reps, cores = cpu_baseline(body)
        cpu = {"value": 1.0 / (t_cpu * per_step), "unit": "timesteps/s", "cores": cores, "kind": "port",
               "sample": (f"FP64 oracle direct sum (OpenMP) of the production body's <3,3> matvec "
                          f"(21,800 particles), median of {reps} after 1 warm-up: {t_cpu * 1e3:.1f} ms "
                          f"({flops3 / t_cpu / 1e9:.1f} GFLOP/s); a CPU timestep is counted as "
                          f"{per_step:.1f} such matvecs, the FP64 GPU solve's mean per step")}

    out = {
        "metric": "timesteps/sec + RPY matvec GB/s at 12.8k segs/9k blobs; mixed vs FP64 @1e-6",
        "value": ts_mixed,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None if mixed is None else mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair tensors with f64 accumulation inside f64 iterative refinement",
        "data": "synthetic (seeded Fibonacci sphere + filaments, Fulford-Blake prescribed beat)",
        "config": {"workload": "prescribed-beat timestep, one body: 640 filaments x 20 segments + "
                               "9,000 blobs, R=50a, dt=T/300, GMRES(50) to 1e-8, dynamic phase "
                               "and shape rotation",
                   "warm_start": "quadratic extrapolation of the last three solutions",
                   "preconditioner": "saddle form with filament self-mobility blocks and 16-blob patches",
                   "inner_tol": INNER_TOL,
                   "num_segs": ns, "num_blobs": nb, "system_size": nsys},
        "timestep": None if mixed is None else {
                     "mixed": summ(mixed), "fp64": summ(fp64),
                     "speedup_mixed_over_fp64": ts_mixed / ts_fp64,
                     "trajectory_dev_over_DL_after_steps": dev,
                     "solution_rel_err_mixed_vs_fp64": sol_err,
                     "warmup_plus_timed_steps": args.warmup + args.steps},
        "trajectory_200": traj,
        "matvec": matvec,
        "flow_field": flow,
        "ensemble": ens,
        "ensemble_timestep": ens_step,
        "elastic_step": elastic,
        "roofline": {"bound": "valu", "achieved": tf32, "peak": PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": tf32 / PEAK_FP32_VALU_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "rpy_pairs_sym_kernel, FP32 <3,3> symmetric form (timestep matvec, 17.11 GFLOP "
                               "of ordered-pair work per launch, each unordered pair evaluated once)",
                     "launches": k32_n, "matvecs_used": used32, "avg_ms": k32_ms / max(used32, 1),
                     "fp64_kernel": {"achieved": tf64, "peak": PEAK_FP64_VALU_TFLOPS,
                                     "frac": tf64 / PEAK_FP64_VALU_TFLOPS,
                                     "avg_ms": k64_ms / max(used64, 1), "launches": k64_n,
                                     "matvecs_used": used64}},
        "cpu_baseline": cpu,
    }
    if copy_cal is not None:
        out["copy_calibration"] = copy_cal
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
