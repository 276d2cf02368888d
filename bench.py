"""Benchmark of the RPY mobility hot path on one MI355X.

Workload: the production body -- 640 filaments x 20 segments (12,800
segments, 6 DOF each: force+torque in, velocity+rotation out) and 9,000 blobs
on a sphere of radius 50a -- with seeded synthetic positions and N(0,1)
forces/torques.  Inputs are resident in HBM before timing starts.

One step = one full grand-mobility matvec (all four blocks, dimension
6*12,800 + 3*9,000) in the mixed precision (FP32 pair tensors, FP64
accumulation); the same matvec in plain FP64 is timed beside it, and the
relative L2 error of the mixed result against the FP64 result is reported.

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
PEAK_FP32_VALU_TFLOPS = 157.3    # MI355X vector FP32, published (packed)
PEAK_HBM_GBPS = 8000.0

# Algorithmic flops per pair (canonical far-field count: +,-,* = 1, FMA = 2,
# sqrt/rsqrt/div = 1): tt 36, tr 16, rt 15, rr 21.
FLOPS_SS6, FLOPS_SB6, FLOPS_BS6, FLOPS_BB = 88, 52, 52, 36


def matvec_flops(ns: int, nb: int, seg_dof: int) -> float:
    if seg_dof == 6:
        return FLOPS_SS6 * ns * ns + FLOPS_SB6 * ns * nb + FLOPS_BS6 * nb * ns + FLOPS_BB * nb * nb
    return 36.0 * (ns + nb) ** 2


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
    pair_ms, launches = handle.profile_read()
    handle.profile(False)
    return {"ms_per_step": total_ms / steps, "pair_kernel_ms": pair_ms / max(launches, 1),
            "launches": launches}


def cpu_baseline(body, budget_s: float = 12.0):
    """FP64 oracle direct sum on the host cores (reported, not gating)."""
    from oracle import oracle as O
    O.full_mobility(body.x_segs[:300], body.x_blobs[:300], body.f_segs[:600], body.f_blobs[:300])
    times = []
    t_begin = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        O.full_mobility(body.x_segs, body.x_blobs, body.f_segs, body.f_blobs, seg_dof=6)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_begin > budget_s or len(times) >= 5:
            break
    return float(np.median(times)), len(times), O.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(multi-GPU target sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from timfilsimmod_amd import _lib as L
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.mobility import RpyHandle

    torch.cuda.set_device(0)
    # A dedicated (default-priority) stream: the library's kernels and torch's
    # timing events go on the same queue.
    tstream = torch.cuda.Stream()
    torch.cuda.set_stream(tstream)
    stream = tstream.cuda_stream
    body = G.production_body()
    ns, nb = body.num_segs, body.num_blobs

    # Resident inputs: upload once, evaluate on the device buffers.
    # The full grand mobility: Mbs accumulates onto Mbb (the reference's
    # accumulate build, cuda_functions.cu:739-743).
    h = RpyHandle(ns, nb, seg_dof=6, precision=L.PRECISION_FP64, bs_accumulate=True, stream=stream)
    h.upload(L.SEG_POSITIONS, body.x_segs)
    h.upload(L.BLOB_POSITIONS, body.x_blobs)
    h.upload(L.SEG_FORCES, body.f_segs)
    h.upload(L.BLOB_FORCES, body.f_blobs)
    h.sync()

    fp64 = matvec_leg(torch, h, L.PRECISION_FP64, args.steps, args.warmup)
    v64s, v64b = np.zeros(6 * ns), np.zeros(3 * nb)
    h.download(L.SEG_VELOCITIES, v64s)
    h.download(L.BLOB_VELOCITIES, v64b)
    h.sync()
    mixed = matvec_leg(torch, h, L.PRECISION_FP32_ACC64, args.steps, args.warmup)
    v32s, v32b = np.zeros(6 * ns), np.zeros(3 * nb)
    h.download(L.SEG_VELOCITIES, v32s)
    h.download(L.BLOB_VELOCITIES, v32b)
    h.sync()
    v64 = np.concatenate([v64s, v64b])
    err = float(np.linalg.norm(np.concatenate([v32s, v32b]) - v64) / np.linalg.norm(v64))

    flops = matvec_flops(ns, nb, 6)
    nbytes = matvec_bytes(ns, nb, 6)
    gbps_mixed = nbytes / (mixed["ms_per_step"] * 1e-3) / 1e9
    gbps_fp64 = nbytes / (fp64["ms_per_step"] * 1e-3) / 1e9
    tf_mixed_kernel = flops / (mixed["pair_kernel_ms"] * 1e-3) / 1e12
    tf_fp64_kernel = flops / (fp64["pair_kernel_ms"] * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu_baseline:
        t_cpu, reps, cores = cpu_baseline(body)
        cpu = {"value": nbytes / t_cpu / 1e9, "unit": "GB/s (matvec, compulsory FP64 bytes)",
               "cores": cores, "kind": "port",
               "sample": f"full production-body matvec (6*12800+3*9000), FP64 oracle, OpenMP, "
                         f"median of {reps} after 1 warm-up; {t_cpu * 1e3:.1f} ms per matvec, "
                         f"{flops / t_cpu / 1e9:.1f} GFLOP/s"}

    out = {
        "metric": "timesteps/sec + RPY matvec GB/s at 12.8k segs/9k blobs; mixed vs FP64 @1e-6",
        "value": gbps_mixed,
        "unit": "GB/s (RPY matvec, compulsory FP64 bytes / time)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair tensors, f64 accumulation",
        "data": "synthetic (seeded Fibonacci sphere + straight filaments, N(0,1) forces/torques)",
        "config": {"workload": "single-body RPY grand-mobility matvec, 12,800 segments (6 DOF) + "
                               "9,000 blobs, R=50a, a=1, mu=1",
                   "num_segs": ns, "num_blobs": nb, "matvec_dim": 6 * ns + 3 * nb},
        "matvec": {
            "mixed": {"ms": mixed["ms_per_step"], "GB_s": gbps_mixed,
                      "pair_kernel_ms": mixed["pair_kernel_ms"], "TFLOP_s": tf_mixed_kernel,
                      "rel_l2_err_vs_fp64": err},
            "fp64": {"ms": fp64["ms_per_step"], "GB_s": gbps_fp64,
                     "pair_kernel_ms": fp64["pair_kernel_ms"], "TFLOP_s": tf_fp64_kernel},
            "speedup_mixed_over_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
            "algorithmic_gflop": flops / 1e9,
        },
        "roofline": {"bound": "valu", "achieved": tf_mixed_kernel, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": tf_mixed_kernel / PEAK_FP32_VALU_TFLOPS,
                     "traffic": None,
                     "kernel": "rpy_pairs_kernel<float,6>",
                     "fp64_kernel": {"achieved": tf_fp64_kernel, "peak": PEAK_FP64_VALU_TFLOPS,
                                     "frac": tf_fp64_kernel / PEAK_FP64_VALU_TFLOPS}},
        "cpu_baseline": cpu,
    }
    h.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
