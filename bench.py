#!/usr/bin/env python3
"""Benchmark: bit-exact biome samples/s, batched vs plain per-call, on one MI355X.

Workload (one step): 4096 seeds from splitmix64(1); each seed a 64x64 quart grid at quart-y
{-16, 0, 16, 48} with its origin uniform in +-2^22 blocks; Overworld 26.3 (float noise path) and
26.1 (double noise path) on identical inputs.  A step is one batched plan run per version: the seeds'
noise state is rebuilt and every sample evaluated (nothing cached between steps).

Prints ONE JSON line.  Besides the contract keys it reports the plain per-call rate (one synchronous
request per seed, like a per-call GPU library), the dominant kernel's roofline figures measured with
HIP events on its own stream, and the reference CPU engine timed on this host (cpu_baseline).
"""
import argparse
import json
import os
import sys
import threading
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "mc-worldgen-seed-lab_amd")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

N_SEEDS, SIDE, QYS, RANGE = 4096, 64, (-16, 0, 16, 48), 1 << 22
VERSIONS = ("26.3", "26.1")

# Algorithmic work per unique column and per sample, counting IEEE add/sub/mul/div the kernel must
# execute (conversions, compares, floors and integer/tree work excluded), derived in DESIGN.md.
# Per 2-D octave evaluation: 8 gradient dots (3 mul + 2 add) = 40, 2 corner offsets (r-1), 2 fades (7
# each), 7 lerps (3 each), amplitude mul + accumulate = 79 in the noise precision; coordinates
# (x*f, z*f, +offset x2, -floor x2) = 6 double.  52 evaluations per column; per level pair one f*1.018;
# column glue (shift *4, ridges folded, spline segment, offset) ~40 float.
FLOPS = {
    "26.3": {"f32": 52 * 79 + 40, "f64": 52 * 6 + 26 + 2},
    "26.1": {"f32": 0, "f64": 52 * 79 + 52 * 6 + 26 + 2 + 40},
}
PER_SAMPLE = {"26.3": {"f32": 3}, "26.1": {"f64": 3}}
PEAK = {"f32": 157.3, "f64": 78.6}     # MI355X published vector peaks, TFLOP/s (FMA counted as 2)


def splitmix64_stream(k, n):
    mask = (1 << 64) - 1
    st = k & mask
    out = []
    for _ in range(n):
        st = (st + 0x9E3779B97F4A7C15) & mask
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out


def signed64(v):
    return v - (1 << 64) if v >= (1 << 63) else v


def workload():
    """(seed, qx0, qz0) per seed: draws 3i, 3i+1, 3i+2 of splitmix64(1); origin = (draw mod (2R+1)) - R blocks."""
    d = splitmix64_stream(1, 3 * N_SEEDS)
    out = []
    for i in range(N_SEEDS):
        seed = signed64(d[3 * i])
        bx0 = d[3 * i + 1] % (2 * RANGE + 1) - RANGE
        bz0 = d[3 * i + 2] % (2 * RANGE + 1) - RANGE
        out.append((seed, bx0 >> 2, bz0 >> 2))
    return out


def cpu_baseline(work, budget_s=12.0):
    """The reference CPU engine (engine/cuda/mcg_host.c built from its sources into oracle/_ref) on a
    bounded sample of the same workload: seeds in order, each its own 64x64x4 quart points, one seed
    per mcg_ref_points call, calls spread over host threads (parallel over seeds, as the reference's
    OpenMP loop); falls back to the oracle restatement where oracle/_ref is absent."""
    import ctypes as C
    import oracle
    kind = "reference" if oracle.RefEngine.available() else "port"
    ncores = min(16, os.cpu_count() or 1)
    try:
        gomp = C.CDLL("libgomp.so.1")
    except OSError:
        gomp = None
    eng = oracle.RefEngine("26.3", "overworld") if kind == "reference" else oracle.Oracle("26.3", "overworld")
    xs, zs = np.meshgrid(np.arange(SIDE), np.arange(SIDE))
    lock = threading.Lock()
    state = {"next": 0, "done": 0, "stop": False}
    t0 = time.perf_counter()

    def worker():
        if gomp is not None:
            gomp.omp_set_num_threads(1)
        while True:
            with lock:
                if state["stop"] or state["next"] >= len(work):
                    return
                i = state["next"]
                state["next"] += 1
            seed, qx0, qz0 = work[i]
            pts = np.concatenate([np.stack([(qx0 + xs).ravel(), np.full(SIDE * SIDE, qy), (qz0 + zs).ravel()], 1)
                                  for qy in QYS]).astype(np.int32)
            eng.points([seed], pts)
            with lock:
                state["done"] += 1
                if time.perf_counter() - t0 > budget_s:
                    state["stop"] = True

    th = [threading.Thread(target=worker) for _ in range(ncores)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    samples = state["done"] * SIDE * SIDE * len(QYS)
    return {"value": samples / dt, "unit": "samples/s", "cores": ncores, "kind": kind,
            "sample": "%d of the %d seeds (each 64x64 quarts x 4 y, its own origin), Overworld 26.3, %.1f s wall, "
                      "one seed per mcg_ref_points call over %d host threads" % (state["done"], N_SEEDS, dt, ncores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tile-w", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--percall-seeds", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import mcbiome
    assert torch.cuda.is_available(), "bench needs a HIP device"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    work = workload()
    regions = [mcbiome.make_region(s, qx0, qz0, SIDE, SIDE, QYS, False, i * SIDE * SIDE * len(QYS))
               for i, (s, qx0, qz0) in enumerate(work)]
    engines, plans, outs = {}, {}, {}
    tuned = {}
    for v in VERSIONS:
        engines[v] = mcbiome.Engine(v, "overworld")
        plans[v] = engines[v].plan(regions, tile_w=args.tile_w)
        outs[v] = torch.empty(plans[v].out_bytes, dtype=torch.uint8, device="cuda")
        if not args.tile_w:      # project autotuner: tile shape, columns per lane, persistent workgroups
            tw, ms = plans[v].autotune(outs[v].data_ptr(), sp)
            tuned[v] = {"tile_w": tw & 0xFFFF, "cols_per_lane": tw >> 16, "grid_ms": round(ms, 3)}
    samples_per_step = sum(plans[v].samples for v in VERSIONS)

    def step():
        for v in VERSIONS:
            plans[v].run(outs[v].data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = samples_per_step * args.steps / dt

    # dominant kernel: the quart-grid launch (stage 1) per version, timed with events on its stream
    kern = {}
    for v in VERSIONS:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        plans[v].run_stage(0, outs[v].data_ptr(), sp)
        reps = 3
        e0.record(stream)
        for _ in range(reps):
            plans[v].run_stage(1, outs[v].data_ptr(), sp)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        cols, samp = plans[v].columns, plans[v].samples
        f32 = FLOPS[v]["f32"] * cols + PER_SAMPLE[v].get("f32", 0) * samp
        f64 = FLOPS[v]["f64"] * cols + PER_SAMPLE[v].get("f64", 0) * samp
        kern[v] = {"ms": ms, "f32_tflops": f32 / ms / 1e9, "f64_tflops": f64 / ms / 1e9,
                   "samples_per_s": samp / (ms / 1e3), "out_gbps": samp / (ms / 1e3) / 1e9}
    # headline kernel: 26.3 grid, FP32 path
    k = kern["26.3"]
    roofline = {"bound": "valu", "achieved": round(k["f32_tflops"], 3), "peak": PEAK["f32"], "unit": "TFLOP/s",
                "frac": round(k["f32_tflops"] / PEAK["f32"], 4), "traffic": None,
                "kernel": "k_grid<26.3 float,overworld>", "kernel_ms": round(k["ms"], 3),
                "f64_side_tflops": round(k["f64_tflops"], 3),
                "issue_peak_note": "FMA-free, unpacked: 1 op/lane/clk = 39.3 T op/s is the attainable ceiling (25% of 157.3)",
                "k_grid_26_1_ms": round(kern["26.1"]["ms"], 3),
                "k_grid_26_1_f64_tflops": round(kern["26.1"]["f64_tflops"], 3),
                "output_write_GBps": round(k["out_gbps"], 3)}

    # plain per-call path: one synchronous request per seed (allocate, launch, copy back, free)
    npc = min(args.percall_seeds, N_SEEDS)
    e = engines["26.3"]
    e.region(work[0][0], work[0][1], work[0][2], SIDE, SIDE, QYS)
    t1 = time.perf_counter()
    for s, qx0, qz0 in work[:npc]:
        e.region(s, qx0, qz0, SIDE, SIDE, QYS)
    per_call = npc * SIDE * SIDE * len(QYS) / (time.perf_counter() - t1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(work)

    line = {
        "metric": "bit-exact biome samples/s batched vs per-call, 1 MI355X; p50/p99 tile latency",
        "value": round(value, 1), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64 (26.3 float noise with f64 coordinates; 26.1 f64)",
        "data": "synthetic (splitmix64-derived seeds and origins)",
        "config": {"workload": "batched Overworld quart grids: 4096 seeds x 64x64 x qy{-16,0,16,48}, 26.3 and 26.1",
                   "samples_per_step": samples_per_step, "columns_per_step": sum(plans[v].columns for v in VERSIONS),
                   "autotuned": tuned, "parallelism": "single GPU"},
        "per_call_samples_per_s": round(per_call, 1),
        "batched_over_per_call_26_3": round(kern["26.3"]["samples_per_s"] / per_call, 2),
        "per_version_kernel_samples_per_s": {v: round(kern[v]["samples_per_s"], 1) for v in VERSIONS},
        "tile_latency_ms": {"p50": None, "p99": None, "note": "serving loop not in this bench"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
