"""Benchmark of the PBF particle substep on one MI355X.

Workload: the 1,048,576-particle dam break (128x64x128 jittered lattice in a
256x128x256 box, h = 2, dt = 1/60, 4 solver iterations, XSPH c = 0.01).  A
step is one full substep: predict + cell key, cell scan, scatter, stable
permute, 4 x (lambda pass, delta-p pass), XSPH + commit.  Inputs are resident
in HBM before the timed region; hipEvents on the library's stream bracket
exactly --steps steps after --warmup untimed steps.

Per-kernel times come from a second run over the same simulation window (the
scene is re-initialised, same warmup and steps) with an event after every
kernel; the roofline entry uses the dominant kernel's average launch time from
that run.  The CPU baseline is the FP32 oracle (one thread) timed on a bounded
sample of the same scene on this host.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PBF particle-steps/s at 1M particles; learned vs plain at equal density error"
PEAK_HBM_GBS = 8000.0

# algorithmic bytes per particle per launch: compulsory float4 SoA traffic,
# neighbour gathers served on chip
PHASE_BYTES = {
    "predict_key": 64 + 24,   # read x, v (predict) + cell key, histogram
    "cell_scan": None,        # 8 bytes per cell, set from the grid
    "scatter": 12,            # read key, slot; write slot index
    "stable_permute": 96,     # read rank, x, x*, v; write sorted x, x*, id, key
    "lambda": 28,             # read x*, cell range; write lambda
    "lambda_build": 28,
    "deltap": 44,             # read x*, lambda, range; write x*
    "xsph": 88,               # read x, x*, v, range; write v, x
}
STEP_BYTES_K4 = 580           # per particle-step at 4 iterations


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(cfg, seconds_target: float = 15.0) -> dict:
    """FP32 oracle, 1 thread, on the same scene: whole steps until ~target s."""
    from oracle import oracle as O
    O.build()
    p = O.pbf_params(cfg)
    x = cfg.positions()
    v = np.zeros_like(x)
    ids = np.arange(cfg.n, dtype=np.int32)
    O.pbf_step(p, x, v, ids, np.float32(cfg.dt))  # warm (first step from rest)
    steps = 0
    t0 = time.perf_counter()
    while True:
        O.pbf_step(p, x, v, ids, np.float32(cfg.dt))
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds_target or steps >= 50:
            break
    return {
        "value": cfg.n * steps / el,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"steps 2-{steps + 1} of the {cfg.n}-particle scene ({el:.1f} s); FP32 oracle, "
                  f"-O2 -fno-fast-math -ffp-contract=off, 1 thread of {cpu_model()} "
                  f"({os.cpu_count()} cpus visible)",
    }


def bench_batch(args, G, scenes) -> int:
    """64 independent dam breaks stepped together: aggregate particle-steps/s
    and the per-scene step latency (every scene of the batch completes its step
    when the batch step completes, so a scene's latency is the batch step time)."""
    b = scenes.cfg4()
    base = b.scenes[0]
    if args.iterations is not None:
        for s in b.scenes:
            s.iterations = args.iterations
    params = G.pbf_params(base, exact=args.exact)
    dt = np.float32(base.dt)
    with G.Gpu(0) as g:
        g.pbf_batch_init(params, b.counts(), b.boxes(), b.positions())
        g.pbf_set_graph(not args.no_graph)
        ms, _, _ = g.pbf_bench(dt, args.warmup, args.steps, phases=False)
        g.pbf_batch_init(params, b.counts(), b.boxes(), b.positions())
        g.pbf_set_graph(not args.no_graph)
        g.pbf_step(dt, args.warmup)
        lat = g.pbf_step_latency(dt, args.steps)
    value = b.n * args.steps / (ms / 1e3)
    out = {
        "metric": "PBF particle-steps/s, 64 batched scenes (aggregate)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": b.name, "scenes": len(b.scenes), "particles": b.n,
                   "scene_particles_min": int(b.counts().min()), "scene_particles_max": int(b.counts().max()),
                   "iterations": base.iterations, "arithmetic": "exact" if args.exact else "fused",
                   "launch": "plain" if args.no_graph else "hipgraph", "parallelism": "single-gpu"},
        "scene_step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                  "max": float(lat.max())},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg2", choices=["cfg0", "cfg2", "cfg2_tail", "cfg4"])
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--exact", action="store_true",
                    help="oracle operation order (bit-identical results) instead of the fused path")
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph replay")
    ap.add_argument("--no-phases", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        return 2

    from lustrine_amd import gpu as G
    from lustrine_amd import scenes

    if G.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.workload == "cfg4":
        return bench_batch(args, G, scenes)
    cfg = {"cfg0": scenes.cfg0, "cfg2": scenes.cfg2, "cfg2_tail": scenes.cfg2_tail}[args.workload]()
    if args.iterations is not None:
        cfg.iterations = args.iterations
    x0 = cfg.positions()
    gx, gy, gz = cfg.grid
    ncell = gx * gy * gz
    params = G.pbf_params(cfg, exact=args.exact)
    dt = np.float32(cfg.dt)
    with G.Gpu(0) as g:
        g.pbf_init(params, x0, None)
        g.pbf_set_graph(not args.no_graph)
        ms, _, _ = g.pbf_bench(dt, args.warmup, args.steps, phases=False)
        maxc = g.pbf_last_max_constraint()
        phase_ms = launches = None
        copy_gbps = None
        if not args.no_phases:
            g.pbf_init(params, x0, None)
            _, phase_ms, launches = g.pbf_bench(dt, args.warmup, args.steps, phases=True)
            copy_gbps, _ = g.stream_copy_gbps(1 << 30, 10)
    ms_per_step = ms / args.steps
    value = cfg.n * args.steps / (ms / 1e3)
    step_bytes = (STEP_BYTES_K4 - 4 * 72 + cfg.iterations * 72) * cfg.n

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg.name, "particles": cfg.n, "lattice": list(cfg.lattice),
                   "box": list(cfg.box), "cells": ncell, "h": cfg.h, "dt": cfg.dt,
                   "iterations": cfg.iterations, "c_xsph": cfg.c_xsph,
                   "arithmetic": "exact" if args.exact else "fused",
                   "launch": "plain" if args.no_graph else "hipgraph", "parallelism": "single-gpu"},
        "step_roofline": {"alg_bytes_per_step": int(step_bytes),
                          "achieved_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "max_density_constraint_last_step": maxc,
    }
    if phase_ms is not None:
        pb = dict(PHASE_BYTES)
        pb["cell_scan"] = 8.0 * (ncell + 1) / cfg.n
        phases = {}
        for name, t in phase_ms.items():
            nl = launches[name]
            if nl == 0:
                continue
            avg_us = t / nl * 1e3
            b = pb[name] * cfg.n
            phases[name] = {"total_ms": round(t, 4), "launches": nl, "avg_launch_us": round(avg_us, 2),
                            "alg_bytes_per_launch": int(b),
                            "alg_GBps": round(b / (avg_us * 1e-6) / 1e9, 1)}
        dom = max(phases, key=lambda k: phases[k]["total_ms"])
        d = phases[dom]
        achieved = d["alg_bytes_per_launch"] / (d["avg_launch_us"] * 1e-6) / 1e9
        traffic = None
        tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{cfg.name}.json")
        if os.path.exists(tfile):
            with open(tfile) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                           "avg_launch_us": d["avg_launch_us"],
                           "alg_bytes_per_launch": d["alg_bytes_per_launch"],
                           "measured_copy_GBps": round(copy_gbps, 1)}
        out["phases"] = phases
    if not args.no_cpu_baseline:
        cb = cpu_baseline(cfg, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
