"""Benchmark of the PBF particle substep on one MI355X.

Workload: the 1,048,576-particle dam break (128x64x128 jittered lattice in a
256x128x256 box, h = 2, dt = 1/60, 4 solver iterations, XSPH c = 0.01).  A
step is one full substep: predict, cell key + histogram, cell scan, stable
scatter/permute, 4 x (lambda pass, delta-p pass), XSPH + commit.  Inputs are
resident in HBM before the timed region; hipEvents on the library's stream
bracket exactly --steps steps after --warmup untimed steps.

Prints one JSON line.  The CPU baseline is the FP32 oracle (one thread) timed
on a bounded sample of the same scene on this host.
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

# algorithmic bytes per particle per launch (compulsory float4 SoA traffic,
# neighbour gathers served on chip)
PHASE_BYTES = {
    "predict_key": 64 + 24,   # read x, v; (predict) + cell key, histogram
    "cell_scan": None,        # 8 bytes per cell, set from the grid
    "scatter": 12,            # read key, slot; write slot index
    "stable_permute": 96,     # read rank, x, x*, v; write sorted x, x*, id, key
    "lambda": 28,             # read x*, cell range; write lambda
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
        "sample": f"{steps} steps of the {cfg.n}-particle scene after 1 warm step, "
                  f"{el:.1f} s, FP32 oracle (-O2, no fast-math, no FMA), {cpu_model()}, "
                  f"{os.cpu_count()} cpus visible",
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg2", choices=["cfg0", "cfg2", "cfg2_tail"])
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is described, not built)", file=sys.stderr)
        return 2

    from lustrine_amd import gpu as G
    from lustrine_amd import scenes

    cfg = {"cfg0": scenes.cfg0, "cfg2": scenes.cfg2, "cfg2_tail": scenes.cfg2_tail}[args.workload]()
    if args.iterations is not None:
        cfg.iterations = args.iterations
    if G.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    x0 = cfg.positions()
    gx, gy, gz = cfg.grid
    ncell = gx * gy * gz
    with G.Gpu(0) as g:
        g.pbf_init(G.pbf_params(cfg), x0, None)
        ms, phase_ms, launches = g.pbf_bench(np.float32(cfg.dt), args.warmup, args.steps, phases=True)
        maxc = g.pbf_last_max_constraint()
    ms_per_step = ms / args.steps
    value = cfg.n * args.steps / (ms / 1e3)

    # per-phase roofline from the event-bracketed second run
    pb = dict(PHASE_BYTES)
    pb["cell_scan"] = 8.0 * (ncell + 1) / cfg.n
    phases = {}
    for name, t in phase_ms.items():
        nl = max(launches[name], 1)
        avg_ms = t / nl
        bytes_launch = pb[name] * cfg.n
        gbs = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        phases[name] = {"total_ms": round(t, 4), "launches": launches[name],
                        "avg_launch_us": round(avg_ms * 1e3, 3), "alg_bytes_per_launch": int(bytes_launch),
                        "alg_GBps": round(gbs, 1)}
    dom = max(phase_ms, key=lambda k: phase_ms[k])
    d = phases[dom]
    achieved = d["alg_GBps"]
    step_bytes = (STEP_BYTES_K4 - 4 * 72 + cfg.iterations * 72) * cfg.n
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{cfg.name}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

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
                   "iterations": cfg.iterations, "c_xsph": cfg.c_xsph, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic},
        "step_roofline": {"alg_bytes_per_step": int(step_bytes),
                          "achieved_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "phases": phases,
        "max_density_constraint_last_step": maxc,
    }
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
