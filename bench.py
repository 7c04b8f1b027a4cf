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
that run.  `learned_vs_plain` compares the learned warm start (network
lustrine_amd/data/warm_start.npz, LEARNED_ITERATIONS plain iterations after
the network) with plain K = 4..16 on the same window and the same
density-error measure.  The CPU baseline is the
FP32 oracle (one thread) timed on a bounded sample of the same scene on this
host.  Prints one JSON line.  --workload cfg3 runs the full learned sweep,
cfg4 the 64-scene batch.
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
    "scatter": None,          # read key, slot; write slot index (12) + zero the histogram, set from the grid
    "stable_permute": 96,     # read rank, x, x*, v; write sorted x, x*, id, key
    "lambda": 28,             # read x*, cell range; write lambda
    "lambda_build": 28,
    "deltap": 44,             # read x*, lambda, range; write x*
    "xsph": 96,               # read x, x*, v, range; write v, x; the next step's key and slot
    "features": 104,          # read x*, x, range; write 32 bf16 features
    "mlp": 96,                # read 32 bf16 features, x*; write x*
    "decide": None,           # one block over the block maxima
    "tiles": None,            # tile list: two cell-start reads per chunk, one int4 per tile
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
    from lustrine_amd import autotune as T
    params = G.pbf_params(base, exact=args.exact)
    dt = np.float32(base.dt)
    counts, boxes, pos = b.counts(), b.boxes(), b.positions()
    with G.Gpu(0) as g:
        def init(launch):
            g.pbf_batch_init(T.cell_params(params, None, launch.cell_scale), counts, boxes, pos)
        tuner = T.Tuner()
        best, table = (T.Launch(), []) if args.no_autotune else tuner.tune(g, init, b.n, dt)
        init(best)
        T.apply(g, best)
        g.pbf_set_graph(not args.no_graph)
        ms, _, _ = g.pbf_bench(dt, args.warmup, args.steps, phases=False)
        init(best)
        T.apply(g, best)
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
        "autotune": {"bucket": T.bucket(b.n), "choice": T.asdict(best),
                     "candidates_ms_per_step": [dict(**T.asdict(c), ms=round(t, 4)) for c, t in table]},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


PLAIN_SWEEP = (1, 2, 3, 4, 6, 8, 12, 16, 32)
# plain Jacobi iterations after the network's correction in the learned step;
# with one, the network's residual error feeds back through the velocities and
# the 1M dam break drifts off the training distribution within ~50 steps
LEARNED_ITERATIONS = 3
# the bounded learned step's plain rerun: enough iterations that a rerun step
# ends below the learned step's typical error
RERUN_ITERATIONS = 16


def measure_mode(g, G, cfg, x0, dt, warmup, steps, iterations, learned_iters=None, bound=float("inf"),
                 exact=False, graph=True):
    """Throughput and per-step density error e_t = max_i max(0, rho_i/rho0 - 1)
    at the final positions of one solver configuration over the same window
    (warmup untimed steps, then `steps` steps).  learned_iters None: plain
    solve with `iterations`; else the learned step with that many plain
    iterations after the network and the plain rerun (`iterations`) above
    `bound`.  Two runs from the same initial state: a timed one (plain: no
    density measurement) and one that records e_t."""
    params = G.pbf_params(cfg, iterations=iterations, exact=exact)
    learned = learned_iters is not None
    g.pbf_init(params, x0, None)
    g.pbf_set_mode(learned, learned_iters or 0, bound, check=False)
    g.pbf_set_graph(graph)
    ms, _, _ = g.pbf_bench(dt, warmup, steps, phases=False)
    g.pbf_init(params, x0, None)
    g.pbf_set_mode(learned, learned_iters or 0, bound, check=True)
    g.pbf_set_graph(graph)
    g.pbf_step(dt, warmup)
    _, r0, _ = g.pbf_learned_stats()
    g.pbf_step(dt, steps)
    _, r1, errs, errm = g.pbf_learned_stats(with_mean=True)
    _, vel, _ = g.pbf_download(order=0)
    e = errs[-steps:].astype(np.float64)
    em = errm[-steps:].astype(np.float64)
    return {"iterations": iterations, "learned_iterations": learned_iters,
            "bound": None if not learned or not np.isfinite(bound) else bound,
            "ms_per_step": ms / steps, "particle_steps_per_s": cfg.n * steps / (ms / 1e3),
            "err_mean": float(e.mean()), "err_p50": float(np.percentile(e, 50)), "err_max": float(e.max()),
            "avg_compression_mean": float(em.mean()),
            "mean_speed_end": float(np.linalg.norm(vel.astype(np.float64), axis=1).mean()),
            "rerun_fraction": (r1 - r0) / steps if learned else None}


def equal_error_speedup(plain, learned_row, key="err_mean"):
    """Learned throughput over the throughput of the cheapest plain K whose
    error is at most the learned error (plain rows sorted by K).  None when no
    plain K reaches it."""
    target = learned_row[key]
    ok = [r for r in plain if r[key] <= target]
    if not ok:
        return None, None
    best = max(ok, key=lambda r: r["particle_steps_per_s"])
    return learned_row["particle_steps_per_s"] / best["particle_steps_per_s"], best["iterations"]


def learned_vs_plain(G, cfg, x0, dt, warmup, steps, net, sweep=(4, 8, 12, 16), graph=True):
    """The compact comparison of the default run: plain K over `sweep` and the
    learned step with LEARNED_ITERATIONS plain iterations, same window, same
    error measure."""
    with G.Gpu(0) as g:
        g.pbf_set_network(net)
        plain = [measure_mode(g, G, cfg, x0, dt, warmup, steps, k, graph=graph) for k in sweep]
        lrn = measure_mode(g, G, cfg, x0, dt, warmup, steps, cfg.iterations, learned_iters=LEARNED_ITERATIONS,
                           graph=graph)
    out = {"learned": lrn, "plain_sweep": plain}
    for key in ("err_mean", "err_max", "avg_compression_mean"):
        sp, k = equal_error_speedup(plain, lrn, key)
        out[f"speedup_at_equal_{key}"] = sp
        out[f"plain_iterations_at_equal_{key}"] = k
    out["note"] = ("error e_t = max_i max(0, rho_i/rho0 - 1) at the final positions per step; speedup None: "
                   f"no plain K in {list(sweep)} reaches the learned error (then the learned throughput over "
                   "the largest K's bounds it from below)")
    return out


def bench_learned(args, G, scenes) -> int:
    """Learned warm start vs the plain solver on the 1M dam break: an
    error-vs-throughput curve (plain K over PLAIN_SWEEP, learned with 1..3
    plain iterations), and the learned step with a rerun bound: 1% (no plain
    K reaches it on this scene, so every step reruns) and the plain K=12
    solve's worst per-step error, both rerunning RERUN_ITERATIONS plain
    iterations."""
    from lustrine_amd import learned as L
    cfg = scenes.cfg2()
    net_path = args.network or L.DEFAULT_NETWORK
    net, meta = L.load(net_path)
    x0 = cfg.positions()
    dt = np.float32(cfg.dt)
    with G.Gpu(0) as g:
        g.pbf_set_network(net)
        plain = [measure_mode(g, G, cfg, x0, dt, args.warmup, args.steps, k) for k in PLAIN_SWEEP]
        lrn = [measure_mode(g, G, cfg, x0, dt, args.warmup, args.steps, cfg.iterations, learned_iters=li)
               for li in (1, 2, 3)]
        guard = next(r for r in plain if r["iterations"] == 12)["err_max"]
        bounded = [measure_mode(g, G, cfg, x0, dt, args.warmup, args.steps, RERUN_ITERATIONS,
                                learned_iters=LEARNED_ITERATIONS, bound=b) for b in (0.01, guard)]
    head = next(r for r in lrn if r["learned_iterations"] == LEARNED_ITERATIONS)
    sp_mean, k_mean = equal_error_speedup(plain, head, "err_mean")
    sp_max, k_max = equal_error_speedup(plain, head, "err_max")
    sp_avg, k_avg = equal_error_speedup(plain, head, "avg_compression_mean")
    out = {
        "metric": METRIC,
        "value": head["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16",
        "data": "synthetic",
        "config": {"workload": "pbf_dam_break_1m_learned", "particles": cfg.n,
                   "learned_iterations": LEARNED_ITERATIONS, "rerun_iterations": RERUN_ITERATIONS,
                   "network": os.path.relpath(net_path, ROOT),
                   "network_training": {k: meta.get(k) for k in ("scenes", "particles_per_scene", "steps",
                                                                   "teacher_iterations", "rows", "epochs")},
                   "error": "per step: max_i max(0, rho_i/rho0 - 1) at the final positions; "
                            "err_mean / err_max over the timed window", "parallelism": "single-gpu"},
        "equal_error": {"by_err_mean": {"speedup": sp_mean, "plain_iterations": k_mean},
                        "by_err_max": {"speedup": sp_max, "plain_iterations": k_max},
                        "by_avg_compression": {"speedup": sp_avg, "plain_iterations": k_avg},
                        "note": "speedup None: no plain K in the sweep reaches the learned error; "
                                "the learned throughput over the largest K's is then a lower bound"},
        "plain_sweep": plain,
        "learned": lrn,
        "learned_bounded": bounded,
        "bounded_equal_err_max": [dict(zip(("speedup", "plain_iterations"),
                                           equal_error_speedup(plain, r, "err_max")), bound=r["bound"])
                                  for r in bounded],
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def bench_sand(args, G, scenes) -> int:
    """The reference's default simulate_fun, simulate_sand (src/Simulate.cpp:156-325:
    predict, sorted-grid lists, 4 Jacobi contact + friction substeps, velocity update),
    bit-exact on the GPU, on the 1,048,576-particle dam-break lattice as sand (no
    solids; h = 3.1 r, init_simulation's default kernel radius).  Sand particle-steps/s
    timed on the host around synchronised batches of steps, the phase slots of one
    profiled step, and the oracle's simulate_sand on one host core."""
    cfg = scenes.cfg2()
    pos = cfg.positions()
    sp = G.sand_params(*cfg.box)
    dt = np.float32(cfg.dt)
    n = pos.shape[0]
    with G.Gpu(0) as g:
        g.sand_init(sp, pos)
        for _ in range(args.warmup):
            g.sand_step(dt)
        g.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.sand_step(dt)
        g.sync()
        sec = time.perf_counter() - t0
        slots = None
        if not args.no_phases:
            g.profiling(True)
            g.sand_step(dt)
            g.sync()
            _, slots = G.profile_slots()
            g.profiling(False)
        # the same steps without the mid-step wait (a fixed list buffer of 48
        # entries per particle; ~33 are used) and replayed as hipGraphs
        g.sand_init(sp, pos)
        cap = 48 * n
        g.sand_set_list_capacity(cap)
        g.sand_set_graph(True)
        for _ in range(args.warmup):
            g.sand_step(dt)
        g.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.sand_step(dt)
        g.sync()
        sec_g = time.perf_counter() - t0
    ms = sec * 1e3
    out = {
        "metric": "simulate_sand particle-steps/s at 1M sand particles",
        "value": n * args.steps / sec, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "sand_dam_break_1m", "particles": n, "lattice": list(cfg.lattice),
                   "box": list(cfg.box), "h": float(sp.kernelRadius), "cells": sp.grid.gridX * sp.grid.gridY * sp.grid.gridZ,
                   "substeps": int(sp.substeps), "dt": cfg.dt, "arithmetic": "exact (bit-identical to the oracle)",
                   "timing": "host clock around synchronised step batches", "parallelism": "single-gpu"},
    }
    out["no_wait_graph"] = {"ms_per_step": sec_g * 1e3 / args.steps, "particle_steps_per_s": n * args.steps / sec_g,
                            "list_capacity": cap,
                            "note": "lus_sand_set_list_capacity + lus_sand_set_graph: no mid-step wait, each step "
                                    "a replayed hipGraph; the same steps bit for bit"}
    if slots is not None:
        out["phases_ms"] = {"step": slots[0] * 1e3, "neighbour_search": slots[3] * 1e3,
                            "substeps": slots[4] * 1e3, "velocity_update": slots[6] * 1e3,
                            "window": f"step {args.warmup + args.steps + 1}"}
    out["adapter_call"] = adapter_call_ms(pos, cfg.box, steps=10)
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        s = O.init_constants(*cfg.box, 0.5, 1.0, 3.1)
        st = O.SandStream(s, pos, None)
        st.step(dt)  # the first call also copies the solids (src/Simulate.cpp:217-223)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            st.step(dt)
            done += 1
        csec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * done / csec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"steps 2-{done + 1} of the sand scene ({csec:.1f} s), FP32 oracle "
                                         f"simulate_sand, 1 thread of {cpu_model()}"}
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))
    return 0


def adapter_call_ms(pos, box, steps=10) -> dict:
    """One lus_adapter::simulate_sand call on a Simulation-shaped host state
    (tests/cpp/adapter_sand.exe, the header-only drop-in): ms per call when
    every call leaves positions, positions_star, velocities and attracted in
    the host arrays, ms per call when they stay on the device, and the ms of
    the one synchronisation that brings them back."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "tests", "cpp", "adapter_sand.exe")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/adapter_sand.exe is not built"}
    n = pos.shape[0]
    total = max(n, int(box[0]) * int(box[1]) * int(box[2]))
    with tempfile.TemporaryDirectory() as td:
        inp, outp = os.path.join(td, "in.bin"), os.path.join(td, "out.txt")
        with open(inp, "wb") as f:
            np.array([n, 0, total], np.int32).tofile(f)
            np.ascontiguousarray(pos, np.float32).tofile(f)
        r = subprocess.run([exe, "bench", inp, outp, str(steps), *map(str, box), "3.1"], capture_output=True,
                           text=True, timeout=300)
        if r.returncode != 0:
            return {"error": r.stderr[-400:]}
        every, lazy, sync = (float(t) for t in open(outp).read().split())
    return {"ms_per_call_host_synced": every, "ms_per_call_on_demand": lazy, "ms_sync_host": sync,
            "particles": n, "calls": steps,
            "note": "lus_adapter::simulate_sand on a Simulation-shaped state; host synced: positions, "
                    "positions_star, velocities, attracted back in the host arrays after every call"}


def bench_fluid(args, G, scenes) -> int:
    """The reference's simulate_fluid (src/Simulate.cpp:27-115) on the GPU: the
    1,048,576-particle lattice in 256 x 128 x 256 with the reference's
    Simulation defaults (rest density 24, epsilon 10, cubic kernel, h = 4 r),
    its sequential pressure sweep run as dependency levels.  Host clock
    around synchronised steps (the sweep synchronises every 32 levels)."""
    from oracle import oracle as O
    O.build()
    cfg = scenes.cfg2()
    s = O.init_constants(*cfg.box, 0.5, 1.0, 4.0)
    sand = cfg.positions()
    n = sand.shape[0]
    total = max(n, int(cfg.box[0]) * int(cfg.box[1]) * int(cfg.box[2]))
    p = np.zeros((total, 3), np.float32)
    p[:n] = sand
    dt = 1.0 / 60.0
    with G.Gpu(0) as g:
        g.fluid_init(G.fluid_params(s), p, p, np.zeros_like(p), (0, n), (total - 1, total - 1))
        for _ in range(args.warmup):
            g.fluid_step(dt)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.fluid_step(dt)
        sec = time.perf_counter() - t0
        levels, entries = g.fluid_stats()
    out = {"metric": "simulate_fluid particle-steps/s at 1M particles", "value": n * args.steps / sec,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": sec * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "fluid_dam_break_1m", "particles": n, "sweep_levels": levels,
                      "list_entries": entries, "parallelism": "single-gpu"}}
    if not args.no_cpu_baseline:
        po, pso, vo, lo = p.copy(), p.copy(), np.zeros_like(p), np.zeros(total, np.float32)
        O.simulate_fluid(s, po, pso, vo, lo, (0, n), (total - 1, total - 1), dt)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            O.simulate_fluid(s, po, pso, vo, lo, (0, n), (total - 1, total - 1), dt)
            done += 1
        csec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * done / csec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"steps 2-{done + 1} ({csec:.1f} s), FP32 oracle simulate_fluid, "
                                         f"1 thread of {cpu_model()}"}
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))
    return 0


def bench_mixed(args, G, scenes) -> int:
    """Mixed sand + fluid over the voxel heightfield: moving particle-steps/s
    (solids are not counted), 3 iterations, hipGraph replay."""
    m = scenes.cfg1()
    pos, ph = m.arrays()
    params = G.pbf_params(m.fluid, exact=args.exact)
    mp = G.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
    dt = np.float32(m.fluid.dt)
    with G.Gpu(0) as g:
        g.mix_init(params, mp, pos, ph)
        g.pbf_set_graph(not args.no_graph)
        ms, _, _ = g.pbf_bench(dt, args.warmup, args.steps, phases=False)
        phase_ms = launches = None
        if not args.no_phases:
            g.mix_init(params, mp, pos, ph)
            _, phase_ms, launches = g.pbf_bench(dt, args.warmup, args.steps, phases=True)
    value = m.n_moving * args.steps / (ms / 1e3)
    out = {
        "metric": "mixed sand+fluid moving particle-steps/s",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": m.name, "moving": m.n_moving, "solids": int((ph == 2).sum()),
                   "sand": int((ph == 1).sum()), "fluid": int((ph == 0).sum()), "box": list(m.fluid.box),
                   "iterations": m.fluid.iterations, "mu_s": m.mu_s, "mu_k": m.mu_k,
                   "arithmetic": "exact" if args.exact else "fused",
                   "launch": "plain" if args.no_graph else "hipgraph", "parallelism": "single-gpu"},
    }
    if phase_ms is not None:
        out["phases"] = {k: {"total_ms": round(t, 4), "launches": launches[k],
                             "avg_launch_us": round(t / launches[k] * 1e3, 2)}
                         for k, t in phase_ms.items() if launches[k]}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        p = O.pbf_params(m.fluid)
        omp = O.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
        x, v = pos.copy(), np.zeros_like(pos)
        ids = np.arange(pos.shape[0], dtype=np.int32)
        phs = ph.copy()
        O.mixed_step(p, omp, x, v, ids, phs, dt)
        steps, t0 = 0, time.perf_counter()
        while True:
            O.mixed_step(p, omp, x, v, ids, phs, dt)
            steps += 1
            el = time.perf_counter() - t0
            if el >= args.cpu_seconds or steps >= 50:
                break
        cb = {"value": m.n_moving * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
              "sample": f"steps 2-{steps + 1} of the mixed scene ({el:.1f} s), FP32 oracle, 1 thread of {cpu_model()}"}
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="cfg2", choices=["cfg0", "cfg1", "cfg2", "cfg2_tail", "cfg3", "cfg4", "sand_1m",
                                                         "fluid_1m"])
    ap.add_argument("--network", default=None, help="warm-start network file (cfg3)")
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--exact", action="store_true",
                    help="oracle operation order (bit-identical results) instead of the fused path")
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph replay")
    ap.add_argument("--no-phases", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="cfg4: default launch instead of the autotuner")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-vs-plain comparison of the default run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: the library runs a scene on one GPU; --gpus {args.gpus} is not supported "
              "(sharding a scene over several GPUs is not implemented)", file=sys.stderr)
        return 2

    from lustrine_amd import gpu as G
    from lustrine_amd import scenes

    if G.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.workload == "cfg4":
        return bench_batch(args, G, scenes)
    if args.workload == "cfg3":
        return bench_learned(args, G, scenes)
    if args.workload == "cfg1":
        return bench_mixed(args, G, scenes)
    if args.workload == "sand_1m":
        return bench_sand(args, G, scenes)
    if args.workload == "fluid_1m":
        return bench_fluid(args, G, scenes)
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
        pb["scatter"] = 12.0 + 4.0 * (ncell + 1) / cfg.n
        phases = {}
        for name, t in phase_ms.items():
            nl = launches[name]
            if nl == 0:
                continue
            avg_us = t / nl * 1e3
            phases[name] = {"total_ms": round(t, 4), "launches": nl, "avg_launch_us": round(avg_us, 2)}
            if pb.get(name) is not None:
                b = pb[name] * cfg.n
                phases[name].update(alg_bytes_per_launch=int(b), alg_GBps=round(b / (avg_us * 1e-6) / 1e9, 1))
        dom = max((p for p in phases if "alg_bytes_per_launch" in phases[p]), key=lambda k: phases[k]["total_ms"])
        d = phases[dom]
        achieved = d["alg_bytes_per_launch"] / (d["avg_launch_us"] * 1e-6) / 1e9
        # HBM traffic per launch is not measurable inside this process (it
        # takes the FETCH_SIZE / WRITE_SIZE counters of two profiled runs):
        # it is read from the committed summary of those runs, whose window
        # and command are reported with it
        traffic = traffic_src = None
        tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{cfg.name}.json")
        if os.path.exists(tfile):
            with open(tfile) as f:
                tj = json.load(f)
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_src = {"file": os.path.relpath(tfile, ROOT), "window": tj.get("window"),
                           "kernel": tj.get("kernel"), "avg_launch_us_profiled": tj.get("avg_launch_us")}
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                           "avg_launch_us": d["avg_launch_us"],
                           "alg_bytes_per_launch": d["alg_bytes_per_launch"],
                           "measured_copy_GBps": round(copy_gbps, 1),
                           "window": f"steps {args.warmup + 1}-{args.warmup + args.steps} after "
                                     f"{args.warmup} warmup, kernels timed by their own dispatch events",
                           "traffic_source": traffic_src}
        out["phases"] = phases
    if not args.no_learned and args.workload == "cfg2" and cfg.iterations >= 1:
        from lustrine_amd import learned as L
        if os.path.exists(L.DEFAULT_NETWORK):
            net, _ = L.load(L.DEFAULT_NETWORK)
            out["learned_vs_plain"] = learned_vs_plain(G, cfg, x0, dt, args.warmup, args.steps, net,
                                                   graph=not args.no_graph)
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
