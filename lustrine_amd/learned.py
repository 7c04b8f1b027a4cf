"""Learned warm start of the PBF solve: training data, fitting, network files.

The network (include/lustrine_gpu.h, "learned warm start") maps 32 features of
a particle's predicted neighbourhood to the position correction that the
plain solver's `teacher_iterations` Jacobi iterations produce in that step.
Training data come from the plain solver itself, on the GPU: a batch of dam
breaks is stepped with lus_pbf_record_step, which appends the features and
targets of every `stride`-th cell-order slot (rotating offset) of each step to
a device-resident training set; lus_mlp_fit then fits the network with the
BF16-MFMA training kernels.  Nothing here runs on the CPU besides the loop.

    python -m lustrine_amd.learned --out warm_start.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import gpu as G
from . import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_NETWORK = os.path.join(HERE, "data", "warm_start.npz")


def generate(g: "G.Gpu", batch: "scenes.BatchConfig", teacher_iterations: int = 4, stride: int = 16,
             steps: int | None = None, exact: bool = False, capacity: int | None = None) -> "G.Trainer":
    """Steps `batch` with the plain solver (teacher_iterations iterations),
    recording 1/stride of the particles of every step.  Returns the trainer."""
    steps = batch.steps if steps is None else steps
    base = batch.scenes[0]
    params = G.pbf_params(base, iterations=teacher_iterations, exact=exact)
    g.pbf_batch_init(params, batch.counts(), batch.boxes(), batch.positions())
    g.pbf_set_mode(False)
    cap = capacity or (batch.n // stride + 1) * steps
    t = g.trainer(cap)
    dt = np.float32(base.dt)
    for _ in range(steps):
        t.record_step(dt, stride)
    return t


def fit(t: "G.Trainer", epochs: int = 24, batch: int = 4096, lr: float = 2e-3, lr_final: float = 5e-5,
        val_fraction: float = 0.05, seed: int = 1):
    tl, vl = t.fit(epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, val_fraction=val_fraction, seed=seed)
    return t.network(), tl, vl


def save(path: str, net: np.ndarray, meta: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, net=np.asarray(net, np.float32), meta=np.array(json.dumps(meta)))


def load(path: str = DEFAULT_NETWORK):
    """(net float32 [NET_FLOATS], meta dict); arrays only, nothing executed."""
    with np.load(path, allow_pickle=False) as z:
        net = z["net"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if net.shape != (G.NET_FLOATS,):
        raise ValueError(f"{path}: network has {net.shape} floats, expected {G.NET_FLOATS}")
    return net, meta


def train_default(device: int = 0, n_scenes: int = 64, n_per_scene: int = 8192, steps: int = 200,
                  teacher_iterations: int = 8, stride: int = 16, epochs: int = 24, seed: int = 3,
                  log=print):
    """The shipped network: 64 dam breaks with box aspect U[0.5, 2], 200 steps,
    targets from 8 plain iterations."""
    b = scenes.training_batch(n_scenes, n_per_scene, seed)
    with G.Gpu(device) as g:
        t0 = time.perf_counter()
        t = generate(g, b, teacher_iterations, stride, steps)
        g.sync()
        t1 = time.perf_counter()
        rows = t.rows
        log(f"training set: {rows} rows from {b.n} particles x {steps} steps ({t1 - t0:.1f} s)")
        net, tl, vl = fit(t, epochs=epochs)
        t2 = time.perf_counter()
        log(f"fit: {epochs} epochs ({t2 - t1:.1f} s); train loss {tl[0]:.4g} -> {tl[-1]:.4g}, "
            f"validation {vl[0]:.4g} -> {vl[-1]:.4g}")
        t.close()
    meta = {"scenes": n_scenes, "particles_per_scene": n_per_scene, "steps": steps,
            "teacher_iterations": teacher_iterations, "stride": stride, "epochs": epochs, "seed": seed,
            "rows": int(rows), "train_loss": [float(x) for x in tl], "val_loss": [float(x) for x in vl],
            "generate_s": t1 - t0, "fit_s": t2 - t1}
    return net, meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_NETWORK)
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--particles", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--teacher-iterations", type=int, default=8)
    ap.add_argument("--stride", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--seed", type=int, default=3)
    a = ap.parse_args(argv)
    net, meta = train_default(0, a.scenes, a.particles, a.steps, a.teacher_iterations, a.stride, a.epochs,
                              a.seed, log=lambda s: print(s, flush=True))
    save(a.out, net, meta)
    print(json.dumps({"out": a.out, **{k: meta[k] for k in ("rows", "generate_s", "fit_s")},
                      "val_loss_last": meta["val_loss"][-1]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
