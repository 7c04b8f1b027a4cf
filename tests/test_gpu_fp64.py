"""Deviation of the GPU step from the FP64 oracle, per configuration and
arithmetic mode, next to the deviation of the FP32 oracle (the reference's
arithmetic order) from the same FP64 step.  Positions are compared as
max |a - b| / max |b|; velocities (x* - x) / dt carry the positions' rounding
divided by dt, so their figure is also given relative to one FP32 ulp of the
largest coordinate over dt.  Each case prints one "FP64 deviation" line."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _by_id(a, ids):
    out = np.empty_like(a)
    out[ids] = a
    return out


def _dev(a, b):
    return float(np.abs(a.astype(np.float64) - b).max()) / max(float(np.abs(b).max()), 1e-30)


def _report(tag, x64, v64, pairs, dt):
    ulp_v = float(np.spacing(np.float32(np.abs(x64).max()))) / dt
    parts = []
    devs = {}
    for name, (x, v) in pairs.items():
        dx, dv = _dev(x, x64), _dev(v, v64)
        dva = float(np.abs(v.astype(np.float64) - v64).max()) / ulp_v
        devs[name] = (dx, dv, dva)
        parts.append(f"{name}: x {dx:.2e} v {dv:.2e} ({dva:.1f} ulp(x)/dt)")
    print(f"FP64 deviation {tag}: " + "; ".join(parts), flush=True)
    return devs


@pytest.mark.parametrize("name,steps", [("cfg0", 1), ("cfg0", 10), ("cfg2", 1)])
def test_pbf_fp64_deviation(gpu_mod, oracle, name, steps):
    cfg = {"cfg0": scenes.cfg0, "cfg2": scenes.cfg2}[name]()
    x0 = cfg.positions()
    n = x0.shape[0]
    p = oracle.pbf_params(cfg)
    x64, v64, i64 = x0.astype(np.float64), np.zeros((n, 3), np.float64), np.arange(n, dtype=np.int32)
    x32, v32, i32 = x0.copy(), np.zeros_like(x0), np.arange(n, dtype=np.int32)
    for _ in range(steps):
        oracle.pbf_step(p, x64, v64, i64, cfg.dt, fp64=True)
        oracle.pbf_step(p, x32, v32, i32, np.float32(cfg.dt))
    x64, v64 = _by_id(x64, i64), _by_id(v64, i64)
    pairs = {"fp32 oracle": (_by_id(x32, i32), _by_id(v32, i32))}
    for exact in (True, False):
        with gpu_mod.Gpu(0) as g:
            g.pbf_init(gpu_mod.pbf_params(cfg, exact=exact), x0, None)
            g.pbf_step(np.float32(cfg.dt), steps)
            gx, gv, gid = g.pbf_download(order=0)
        pairs["gpu exact" if exact else "gpu fused"] = (_by_id(gx, gid), _by_id(gv, gid))
    devs = _report(f"{name} ({n} particles) after {steps} step(s)", x64, v64, pairs, cfg.dt)
    _check(devs)


def _check(devs):
    # positions: within the 1e-5 bar of the FP64 step and no further from it
    # than a small multiple of the reference-order FP32 path; velocities: in
    # units of ulp(max |x|) / dt, at most twice the FP32 path's figure plus 2
    ref_x, _, ref_va = devs["fp32 oracle"]
    for mode in ("gpu exact", "gpu fused"):
        dx, _, dva = devs[mode]
        assert dx <= 1e-5 and dx <= 4.0 * ref_x + 1e-7, (mode, dx, ref_x)
        assert dva <= 2.0 * ref_va + 2.0, (mode, dva, ref_va)


def test_mixed_fp64_deviation(gpu_mod, oracle):
    m = scenes.cfg1()
    pos, ph = m.arrays()
    n = pos.shape[0]
    p = oracle.pbf_params(m.fluid)
    mp = oracle.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
    x64, v64, i64, p64 = pos.astype(np.float64), np.zeros((n, 3), np.float64), np.arange(n, dtype=np.int32), ph.copy()
    x32, v32, i32, p32 = pos.copy(), np.zeros_like(pos), np.arange(n, dtype=np.int32), ph.copy()
    oracle.mixed_step(p, mp, x64, v64, i64, p64, m.fluid.dt, fp64=True)
    oracle.mixed_step(p, mp, x32, v32, i32, p32, np.float32(m.fluid.dt))
    x64, v64 = _by_id(x64, i64), _by_id(v64, i64)
    pairs = {"fp32 oracle": (_by_id(x32, i32), _by_id(v32, i32))}
    gmp = gpu_mod.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
    for exact in (True, False):
        with gpu_mod.Gpu(0) as g:
            g.mix_init(gpu_mod.pbf_params(m.fluid, exact=exact), gmp, pos, ph)
            g.pbf_step(np.float32(m.fluid.dt), 1)
            gx, gv, gid = g.pbf_download(order=0)
        pairs["gpu exact" if exact else "gpu fused"] = (_by_id(gx, gid), _by_id(gv, gid))
    _check(_report(f"cfg1 mixed ({n} particles) after 1 step", x64, v64, pairs, m.fluid.dt))
