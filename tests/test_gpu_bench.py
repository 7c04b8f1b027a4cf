"""The timing entry points the benchmark uses (lus_pbf_bench, lus_pbf_step_latency,
lus_pbf_time_kernels): what they count, that their phase times add up, and that
timing changes no result."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _init(gpu_mod, g, cfg, x0, exact):
    g.pbf_init(gpu_mod.pbf_params(cfg, exact=exact), x0, None)


@pytest.mark.parametrize("fused_keys", [True, False])
@pytest.mark.parametrize("exact", [True, False])
def test_phase_run_counts_and_sums(gpu_mod, exact, fused_keys):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    steps, K = 6, cfg.iterations
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, exact)
        g.pbf_set_key_fusion(fused_keys)
        ms, ph, nl = g.pbf_bench(np.float32(cfg.dt), 2, steps, phases=True)
    # one launch per phase per step; the first lambda pass of a step builds the
    # lists; with fused keys every timed step's keys come from the XSPH pass
    # of the step before it (the warm-up's last step included)
    assert nl["predict_key"] == (0 if fused_keys else steps) and nl["cell_scan"] == steps
    assert nl["scatter"] == steps and nl["stable_permute"] == steps
    assert nl["lambda_build"] == steps and nl["lambda"] == (K - 1) * steps
    assert nl["deltap"] == K * steps and nl["xsph"] == steps
    for name in ("features", "mlp", "decide", "tiles"):
        assert nl[name] == 0 and ph[name] == 0.0
    # per-kernel device times: positive, bounded, and *ms is their sum
    for name, n in nl.items():
        if n:
            assert 0.0 < ph[name] / n < 50.0, (name, ph[name], n)
    assert ms == pytest.approx(sum(ph.values()), rel=1e-4)


def test_timing_changes_no_result(gpu_mod):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    dt = np.float32(cfg.dt)
    out = []
    # the same 12 steps: plain steps, a graph-replayed bench, a phase-timed
    # bench, the latency probe, and a phase-timed bench after time_kernels
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, False)
        g.pbf_step(dt, 12)
        out.append(g.pbf_download(order=1)[:2])
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, False)
        g.pbf_set_graph(True)
        g.pbf_bench(dt, 2, 10, phases=False)
        out.append(g.pbf_download(order=1)[:2])
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, False)
        g.pbf_bench(dt, 4, 8, phases=True)
        out.append(g.pbf_download(order=1)[:2])
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, False)
        g.pbf_set_graph(True)
        lat = g.pbf_step_latency(dt, 12)
        assert len(lat) == 12 and all(t > 0.0 for t in lat)
        out.append(g.pbf_download(order=1)[:2])
    with gpu_mod.Gpu(0) as g:
        _init(gpu_mod, g, cfg, x0, False)
        g.pbf_step(dt, 3)
        us = g.pbf_time_kernels(reps=5)
        assert all(0.0 < v < 10000.0 for v in us.values()), us
        g.pbf_bench(dt, 0, 9, phases=True)
        out.append(g.pbf_download(order=1)[:2])
    for pos, vel in out[1:]:
        assert np.array_equal(pos, out[0][0]) and np.array_equal(vel, out[0][1])
