"""Long runs of the benchmark scenes: 1,000 steps of the 32K dam break and 400
of the 1M one (the later, compressed part of the fall, where neighbour lists
grow toward their capacity and some particles take the overflow path) stay
finite, inside the clamp box, a permutation of the particles, with a bounded
density constraint."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,steps", [("cfg0", 1000), ("cfg2", 400)])
def test_pbf_long_run_stays_valid(gpu_mod, name, steps):
    cfg = {"cfg0": scenes.cfg0, "cfg2": scenes.cfg2}[name]()
    x0 = cfg.positions()
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0, None)
        g.pbf_set_graph(True)
        g.pbf_step(np.float32(cfg.dt), steps)
        x, v, ids = g.pbf_download(order=0)
        maxc = g.pbf_last_max_constraint()
    assert np.array_equal(np.sort(ids), np.arange(cfg.n))
    assert np.isfinite(x).all() and np.isfinite(v).all()
    lo, hi = 0.5, np.array(cfg.box, np.float32) - 0.5
    assert (x >= lo).all() and (x <= hi).all()
    # the fluid has settled into a compressed pool, not blown up: the max
    # density constraint of the last lambda pass is finite and bounded
    assert np.isfinite(maxc) and maxc < 5.0, maxc
    print(f"{name} after {steps} steps: max C {maxc:.3f}, max |v| {np.abs(v).max():.2f}")


def test_sand_long_run_stays_valid(gpu_mod):
    # simulate_sand on the 1M lattice for 200 steps: the heap spreads on the
    # floor; positions stay finite and inside [r, D - r]
    cfg = scenes.cfg2()
    pos = cfg.positions()
    sp = gpu_mod.sand_params(*cfg.box)
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos)
        for _ in range(200):
            g.sand_step(np.float32(cfg.dt))
        x, xs, v, att, perm = g.sand_download()
    assert np.array_equal(np.sort(perm), np.arange(pos.shape[0]))
    assert np.isfinite(x).all() and np.isfinite(v).all()
    assert (x >= 0.5).all() and (x <= np.array(cfg.box, np.float32) - 0.5).all()
    print(f"sand after 200 steps: max |v| {np.abs(v).max():.2f}, lowest y {x[:, 1].min():.3f}")
