"""The launch autotuner: choice per size bucket, cache, cell-size parameters
(host logic only; the GPU tests check that every candidate gives the same
results)."""
import numpy as np
import pytest

from lustrine_amd import autotune as T
from lustrine_amd import gpu as G
from lustrine_amd import scenes


class FakeGpu:
    """Times a candidate from a table instead of running it."""

    def __init__(self, cost):
        self.cost = cost
        self.cur = None
        self.launch = (None, None, None, None)

    def pbf_set_launch(self, block):
        self.launch = (block,) + self.launch[1:]

    def pbf_set_tiling(self, tiled, zt):
        self.launch = self.launch[:1] + (tiled, zt) + self.launch[3:]

    def pbf_set_sort_tile(self, cells):
        self.launch = self.launch[:3] + (cells,)

    def pbf_set_graph(self, on):
        pass

    def pbf_bench(self, dt, warmup, steps, phases=False):
        return self.cost(self.cur, self.launch) * steps, None, None


def test_tuner_picks_fastest_and_caches(tmp_path):
    seen = []

    def cost(cand, launch):
        # fastest: block 128, cell 1.125, not tiled, sort tile 4096
        t = 1.0 + 0.1 * abs(launch[0] - 128) / 64 + (0.0 if cand.cell_scale > 1 else 0.05) + (0.5 if launch[1] else 0)
        return t + (0.0 if launch[3] == 4096 else 0.01)

    g = FakeGpu(cost)

    def init(cand):
        g.cur = cand
        seen.append(cand)

    path = str(tmp_path / "tune.json")
    tu = T.Tuner(path)
    best, table = tu.tune(g, init, 3_000_000, np.float32(1 / 60))
    assert best == T.Launch(128, 1.125, False, 32, 4096)
    # the launch candidates at the default sort tile, then the other sort
    # tiles for the fastest of them
    per = len(T.CANDIDATES) + len(T.SORT_TILES) - 1
    assert len(table) == per == len(seen)
    assert [c.sort_tile for c, _ in table[len(T.CANDIDATES):]] == [1024, 4096]
    assert min(t for _, t in table) == pytest.approx(1.0)
    # cached per bucket: same bucket, no retiming
    best2, table2 = tu.tune(g, init, 2_500_000, np.float32(1 / 60))
    assert best2 == best and table2 == [] and len(seen) == per
    # another bucket is timed
    tu.tune(g, init, 10_000, np.float32(1 / 60))
    assert len(seen) == 2 * per
    # persisted
    tu2 = T.Tuner(path)
    assert tu2.choice(3_100_000) == best and tu2.choice(9_000) is not None


def test_bucket_and_cell_params():
    assert T.bucket(1) == 0 and T.bucket(4096) == 12 and T.bucket(4095) == 11
    cfg = scenes.cfg2()
    p = G.pbf_params(cfg)
    q = T.cell_params(p, cfg.box, 1.125)
    assert q.cell_size == pytest.approx(2.25)
    assert (q.gridX, q.gridY, q.gridZ) == (int(256 / 2.25) + 1, int(128 / 2.25) + 1, int(256 / 2.25) + 1)
    assert q.h == p.h and q.rest_density == p.rest_density and q.iterations == p.iterations
    assert p.cell_size == pytest.approx(2.0)  # the base is not modified
