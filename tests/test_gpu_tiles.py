"""Tiled neighbour passes (LDS staging of each tile's 27-cell neighbourhood)
against the one-thread-per-slot passes: the same visiting order and the same
arithmetic, so the same bits, for every chunk length, in both arithmetic
modes, with learned steps and with batched scenes."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _run(gpu_mod, cfg, x0, steps, exact, tiled, zt=16, graph=False):
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=exact), x0)
        g.pbf_set_lanes(1)  # the tiled passes walk with one lane per particle
        g.pbf_set_tiling(tiled, zt)
        g.pbf_set_graph(graph)
        g.pbf_step(np.float32(cfg.dt), steps)
        x, v, ids = g.pbf_download(order=0)
        return x, v, ids, g.pbf_last_max_constraint()


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("zt", [1, 7, 16, 32])
def test_tiled_equals_flat(gpu_mod, exact, zt):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    fx, fv, fid, fmc = _run(gpu_mod, cfg, x0, 6, exact, False)
    tx, tv, tid, tmc = _run(gpu_mod, cfg, x0, 6, exact, True, zt, graph=True)
    assert np.array_equal(fid, tid)
    assert np.array_equal(fx, tx) and np.array_equal(fv, tv)
    assert fmc == tmc


def test_tiled_dense_cluster_and_overflow_rows(gpu_mod):
    # a compressed block: stencil rows over 64 slots (overflow path) and
    # tiles whose staged rows exceed the LDS capacity (global fallback)
    cfg = scenes.small(12, 12, 12, box=(24, 24, 24), seed=3)
    x0 = (cfg.positions() * np.float32(0.35) + np.float32(4.0)).astype(np.float32)
    fx, fv, _, _ = _run(gpu_mod, cfg, x0, 3, True, False)
    tx, tv, _, _ = _run(gpu_mod, cfg, x0, 3, True, True, 32)
    assert np.array_equal(fx, tx) and np.array_equal(fv, tv)


def test_tiled_matches_oracle_one_step(gpu_mod, oracle):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    tx, tv, tid, _ = _run(gpu_mod, cfg, x0, 1, True, True)
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(cfg.n, dtype=np.int32)
    oracle.pbf_step(oracle.pbf_params(cfg), x, v, ids, np.float32(cfg.dt))
    assert np.array_equal(tid, ids)
    assert np.array_equal(tx, x) and np.array_equal(tv, v)


def test_tiled_learned_step_equals_flat(gpu_mod):
    from oracle import mlp as M
    cfg = scenes.small(16, 12, 16, box=(32, 24, 32), seed=4)
    x0 = cfg.positions()
    net = np.concatenate([np.zeros(32), np.ones(32), M.random_params(np.random.default_rng(5)), [0.02]])
    out = []
    for tiled in (False, True):
        with gpu_mod.Gpu(0) as g:
            g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
            g.pbf_set_network(net.astype(np.float32))
            g.pbf_set_mode(learned=True, learned_iterations=1, bound=np.inf)
            g.pbf_set_tiling(tiled, 16)
            g.pbf_step(np.float32(cfg.dt), 4)
            x, v, _ = g.pbf_download(order=1)
            _, _, errs, errm = g.pbf_learned_stats(with_mean=True)
        out.append((x, v, errs, errm))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert np.allclose(out[0][3], out[1][3], rtol=1e-5)


def test_tiled_batch_equals_flat(gpu_mod):
    b = scenes.BatchConfig(name="t", scenes=[scenes.dam_break_of_size(s, 40 + k)
                                             for k, s in enumerate([500, 3000, 1200, 7000])])
    res = []
    for tiled in (False, True):
        with gpu_mod.Gpu(0) as g:
            g.pbf_batch_init(gpu_mod.pbf_params(b.scenes[0], exact=True), b.counts(), b.boxes(), b.positions())
            g.pbf_set_tiling(tiled, 16)
            g.pbf_step(np.float32(b.scenes[0].dt), 3)
            res.append(g.pbf_download(order=1)[:2])
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_set_tiling_rejects_bad_chunk(gpu_mod):
    with gpu_mod.Gpu(0) as g:
        with pytest.raises(gpu_mod.LusError):
            g.pbf_set_tiling(True, 0)
        with pytest.raises(gpu_mod.LusError):
            g.pbf_set_tiling(True, 33)


@pytest.mark.parametrize("block", [64, 128])
@pytest.mark.parametrize("exact", [True, False])
def test_block_sizes_equal_default(gpu_mod, block, exact):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    ref = _run(gpu_mod, cfg, x0, 4, exact, False)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=exact), x0)
        g.pbf_set_lanes(1)  # as _run: the block size of the one-lane walks
        g.pbf_set_launch(block)
        g.pbf_set_graph(True)
        g.pbf_step(np.float32(cfg.dt), 4)
        x, v, ids = g.pbf_download(order=0)
        mc = g.pbf_last_max_constraint()
    assert np.array_equal(ids, ref[2]) and np.array_equal(x, ref[0]) and np.array_equal(v, ref[1])
    assert mc == ref[3]
    with gpu_mod.Gpu(0) as g:
        with pytest.raises(gpu_mod.LusError):
            g.pbf_set_launch(96)


def test_batch_with_larger_cells_matches_oracle(gpu_mod, oracle):
    from lustrine_amd import autotune as T
    b = scenes.BatchConfig(name="t", scenes=[scenes.dam_break_of_size(s, 60 + k)
                                             for k, s in enumerate([700, 2600, 1500])])
    base = gpu_mod.pbf_params(b.scenes[0], exact=True)
    with gpu_mod.Gpu(0) as g:
        g.pbf_batch_init(T.cell_params(base, None, 1.125), b.counts(), b.boxes(), b.positions())
        g.pbf_set_launch(128)
        g.pbf_step(np.float32(b.scenes[0].dt), 3)
        gx, gv, _ = g.pbf_download(order=1)
    off = 0
    for s in b.scenes:
        p = oracle.pbf_params(s)
        p.cell_size = 2.25
        p.gridX, p.gridY, p.gridZ = (int(np.float32(d) / np.float32(2.25)) + 1 for d in s.box)
        x, v = s.positions(), np.zeros((s.n, 3), np.float32)
        ids = np.arange(s.n, dtype=np.int32)
        for _ in range(3):
            oracle.pbf_step(p, x, v, ids, np.float32(s.dt))
        ox, ov = np.zeros_like(x), np.zeros_like(v)
        ox[ids], ov[ids] = x, v
        assert np.array_equal(gx[off:off + s.n], ox) and np.array_equal(gv[off:off + s.n], ov), s.name
        off += s.n


def test_sort_tiles_equal_default(gpu_mod):
    # the counting sort's prefix scan in tiles of 1024, 2048 or 4096 cells
    # (and the default, 0: chosen by the cell count): the same cell starts,
    # so the same steps bit for bit -- on the 32K scene
    # (18 to 71 tiles) and on 1,000 particles in a 420 x 200 x 420 box
    # (4,496,621 cells: 4,392 tiles of 1024, above the 4,096 tiles up to which
    # the down pass sums the tile sums itself)
    for cfg in (scenes.cfg0(), scenes.small(10, 10, 10, box=(420, 200, 420))):
        x0 = cfg.positions()
        res = []
        for tile in (2048, 1024, 4096, 0):
            with gpu_mod.Gpu(0) as g:
                g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
                g.pbf_set_sort_tile(tile)
                g.pbf_set_graph(True)
                g.pbf_step(np.float32(cfg.dt), 4)
                res.append(g.pbf_download(order=0))
        for x, v, ids in res[1:]:
            assert np.array_equal(ids, res[0][2]) and np.array_equal(x, res[0][0]) and np.array_equal(v, res[0][1])
    with gpu_mod.Gpu(0) as g:
        with pytest.raises(gpu_mod.LusError):
            g.pbf_set_sort_tile(512)
