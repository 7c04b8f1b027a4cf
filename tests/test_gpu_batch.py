"""Batched scenes and hipGraph replay of the PBF substep.

A batch stacks the scene grids along y with an empty cell layer between
scenes; each scene must compute exactly what the oracle computes for it alone
(exact arithmetic, bit for bit).  Graph replay must reproduce eager stepping
on every replay, with downloads and other work between replays."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _small_batch(sizes, seed=7):
    return scenes.BatchConfig(name="t", scenes=[scenes.dam_break_of_size(s, seed * 100 + k)
                                                for k, s in enumerate(sizes)])


def _oracle_scene(oracle, cfg, steps):
    p = oracle.pbf_params(cfg)
    x = cfg.positions()
    v = np.zeros_like(x)
    ids = np.arange(cfg.n, dtype=np.int32)
    for _ in range(steps):
        oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    bx, bv = np.zeros_like(x), np.zeros_like(v)
    bx[ids], bv[ids] = x, v
    return bx, bv


@pytest.mark.parametrize("graph", [False, True])
def test_batch_scenes_match_oracle_alone(gpu_mod, oracle, graph):
    b = _small_batch([300, 1200, 4100, 650, 2500, 900])
    base = b.scenes[0]
    with gpu_mod.Gpu(0) as g:
        g.pbf_batch_init(gpu_mod.pbf_params(base, exact=True), b.counts(), b.boxes(), b.positions())
        g.pbf_set_graph(graph)
        g.pbf_step(np.float32(base.dt), 3)
        gx, gv, _ = g.pbf_download(order=1)
    off = 0
    for s in b.scenes:
        ox, ov = _oracle_scene(oracle, s, 3)
        assert np.array_equal(gx[off:off + s.n], ox), s.name
        assert np.array_equal(gv[off:off + s.n], ov), s.name
        off += s.n


def test_graph_replays_match_eager(gpu_mod):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    dt = np.float32(cfg.dt)
    with gpu_mod.Gpu(0) as ge, gpu_mod.Gpu(0) as gg:
        ge.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0)
        gg.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0)
        gg.pbf_set_graph(True)
        for rep in range(4):  # each pbf_step(2) is one graph replay
            ge.pbf_step(dt, 2)
            gg.pbf_step(dt, 2)
            ex, ev, eid = ge.pbf_download(order=0)
            gx, gv, gid = gg.pbf_download(order=0)
            assert np.array_equal(eid, gid), rep
            assert np.array_equal(ex, gx) and np.array_equal(ev, gv), rep
        # an odd step between replays switches the buffer parity: recaptured
        ge.pbf_step(dt, 3)
        gg.pbf_step(dt, 3)
        gg.pbf_step(dt, 2)
        ge.pbf_step(dt, 2)
        assert np.array_equal(ge.pbf_download(1)[0], gg.pbf_download(1)[0])


def test_batch_full_size_runs(gpu_mod):
    # the 64-scene batch (~4.2 M particles): ids form a permutation, every scene
    # stays inside its own box, nothing is NaN
    b = scenes.cfg4()
    with gpu_mod.Gpu(0) as g:
        g.pbf_batch_init(gpu_mod.pbf_params(b.scenes[0], exact=False), b.counts(), b.boxes(), b.positions())
        g.pbf_set_graph(True)
        g.pbf_step(np.float32(b.scenes[0].dt), 4)
        gx, gv, _ = g.pbf_download(order=1)
        _, _, sid_order = g.pbf_download(order=0)
    assert np.array_equal(np.sort(sid_order), np.arange(b.n))
    assert np.isfinite(gx).all() and np.isfinite(gv).all()
    off = 0
    for s in b.scenes:
        p = gx[off:off + s.n]
        assert (p >= 0.5).all() and (p <= np.array(s.box, np.float32) - 0.5).all(), s.name
        off += s.n
