"""GPU path against the CPU oracle on identical inputs, through the C ABI.

Bar: bit-exact for cell ids, the counting-sort permutation and the neighbor
lists.  Exact arithmetic mode follows the oracle's operation order with no FMA
contraction and IEEE fp32 divide and sqrt: its positions and velocities are
compared bit for bit.  Fused mode (FMA, hardware reciprocal square root) is
held to 1e-5 relative on positions (max |gpu - cpu| over the array divided by
max |cpu|) and to a few ulp of the largest coordinate over dt on velocities,
the size its position rounding takes through v = (x* - x) / dt.
"""
import ctypes

import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu

REL = 1e-5


def rel_err(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) / scale


def max_ulp(a, b):
    ai = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return int(np.abs(ai - bi).max()) if ai.size else 0


def ulp_of_max(x):
    return float(np.spacing(np.float32(np.abs(x).max())))


def glibc_rand_keys(num, mod):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    return np.array([libc.rand() % mod for _ in range(num)], np.int32)


# ------------------------------------------------------------------ grid + sort
def test_cell_ids(gpu_mod, oracle):
    s = oracle.init_constants(64, 64, 64, 0.5, 1.0, 4.0)
    g = gpu_mod.grid(s.cell_size, s.gridX, s.gridY, s.gridZ)
    rng = np.random.default_rng(0)
    pts = rng.uniform(0.5, 63.5, (100000, 3)).astype(np.float32)
    k = np.arange(1, 32, dtype=np.float32) * np.float32(2.0)
    faces = np.stack([k, np.nextafter(k, np.float32(0)), k], axis=1)
    pts = np.concatenate([pts, faces])
    assert np.array_equal(gpu_mod.cell_ids(g, pts), oracle.cell_ids(s, pts))


def test_counting_sort_reference_unit_test(gpu_mod, oracle):
    # experiments/unit_tests/main.cpp:46-89
    keys = glibc_rand_keys(20000, 1987)
    perm, counts = gpu_mod.counting_sort(keys, 1987)
    operm, ocounts = oracle.counting_sort(keys, 1987)
    assert np.array_equal(perm, operm)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(keys[perm], np.sort(keys))


# the last two: 4096 scan tiles of 2048 cells (down blocks sum the tile sums
# themselves) and 4395 tiles (the tile sums are scanned by a separate kernel)
@pytest.mark.parametrize("n,c", [(0, 7), (1, 1), (5000, 1), (33, 2), (100000, 4096),
                                 (200000, 8388607), (200000, 9000001)])
def test_counting_sort_edges(gpu_mod, oracle, n, c):
    keys = np.random.default_rng(n).integers(0, c, n).astype(np.int32)
    perm, counts = gpu_mod.counting_sort(keys, c)
    operm, ocounts = oracle.counting_sort(keys, c)
    assert np.array_equal(perm, operm) and np.array_equal(counts, ocounts)


def test_counting_sort_full_size_cells(gpu_mod, oracle):
    # the 1M-particle scene's predicted cell keys over its 1,081,665 cells
    cfg = scenes.cfg2()
    s = oracle.init_constants(*cfg.box, 0.5, 1.0, 4.0)
    keys = oracle.cell_ids(s, cfg.positions())
    perm, counts = gpu_mod.counting_sort(keys, s.num_grid_cells)
    operm, ocounts = oracle.counting_sort(keys, s.num_grid_cells)
    assert np.array_equal(perm, operm) and np.array_equal(counts, ocounts)


# ------------------------------------------------------------------ neighbor lists
def _scene_with_solids(n_side=10, seed=4, box=20):
    pos = scenes.jittered_lattice(n_side, n_side, n_side, origin=(2.0, 3.0, 2.0), amplitude=0.25, seed=seed)
    sol = scenes.lattice(box, 2, box, origin=(0.0, 0.0, 0.0))
    return pos, sol


def test_fluid_neighbor_lists_exact(gpu_mod, oracle):
    pos, sol = _scene_with_solids()
    s = oracle.init_constants(20, 20, 20, 0.5, 1.0, 4.0)
    n, ns = pos.shape[0], sol.shape[0]
    total = 8000 + 0
    total = max(total, n + ns)
    ps = np.zeros((total, 3), np.float32)
    ps[:n] = pos
    solid_slots = np.arange(total - ns, total, dtype=np.int32)
    ps[solid_slots] = sol
    ooff, oidx = oracle.neighbors_fluid(s, ps, (0, n), (total - ns, total))
    g = gpu_mod.grid(s.cell_size, s.gridX, s.gridY, s.gridZ)
    off, idx = gpu_mod.neighbors_fluid(g, s.kernelRadius, pos, 0, sol, solid_slots)
    assert np.array_equal(off, ooff)
    assert np.array_equal(idx, oidx)


# ------------------------------------------------------------------ PBF substep
def _pbf_compare(gpu_mod, oracle, cfg, steps=1, K=None, exact=True):
    x0 = cfg.positions()
    v0 = np.zeros_like(x0)
    p = oracle.pbf_params(cfg, iterations=K)
    x, v = x0.copy(), v0.copy()
    ids = np.arange(x.shape[0], dtype=np.int32)
    for _ in range(steps):
        omc = oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, iterations=K, exact=exact), x0, v0)
        g.pbf_step(np.float32(cfg.dt), steps)
        gx, gv, gid = g.pbf_download(order=0)
        gmc = g.pbf_last_max_constraint()
    return (gx, gv, gid, gmc), (x, v, ids, omc)


@pytest.mark.parametrize("K", [0, 1, 4])
def test_pbf_one_step_small(gpu_mod, oracle, K):
    cfg = scenes.small(12, 12, 12, box=(24, 24, 24))
    (gx, gv, gid, gmc), (x, v, ids, omc) = _pbf_compare(gpu_mod, oracle, cfg, K=K)
    assert np.array_equal(gid, ids)
    print(f"K={K} ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)
    if K:
        assert gmc == omc


# particle counts that are not multiples of a wave (64) or of a block (256):
# every kernel's tail lanes, the XCD block remap's padding blocks and the
# neighbour-list runs of a partial 64-slot group
RAGGED = [((1, 1, 1), (16, 16, 16)), ((1, 1, 2), (16, 16, 16)), ((3, 3, 7), (16, 16, 16)),
          ((5, 13, 1), (16, 16, 16)), ((10, 10, 10), (16, 16, 16)), ((17, 17, 17), (24, 24, 24))]


@pytest.mark.parametrize("lat,box", RAGGED)
def test_pbf_ragged_sizes_exact_bits(gpu_mod, oracle, lat, box):
    cfg = scenes.small(*lat, box=box, seed=sum(lat))
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg, steps=3)
    assert np.array_equal(gid, ids)
    assert np.array_equal(gx, x) and np.array_equal(gv, v), (max_ulp(gx, x), max_ulp(gv, v))


@pytest.mark.parametrize("lat,box", RAGGED[2:])
def test_pbf_ragged_sizes_fused_and_graph(gpu_mod, oracle, lat, box):
    cfg = scenes.small(*lat, box=box, seed=sum(lat))
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg, steps=3, exact=False)
    assert np.array_equal(gid, ids)
    assert rel_err(gx, x) <= REL
    # three steps under hipGraph replay (one captured pair plus one plain step)
    # give the plain launches' bits
    x0 = cfg.positions()
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0, None)
        g.pbf_set_graph(True)
        g.pbf_step(np.float32(cfg.dt), 3)
        hx, hv, hid = g.pbf_download(order=0)
    assert np.array_equal(hid, gid) and np.array_equal(hx, gx) and np.array_equal(hv, gv)


def test_pbf_one_step_dam_break_32k(gpu_mod, oracle):
    # the 32,768-particle dam break, one step from the initial state: exact
    # mode follows the oracle's operation order, so every bit agrees
    cfg = scenes.cfg0()
    (gx, gv, gid, gmc), (x, v, ids, omc) = _pbf_compare(gpu_mod, oracle, cfg)
    assert np.array_equal(gid, ids)
    print(f"32k ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)
    assert gmc == omc


def test_pbf_ten_steps_32k(gpu_mod, oracle):
    # after 10 steps the two paths still agree bit for bit
    cfg = scenes.cfg0()
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg, steps=10)
    assert np.array_equal(gid, ids)
    print(f"32k x10 ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)


def test_pbf_fp64_deviation_32k(gpu_mod, oracle):
    # the FP32 GPU step against the FP64 oracle (reported; bound is loose)
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    p = oracle.pbf_params(cfg)
    x, v = x0.astype(np.float64), np.zeros((x0.shape[0], 3), np.float64)
    ids = np.arange(x0.shape[0], dtype=np.int32)
    oracle.pbf_step(p, x, v, ids, cfg.dt, fp64=True)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
        g.pbf_step(np.float32(cfg.dt), 1)
        gx, gv, gid = g.pbf_download(order=0)
    assert np.array_equal(gid, ids)
    ex, ev = rel_err(gx, x), rel_err(gv, v)
    print(f"32k fp64 deviation: x {ex:.3e} v {ev:.3e}")
    # v = (x* - x) / dt turns the fp32 rounding of positions near 30 (ulp 2e-6)
    # into ~1e-4 absolute on velocities of ~0.3: the bound is 1e-3 relative
    assert ex < 1e-5 and ev < 1e-3


def test_pbf_full_size_one_step_1m(gpu_mod, oracle):
    # the 1,048,576-particle scene: one step against the oracle, plus
    # size-independent properties (ids a permutation, cell order, box)
    cfg = scenes.cfg2()
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg)
    assert np.array_equal(np.sort(gid), np.arange(cfg.n))
    assert np.array_equal(gid, ids)
    print(f"1m ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)
    assert np.isfinite(gx).all() and np.isfinite(gv).all()
    lo, hi = 0.5, np.array(cfg.box, np.float32) - 0.5
    assert (gx >= lo).all() and (gx <= hi).all()


@pytest.mark.parametrize("scene", ["small", "cfg0", "cfg2"])
def test_pbf_fast_mode_one_step(gpu_mod, oracle, scene):
    # fused fast path: within the 1e-5 relative bar of the FP32 oracle, same
    # cell order (its neighbour test, by fused multiply-adds, differs from the
    # exact one only for pairs within rounding of r = h)
    cfg = {"small": lambda: scenes.small(12, 12, 12, box=(24, 24, 24)), "cfg0": scenes.cfg0,
           "cfg2": scenes.cfg2}[scene]()
    (gx, gv, gid, gmc), (x, v, ids, omc) = _pbf_compare(gpu_mod, oracle, cfg, exact=False)
    assert np.array_equal(gid, ids)
    ex, ev = rel_err(gx, x), rel_err(gv, v)
    print(f"fast {scene}: x {ex:.2e} v {ev:.2e} ulp(x)={max_ulp(gx, x)}")
    assert ex <= REL
    # v = (x* - x) / dt: a position that differs by k ulp of the largest
    # coordinate gives a velocity k ulp / dt off.  One step of fused
    # arithmetic moves positions by at most 2 ulp; the bar is 4 ulp / dt.
    assert np.abs(gv - v).max() <= 4 * ulp_of_max(x) / np.float32(cfg.dt)
    assert abs(gmc - omc) <= 1e-4 * max(1.0, abs(omc))


def test_pbf_fast_mode_ten_steps_32k(gpu_mod, oracle):
    # ten steps let the fused rounding compound through the neighbours: the
    # positions stay within the 1e-5 bar, the velocities of the last step
    # within 64 ulp / dt (the last step's position difference over dt)
    cfg = scenes.cfg0()
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg, steps=10, exact=False)
    assert np.array_equal(gid, ids)
    assert rel_err(gx, x) <= REL
    dv = np.abs(gv - v).max() / (ulp_of_max(x) / np.float32(cfg.dt))
    print(f"fast 32k x10: v {dv:.1f} ulp/dt")
    assert dv <= 64


def test_pbf_dense_cluster_rows_over_64(gpu_mod, oracle):
    # a block at 8x rest density: stencil rows hold > 64 slots, which takes the
    # re-test path instead of the 64-bit row masks; the step must still match
    cfg = scenes.small(8, 8, 8, box=(24, 24, 24), origin=(8.0, 8.0, 8.0))
    x0 = scenes.jittered_lattice(8, 8, 8, cell=0.5, origin=(8.0, 8.0, 8.0), amplitude=0.01, seed=3)
    p = oracle.pbf_params(cfg)
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(x0.shape[0], dtype=np.int32)
    oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
        g.pbf_step(np.float32(cfg.dt), 1)
        gx, gv, gid = g.pbf_download(order=0)
    assert np.array_equal(gid, ids)
    print(f"dense ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)


def test_pbf_neighbour_list_overflow_mixed(gpu_mod, oracle):
    # a block at rest spacing (about 30 neighbours, stored lists) around a core
    # at spacing 0.6 (about 150 neighbours, more than the 128-entry list: the
    # re-test path), so waves hold both kinds; exact mode, bit for bit
    outer = scenes.jittered_lattice(16, 12, 16, cell=1.0, origin=(4.0, 4.0, 4.0), amplitude=0.05, seed=11)
    c = np.array([12.0, 10.0, 12.0], np.float32)
    keep = np.abs(outer - c).max(axis=1) > 2.6
    core = scenes.jittered_lattice(8, 8, 8, cell=0.6, origin=tuple(c - 2.1), amplitude=0.02, seed=12)
    x0 = np.concatenate([outer[keep], core]).astype(np.float32)
    cfg = scenes.small(8, 8, 8, box=(28, 24, 28))
    p = oracle.pbf_params(cfg)
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(x0.shape[0], dtype=np.int32)
    for _ in range(2):
        oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
        g.pbf_step(np.float32(cfg.dt), 2)
        gx, gv, gid = g.pbf_download(order=0)
    assert np.array_equal(gid, ids)
    assert np.array_equal(gx, x) and np.array_equal(gv, v)


def test_pbf_list_entries_beyond_16_bits(gpu_mod, oracle):
    # a slab three cell columns wide (x), one cell high (y) and 4,200 cells deep
    # (z): in cell order a particle of the middle column has neighbours in the
    # columns either side whose slots lie one column's population (33,600)
    # before and after, 67,200 apart -- more than a 16-bit list entry holds
    # past its layer's first neighbour.  Those particles re-test their
    # candidates in every pass while the outer columns walk lists; exact mode,
    # bit for bit over two steps.
    cfg = scenes.small(6, 2, 8400, box=(6, 8, 16800))
    x0 = cfg.positions()
    cell = np.floor(x0 / np.float32(cfg.h)).astype(np.int64)
    assert set(cell[:, 1]) == {0} and set(cell[:, 0]) == {0, 1, 2}
    pop = np.bincount(cell[:, 0])
    assert pop[0] + pop[1] > 65535, pop
    p = oracle.pbf_params(cfg)
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(x0.shape[0], dtype=np.int32)
    for _ in range(2):
        oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
        g.pbf_step(np.float32(cfg.dt), 1)
        retest = g.pbf_retest_count()
        g.pbf_step(np.float32(cfg.dt), 1)
        gx, gv, gid = g.pbf_download(order=0)
    assert np.array_equal(gid, ids)
    assert np.array_equal(gx, x) and np.array_equal(gv, v), (max_ulp(gx, x), max_ulp(gv, v))
    # first step: the middle column's particles whose neighbours span more
    # than 65,535 slots (about half of them, 16,819 measured) are counted
    mid = int(pop[1])
    assert 0.25 * mid <= retest <= mid, (retest, mid)


def test_pbf_retest_count_zero_on_dam_break(gpu_mod):
    cfg = scenes.cfg0()
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=False), cfg.positions(), None)
        g.pbf_step(np.float32(cfg.dt), 5)
        assert g.pbf_retest_count() == 0


def test_pbf_original_order_download(gpu_mod):
    cfg = scenes.small(10, 10, 10, box=(20, 20, 20))
    x0 = cfg.positions()
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
        g.pbf_step(np.float32(cfg.dt), 3)
        sx, sv, sid = g.pbf_download(order=0)
        ox, ov, oid = g.pbf_download(order=1)
    assert np.array_equal(oid, np.arange(cfg.n))
    assert np.array_equal(ox[sid], sx) and np.array_equal(ov[sid], sv)


# ------------------------------------------------------------------ sand (simulate_sand)
def _sand_compare(gpu_mod, oracle, pos, sol, steps, box=(20, 20, 20), flags=None, h_scale=3.1):
    s = oracle.init_constants(*box, 0.5, 1.0, h_scale)
    st = oracle.SandStream(s, pos, sol)
    n, ns = pos.shape[0], (0 if sol is None else sol.shape[0])
    total = st.total
    solid_index = np.arange(total - 1, total - 1 - ns, -1, dtype=np.int32)  # written downward
    sp = gpu_mod.sand_params(*box, kernel_radius_scale=h_scale)
    sp.ptr_solid_start = total - ns if ns else total - 1
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos, None, None, sol, solid_index)
        for k in range(steps):
            fl = flags[k] if flags else (False, False, None)
            st.st.attract_flag, st.st.blow_flag = int(fl[0]), int(fl[1])
            if fl[2] is not None:
                st.st.player_position[:] = list(fl[2])
            st.step(np.float32(0.016))
            g.sand_step(np.float32(0.016), fl[0], fl[1], fl[2])
        got = g.sand_download()
        goff, gidx = g.sand_neighbors()
    want = (st.p[:n], st.ps[:n], st.v[:n], st.att[:n], st.perm())
    ooff, oidx = st.neighbors()
    return got, want, (goff, gidx), (ooff, oidx)


def test_sand_steps_match_reference(gpu_mod, oracle):
    pos, sol = _scene_with_solids(n_side=12, seed=7)
    got, want, (goff, gidx), (ooff, oidx) = _sand_compare(gpu_mod, oracle, pos, sol, steps=5)
    assert np.array_equal(got[4], want[4])  # last counting-sort permutation
    assert np.array_equal(goff, ooff) and np.array_equal(gidx, oidx)  # neighbor lists, in order
    assert np.array_equal(got[3], want[3])
    for a, b in zip(got[:3], want[:3]):
        assert rel_err(a, b) <= REL
    print(f"sand x5 ulp(x)={max_ulp(got[0], want[0])} ulp(v)={max_ulp(got[2], want[2])}")


def test_sand_dense_clump_long_lists(gpu_mod, oracle):
    # a clump at spacing 0.3 (about 200 neighbours within h, lists of 216
    # entries) beside a sparse layer (lists of 5 to 20), over solids
    clump = scenes.jittered_lattice(6, 6, 6, cell=0.3, origin=(8.0, 8.0, 8.0), amplitude=0.02, seed=11)
    layer = scenes.jittered_lattice(10, 2, 10, origin=(2.0, 3.0, 2.0), amplitude=0.2, seed=12)
    pos = np.concatenate([layer, clump]).astype(np.float32)
    sol = scenes.lattice(20, 2, 20, origin=(0.0, 0.0, 0.0))
    # one step: the lists are those of the first prediction, the clump intact
    got, want, (goff, gidx), (ooff, oidx) = _sand_compare(gpu_mod, oracle, pos, sol, steps=1)
    assert int(np.diff(ooff).max()) > 64
    assert np.array_equal(got[4], want[4])
    assert np.array_equal(goff, ooff) and np.array_equal(gidx, oidx)
    for a, b in zip(got[:3], want[:3]):
        assert rel_err(a, b) <= REL


@pytest.mark.parametrize("lat", [(1, 1, 1), (1, 1, 2), (5, 13, 1)])
def test_sand_ragged_sizes(gpu_mod, oracle, lat):
    # 1, 2 and 65 sand particles over solids: tail lanes of every sand kernel
    pos = scenes.jittered_lattice(*lat, origin=(6.0, 3.0, 6.0), amplitude=0.2, seed=sum(lat))
    sol = scenes.lattice(20, 2, 20, origin=(0.0, 0.0, 0.0))
    got, want, (goff, gidx), (ooff, oidx) = _sand_compare(gpu_mod, oracle, pos, sol, steps=3)
    assert np.array_equal(got[4], want[4])
    assert np.array_equal(goff, ooff) and np.array_equal(gidx, oidx)
    for a, b in zip(got[:3], want[:3]):
        assert rel_err(a, b) <= REL


def test_sand_attract_and_blow_flags(gpu_mod, oracle):
    pos, sol = _scene_with_solids(n_side=8, seed=9)
    player = np.array([6.0, 6.0, 6.0], np.float32)
    flags = [(True, False, player), (True, False, player), (False, True, player), (False, False, player)]
    got, want, (goff, gidx), (ooff, oidx) = _sand_compare(gpu_mod, oracle, pos, sol, steps=4, flags=flags)
    assert np.array_equal(got[4], want[4])
    assert np.array_equal(gidx, oidx)
    assert np.array_equal(got[3], want[3])
    for a, b in zip(got[:3], want[:3]):
        assert rel_err(a, b) <= REL


def test_sand_no_solids(gpu_mod, oracle):
    pos = scenes.jittered_lattice(10, 10, 10, origin=(3.0, 1.0, 3.0), amplitude=0.2, seed=2)
    got, want, (goff, gidx), (ooff, oidx) = _sand_compare(gpu_mod, oracle, pos, None, steps=3)
    assert np.array_equal(got[4], want[4]) and np.array_equal(gidx, oidx)
    for a, b in zip(got[:3], want[:3]):
        assert rel_err(a, b) <= REL


def test_pbf_tail_one_step_4m(gpu_mod, oracle):
    # the 4,194,304-particle tail (256 x 64 x 256 in 512 x 128 x 512, 4.29 M
    # cells): one exact-mode step bit for bit against the oracle, then a fused
    # step's size-independent properties (ids a permutation, cell order,
    # every position in the clamp box and finite)
    cfg = scenes.cfg2_tail()
    (gx, gv, gid, _), (x, v, ids, _) = _pbf_compare(gpu_mod, oracle, cfg)
    assert np.array_equal(gid, ids)
    print(f"4m ulp(x)={max_ulp(gx, x)} ulp(v)={max_ulp(gv, v)}")
    assert np.array_equal(gx, x) and np.array_equal(gv, v)
    x0 = cfg.positions()
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0, None)
        g.pbf_step(np.float32(cfg.dt), 3)
        fx, fv, fid = g.pbf_download(order=0)
    assert np.array_equal(np.sort(fid), np.arange(cfg.n))
    lo, hi = np.float32(0.5), np.array(cfg.box, np.float32) - np.float32(0.5)
    assert np.isfinite(fx).all() and np.isfinite(fv).all()
    assert (fx >= lo).all() and (fx <= hi).all()
    gxs, gys, gzs = cfg.grid
    cs = np.float32(cfg.cell_size)
    key = ((fx[:, 1] / cs).astype(np.int64) * gxs * gzs + (fx[:, 0] / cs).astype(np.int64) * gzs
           + (fx[:, 2] / cs).astype(np.int64))
    # slots were sorted by the cells of the step's predicted positions; the
    # final positions stay within a cell or so of them: the cell keys of the
    # committed positions are nearly sorted
    assert np.mean(np.diff(key) >= 0) > 0.95


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("scene", ["small", "sheet", "dense", "cfg0", "cfg2"])
def test_pbf_staged_build_matches_gathers(gpu_mod, scene, exact):
    # the wave-staged list build (lus_pbf_set_build 1, the default) against
    # the per-lane gather build: the same bits after three steps.  "sheet" (a
    # 2-particle-thick layer: a wave spans dozens of cell columns) and
    # "dense" (8x rest density: rows over 64 slots, ranges over the staging
    # capacity) take the gather fallback for some rows
    cfg = {"small": lambda: scenes.small(12, 12, 12, box=(24, 24, 24)),
           "sheet": lambda: scenes.small(40, 2, 40, box=(48, 8, 48), seed=4),
           "dense": lambda: scenes.small(8, 8, 8, box=(24, 24, 24), origin=(8.0, 8.0, 8.0)),
           "cfg0": scenes.cfg0, "cfg2": scenes.cfg2}[scene]()
    x0 = cfg.positions()
    if scene == "dense":
        x0 = scenes.jittered_lattice(8, 8, 8, cell=0.5, origin=(8.0, 8.0, 8.0), amplitude=0.01, seed=3)
    out = []
    with gpu_mod.Gpu(0) as g:
        for staged in (False, True):
            g.pbf_init(gpu_mod.pbf_params(cfg, exact=exact), x0)
            g.pbf_set_build(staged)
            g.pbf_step(np.float32(cfg.dt), 3)
            out.append(g.pbf_download(order=0) + (g.pbf_last_max_constraint(), g.pbf_retest_count()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_pbf_four_lane_walks(gpu_mod, oracle):
    # fused steps of a small scene walk each list with four lanes (every
    # fourth group of eight entries, partial sums combined in a fixed order):
    # against the one-lane walks and the oracle, within the fused bars; a
    # dense block (lists past 128 entries: the first lane's re-test walk)
    # against the oracle
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    out = []
    with gpu_mod.Gpu(0) as g:
        for lanes in (1, 4):
            g.pbf_init(gpu_mod.pbf_params(cfg, exact=False), x0)
            g.pbf_set_lanes(lanes)
            g.pbf_step(np.float32(cfg.dt), 3)
            out.append(g.pbf_download(order=1))
    x1, v1, _ = out[0]
    x4, v4, _ = out[1]
    assert rel_err(x4, x1) <= 1e-6
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(cfg.n, dtype=np.int32)
    p = oracle.pbf_params(cfg)
    for _ in range(3):
        oracle.pbf_step(p, x, v, ids, np.float32(cfg.dt))
    xo, vo = np.zeros_like(x), np.zeros_like(v)
    xo[ids], vo[ids] = x, v
    assert rel_err(x4, xo) <= REL
    print(f"4 lanes 32k x3: x vs 1 lane {rel_err(x4, x1):.2e}, vs oracle {rel_err(x4, xo):.2e}")

    dcfg = scenes.small(8, 8, 8, box=(24, 24, 24), origin=(8.0, 8.0, 8.0))
    d0 = scenes.jittered_lattice(8, 8, 8, cell=0.5, origin=(8.0, 8.0, 8.0), amplitude=0.01, seed=3)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(dcfg, exact=False), d0)
        g.pbf_set_lanes(4)
        g.pbf_step(np.float32(dcfg.dt), 1)
        gx, gv, gid = g.pbf_download(order=0)
        retest = g.pbf_retest_count()
    x, v = d0.copy(), np.zeros_like(d0)
    ids = np.arange(d0.shape[0], dtype=np.int32)
    oracle.pbf_step(oracle.pbf_params(dcfg), x, v, ids, np.float32(dcfg.dt))
    assert retest > 0
    assert np.array_equal(gid, ids)
    assert rel_err(gx, x) <= REL


def test_pbf_keys_from_xsph_match_key_kernel(gpu_mod):
    # the XSPH pass writing the next step's cell keys (default) against every
    # step running the key kernel: the same bits, with graphs and without, on
    # the 32K scene (four-lane passes), a 1,000-particle scene with one-lane
    # passes, and a batch of scenes; a call with another dt in between runs
    # the key kernel again
    cfg = scenes.cfg0()
    small = scenes.small(10, 10, 10, box=(20, 20, 20))
    b = scenes.BatchConfig(name="t", scenes=[scenes.dam_break_of_size(s, 70 + k)
                                             for k, s in enumerate([900, 3000, 1700])])
    for case in ("cfg0", "small", "batch"):
        out = []
        for fused in (False, True):
            for graph in (False, True):
                with gpu_mod.Gpu(0) as g:
                    if case == "batch":
                        g.pbf_batch_init(gpu_mod.pbf_params(b.scenes[0], exact=True), b.counts(), b.boxes(),
                                         b.positions())
                        dt = b.scenes[0].dt
                    else:
                        c = cfg if case == "cfg0" else small
                        g.pbf_init(gpu_mod.pbf_params(c, exact=(case != "cfg0")), c.positions())
                        dt = c.dt
                        if case == "small":
                            g.pbf_set_lanes(1)
                    g.pbf_set_key_fusion(fused)
                    g.pbf_set_graph(graph)
                    g.pbf_step(np.float32(dt), 5)
                    g.pbf_step(np.float32(dt * 0.5), 3)
                    g.pbf_step(np.float32(dt), 4)
                    out.append(g.pbf_download(order=0))
        for x, v, ids in out[1:]:
            assert np.array_equal(ids, out[0][2]) and np.array_equal(x, out[0][0]) and np.array_equal(v, out[0][1])
