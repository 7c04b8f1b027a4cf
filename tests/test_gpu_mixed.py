"""Mixed sand + fluid over static solids on the GPU against the oracle's mixed
step (oracle/mixed_core.inc, pinned by tests/test_oracle_mixed.py): exact
arithmetic bit for bit, the fused path within tolerance, the full 262k-particle
configuration one step bit for bit plus invariants over several steps."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, m, pos, ph, steps):
    p = oracle.pbf_params(m.fluid)
    mp = oracle.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
    x, v = pos.copy(), np.zeros_like(pos)
    ids = np.arange(pos.shape[0], dtype=np.int32)
    phs = ph.copy()
    errs = []
    for _ in range(steps):
        errs.append(oracle.mixed_step(p, mp, x, v, ids, phs, np.float32(m.fluid.dt)))
    return x, v, ids, errs


def _gpu_run(gpu_mod, m, pos, ph, steps, exact=True, graph=False, check=False, staged=True):
    mp = gpu_mod.mix_params(mu_s=m.mu_s, mu_k=m.mu_k, collision=m.collision, friction=m.friction)
    with gpu_mod.Gpu(0) as g:
        g.mix_init(gpu_mod.pbf_params(m.fluid, exact=exact), mp, pos, ph)
        g.pbf_set_build(staged)
        if check:
            g.pbf_set_mode(learned=False, check=True)
        g.pbf_set_graph(graph)
        g.pbf_step(np.float32(m.fluid.dt), steps)
        x, v, ids = g.pbf_download(order=0)
        stats = g.pbf_learned_stats(with_mean=True) if check else None
    return x, v, ids, stats


@pytest.mark.parametrize("steps", [1, 5])
def test_mixed_small_matches_oracle_bits(gpu_mod, oracle, steps):
    m = scenes.mixed_small(seed=6)
    pos, ph = m.arrays()
    ox, ov, oid, oerr = _oracle_run(oracle, m, pos, ph, steps)
    gx, gv, gid, st = _gpu_run(gpu_mod, m, pos, ph, steps, check=True)
    assert np.array_equal(gid, oid)
    assert np.array_equal(gx, ox) and np.array_equal(gv, ov)
    n, _, errs, errm = st
    assert n == steps
    assert np.array_equal(errs, np.array([e[0] for e in oerr], np.float32))
    assert np.allclose(errm, [e[1] for e in oerr], rtol=1e-5, atol=1e-9)


def test_mixed_graph_equals_eager(gpu_mod):
    m = scenes.mixed_small(12, 8, 12, box=(28, 24, 28), field=(28, 28), seed=7)
    pos, ph = m.arrays()
    a = _gpu_run(gpu_mod, m, pos, ph, 6, exact=False, graph=False)
    b = _gpu_run(gpu_mod, m, pos, ph, 6, exact=False, graph=True)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_mixed_fused_close_to_oracle(gpu_mod, oracle):
    m = scenes.mixed_small(seed=8)
    pos, ph = m.arrays()
    ox, ov, oid, _ = _oracle_run(oracle, m, pos, ph, 1)
    gx, gv, gid, _ = _gpu_run(gpu_mod, m, pos, ph, 1, exact=False)
    assert np.array_equal(gid, oid)
    scale = np.abs(ox).max()
    assert np.abs(gx - ox).max() <= 1e-5 * scale
    assert np.abs(gv - ov).max() <= 1e-5 * scale / m.fluid.dt


def test_mixed_full_size_one_step_bits(gpu_mod, oracle):
    m = scenes.cfg1()
    pos, ph = m.arrays()
    ox, ov, oid, _ = _oracle_run(oracle, m, pos, ph, 1)
    gx, gv, gid, _ = _gpu_run(gpu_mod, m, pos, ph, 1)
    assert np.array_equal(gid, oid)
    assert np.array_equal(gx, ox) and np.array_equal(gv, ov)


def test_mixed_full_size_invariants(gpu_mod):
    m = scenes.cfg1()
    pos, ph = m.arrays()
    gx, gv, gid, _ = _gpu_run(gpu_mod, m, pos, ph, 10, exact=False, graph=True)
    x = np.zeros_like(gx)
    x[gid] = gx
    v = np.zeros_like(gv)
    v[gid] = gv
    solid = ph == 2
    assert np.array_equal(x[solid], pos[solid]) and not v[solid].any()
    assert np.isfinite(x).all() and np.isfinite(v).all()
    mov = x[~solid]
    assert (mov >= 0.5).all() and (mov <= 127.5).all()
    assert np.array_equal(np.sort(gid), np.arange(pos.shape[0]))


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("scene", ["small", "cfg1"])
def test_mixed_staged_build_matches_gathers(gpu_mod, scene, exact):
    # the wave-staged list build of the mixed scene (phase runs staged with
    # their w, solids without lists) against the per-lane gather build: the
    # same bits after three steps
    m = scenes.mixed_small(12, 8, 12, box=(28, 24, 28), field=(28, 28), seed=9) if scene == "small" else scenes.cfg1()
    pos, ph = m.arrays()
    a = _gpu_run(gpu_mod, m, pos, ph, 3, exact=exact, staged=False)
    b = _gpu_run(gpu_mod, m, pos, ph, 3, exact=exact, staged=True)
    for u, w in zip(a[:3], b[:3]):
        assert np.array_equal(u, w)


def test_mixed_dense_sand_many_contacts_bits(gpu_mod, oracle):
    # a sand clump at twice the lattice density (about 30 contacts within d
    # per particle, more than the projection loop's buffer of 16) next to a
    # fluid block, over the solid field: exact mode bit for bit
    m = scenes.mixed_small(8, 6, 8, box=(24, 24, 24), field=(24, 24), seed=11)
    pos, ph = m.arrays()
    clump = scenes.jittered_lattice(6, 6, 6, cell=0.5, origin=(4.0, 8.0, 4.0), amplitude=0.01, seed=5)
    pos = np.concatenate([clump, pos]).astype(np.float32)
    ph = np.concatenate([np.ones(clump.shape[0], np.int32), ph])
    ox, ov, oid, _ = _oracle_run(oracle, m, pos, ph, 2)
    gx, gv, gid, _ = _gpu_run(gpu_mod, m, pos, ph, 2)
    assert np.array_equal(gid, oid)
    assert np.array_equal(gx, ox) and np.array_equal(gv, ov)
