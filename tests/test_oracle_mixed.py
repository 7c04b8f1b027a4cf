"""Pins the oracle's mixed sand + fluid step (oracle/mixed_core.inc): it
reduces to the PBF step when every particle is fluid, the contact projection
of simulate_sand (src/Simulate.cpp:240-300) has its closed form on a head-on
pair, solids never move, sand in contact with fluid gets the contact term only,
friction slows sliding on solids, and FP32 agrees with FP64."""
import numpy as np
import pytest

from lustrine_amd import scenes


def _params(oracle, cfg, iterations=None, gravity=None):
    p = oracle.pbf_params(cfg, iterations)
    if gravity is not None:
        for a in range(3):
            p.gravity[a] = gravity[a]
    return p


def test_all_fluid_equals_pbf_step(oracle):
    cfg = scenes.small(8, 8, 8, box=(16, 16, 16), seed=3)
    x0 = cfg.positions()
    p = oracle.pbf_params(cfg)
    xa, va = x0.copy(), np.zeros_like(x0)
    ida = np.arange(cfg.n, dtype=np.int32)
    xb, vb, idb = x0.copy(), np.zeros_like(x0), ida.copy()
    ph = np.zeros(cfg.n, np.int32)
    for _ in range(3):
        _, e1, m1 = oracle.pbf_step(p, xa, va, ida, np.float32(cfg.dt), final_err=True)
        e2, m2 = oracle.mixed_step(p, oracle.mix_params(), xb, vb, idb, ph, np.float32(cfg.dt))
        assert e1 == e2 and m1 == pytest.approx(m2, rel=1e-6)
    assert np.array_equal(ida, idb) and np.array_equal(xa, xb) and np.array_equal(va, vb)


def _pair(oracle, phase_b, gap, iterations, mp=None):
    cfg = scenes.small(2, 1, 1, box=(16, 16, 16))
    x = np.array([[8.0, 8.0, 8.0], [8.0 + gap, 8.0, 8.0]], np.float32)
    v = np.zeros_like(x)
    ids = np.arange(2, dtype=np.int32)
    ph = np.array([1, phase_b], np.int32)
    p = _params(oracle, cfg, iterations, gravity=(0.0, 0.0, 0.0))
    oracle.mixed_step(p, mp or oracle.mix_params(), x, v, ids, ph, np.float32(cfg.dt))
    out = np.zeros_like(x)
    out[ids] = x
    return out


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_sand_pair_contact_closed_form(oracle, iterations):
    # each Jacobi projection moves both sand particles by 0.8 * 1/2 of the
    # overlap: the overlap shrinks by (1 - 0.8) per iteration
    out = _pair(oracle, 1, 0.6, iterations)
    gap = out[1, 0] - out[0, 0]
    assert gap == pytest.approx(1.0 - 0.4 * 0.2 ** iterations, abs=2e-6)
    assert (out[0, 0] + out[1, 0]) / 2 == pytest.approx(8.3, abs=2e-6)   # centre fixed
    assert np.allclose(out[:, 1:], 8.0)


def test_sand_against_solid_full_weight(oracle):
    # a solid neighbour takes no share: the sand particle moves 0.8 of the overlap
    out = _pair(oracle, 2, 0.6, 1)
    assert out[1, 0] == pytest.approx(8.6, abs=1e-6)            # the solid stays
    assert out[0, 0] == pytest.approx(8.0 - 0.8 * 0.4, abs=2e-6)


def test_sand_against_fluid_contact_only(oracle):
    # mixed pair: the sand particle moves 0.8 * 1/2 of the overlap (no friction term)
    out = _pair(oracle, 0, 0.6, 1)
    assert out[0, 0] == pytest.approx(8.0 - 0.8 * 0.5 * 0.4, abs=2e-6)
    assert np.allclose(out[0, 1:], 8.0)


def test_friction_slows_sliding_on_solids(oracle):
    # a sand particle resting on a solid floor particle, sliding along x
    cfg = scenes.small(2, 1, 1, box=(16, 16, 16))
    res = []
    for mu in (0.0, 5.0):
        x = np.array([[8.0, 8.9, 8.0], [8.0, 8.0, 8.0]], np.float32)
        v = np.array([[3.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
        ids = np.arange(2, dtype=np.int32)
        ph = np.array([1, 2], np.int32)
        p = _params(oracle, cfg, 3, gravity=(0.0, -9.81, 0.0))
        oracle.mixed_step(p, oracle.mix_params(mu_s=mu, mu_k=mu), x, v, ids, ph, np.float32(cfg.dt))
        out = np.zeros_like(x)
        out[ids] = x
        assert np.array_equal(out[1], [8.0, 8.0, 8.0])
        res.append(out[0, 0] - 8.0)
    # frictionless: at least the free slide (the contact normal tilts along +x
    # as the particle moves and adds a little)
    assert res[0] >= 3.0 / 60.0 - 1e-6
    assert 0.0 <= res[1] < 0.5 * 3.0 / 60.0


def test_solids_never_move_and_fp64_agrees(oracle):
    m = scenes.mixed_small(seed=4)
    pos, ph = m.arrays()
    p = oracle.pbf_params(m.fluid)
    mp = oracle.mix_params(mu_s=m.mu_s, mu_k=m.mu_k)
    x32, v32 = pos.copy(), np.zeros_like(pos)
    x64, v64 = pos.astype(np.float64), np.zeros(pos.shape, np.float64)
    i32 = np.arange(pos.shape[0], dtype=np.int32)
    i64 = i32.copy()
    p32, p64 = ph.copy(), ph.copy()
    for _ in range(4):
        oracle.mixed_step(p, mp, x32, v32, i32, p32, np.float32(m.fluid.dt))
        oracle.mixed_step(p, mp, x64, v64, i64, p64, m.fluid.dt, fp64=True)
    a32 = np.zeros_like(x32)
    a32[i32] = x32
    a64 = np.zeros_like(x64)
    a64[i64] = x64
    solid = ph == 2
    assert np.array_equal(a32[solid], pos[solid])
    vs = np.zeros_like(v32)
    vs[i32] = v32
    assert not vs[solid].any()
    assert np.array_equal(p32, ph[i32])                          # phases ride along
    # FP32 vs FP64 over four steps: within 1e-5 of the largest coordinate
    assert np.abs(a32 - a64).max() < 1e-5 * np.abs(a64).max()
    assert (a32[~solid] >= 0.5).all() and (a32[~solid] <= 23.5).all()
