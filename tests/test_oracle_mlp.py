"""Pins the CPU oracle of the learned warm start (oracle/mlp.py): bf16
rounding known answers, closed forms of the network and its gradient,
finite differences of the float64 backward pass, and hand-derived features.
The learned warm start has no reference counterpart (the reference's fluid
solve is one uncorrected pass, src/Simulate.cpp:83-154)."""
import numpy as np
import pytest

from oracle import mlp as M


def test_bf16_rounding_known_answers():
    e = 2.0 ** -8
    x = np.array([1.0, 1 + e, 1 + 3 * e, 1 + e + 2.0 ** -20, -2.5, 0.0], np.float32)
    got = M.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie -> even (mantissa 0)
    assert got[2] == 1 + 4 * e                # tie between odd 1+2e and even 1+4e
    assert got[3] == 1 + 2 * e                # above the tie -> up
    assert got[4] == -2.5 and got[5] == 0.0
    # every result is a bf16 value: low 16 bits of the float32 pattern are zero
    bits = got.astype(np.float32).view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)
    assert np.isnan(M.bf16(np.float32(np.nan)))
    # 8 significant bits: relative error at most half a unit in the 8th bit, 2^-8
    r = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    rel = np.abs(M.bf16(r) - r) / np.abs(r)
    assert rel.max() <= 2.0 ** -8
    assert rel.max() > 2.0 ** -9  # and the bound is attained in order of magnitude


def test_zero_weights_closed_form():
    # W = 0: o = b3 for every row; dL/db3 = 2 mean(b3 - y s); dL/dW3[k][c] =
    # silu(b2_k) dL/db3_c; every other gradient is zero
    rng = np.random.default_rng(1)
    b1, b2, b3 = rng.normal(size=64), rng.normal(size=64), rng.normal(size=3)
    p = M.pack(np.zeros((32, 64)), b1, np.zeros((64, 64)), b2, np.zeros((64, 3)), b3)
    X = rng.normal(size=(50, 32))
    Y = rng.normal(size=(50, 3))
    s = 0.5
    loss, g = M.loss_grad(p, X, Y, s, bf16_round=False)
    want_loss = np.mean(np.sum((b3 - Y * s) ** 2, axis=1))
    assert loss == pytest.approx(want_loss, rel=1e-12)
    gb3 = 2 * np.mean(b3 - Y * s, axis=0)
    gw1, gbb1, gw2, gbb2, gw3, gbb3 = M.unpack(g)
    assert np.allclose(gbb3, gb3, rtol=1e-12)
    assert np.allclose(gw3, np.outer(M.silu(b2), gb3), rtol=1e-12)
    assert not gw1.any() and not gbb1.any() and not gw2.any() and not gbb2.any()


def test_silu_and_derivative():
    z = np.linspace(-8, 8, 101)
    assert np.allclose(M.silu(z), z * (1 / (1 + np.exp(-z))))
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.allclose(M.silu_grad(z), fd, atol=1e-8)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = M.random_params(rng)
    X = rng.normal(size=(40, 32))
    Y = rng.normal(size=(40, 3))
    loss, g = M.loss_grad(p, X, Y, 0.7, bf16_round=False)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 60, replace=False),
                          [M.B1, M.B2 + 5, M.B3, M.B3 + 2, M.W3, M.W1]])
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (M.loss_grad(pp, X, Y, 0.7, False)[0] - M.loss_grad(pm, X, Y, 0.7, False)[0]) / (2 * eps)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_bf16_gradient_close_to_fp64():
    # the bf16-rounded pass stays within a few 2^-8 of the exact one
    rng = np.random.default_rng(3)
    p = M.random_params(rng)
    X = rng.normal(size=(256, 32))
    Y = rng.normal(size=(256, 3))
    l64, g64 = M.loss_grad(p, X, Y, 1.0, bf16_round=False)
    l16, g16 = M.loss_grad(p, X, Y, 1.0, bf16_round=True)
    assert abs(l16 - l64) / l64 < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_infer_applies_normalisation_and_scale():
    rng = np.random.default_rng(4)
    p = M.random_params(rng)
    mu = rng.normal(size=32)
    inv = rng.uniform(0.5, 2.0, 32)
    net = np.concatenate([mu, inv, p, [0.25]])
    assert net.shape[0] == M.NET_FLOATS
    f = rng.normal(size=(10, 32))
    out = M.infer(net, f, bf16_round=False)
    xn = (f.astype(np.float32) - mu.astype(np.float32)) * inv.astype(np.float32)
    want = M.forward(p, xn, False)[0] * 0.25
    assert np.allclose(out, want)


def _two_particles(r_over_h, h=2.0):
    xp = np.array([[10.0, 10.0, 10.0], [10.0 + r_over_h * h, 10.0, 10.0]], np.float32)
    return xp


def test_features_two_particles_closed_form():
    h, f, m, rho0 = 2.0, 0.5, 5.0, 40.0
    xp = _two_particles(0.6, h)
    xold = xp - np.float32(0.25)
    lo, hi = (9.5, 0.5, 0.5), (30.0, 30.0, 30.0)   # nearest wall of particle 0: x = 9.5 -> 0.5 away
    F = M.features(xp, xold, h, f, m, rho0, lo, hi)
    hf = h * f
    p6 = np.float32(315.0 / (64.0 * 3.14 * hf ** 9))
    r = 0.6 * h
    rho = (hf * hf) ** 3 + (hf * hf - f * f * r * r) ** 3
    assert F[0, 0] == pytest.approx(m * float(p6) / rho0 * rho - 1.0, rel=1e-5)
    assert np.allclose(F[0, 1:4], 0.125)             # 0.25 / h
    b = int(8 * 0.6)                                  # bin 4
    assert np.allclose(F[0, 4 + 3 * b:7 + 3 * b], [0.6, 0.0, 0.0], atol=1e-6)
    assert np.allclose(F[1, 4 + 3 * b:7 + 3 * b], [-0.6, 0.0, 0.0], atol=1e-6)
    others = [k for k in range(8) if k != b]
    for k in others:
        assert not F[0, 4 + 3 * k:7 + 3 * k].any()
    assert F[0, 28] == pytest.approx(1 / 32)
    assert F[0, 29] == pytest.approx(0.6, rel=1e-6)
    assert F[0, 30] == pytest.approx(0.36, rel=1e-6)
    assert F[0, 31] == pytest.approx(0.25)           # 0.5 from the x = 9.5 wall, / h
    assert F[1, 31] == pytest.approx(0.85)           # 1.7 from the same wall
    F2 = M.features(xp + np.float32(8.0), xold, h, f, m, rho0, lo, hi)
    assert F2[0, 31] == 1.0                          # clipped at one kernel radius


def test_features_isolated_and_boundary_bins():
    h = 2.0
    # particle at exactly r = h belongs to bin 7; beyond h it is no neighbour
    xp = np.array([[5.0, 5.0, 5.0], [5.0, 5.0 + h, 5.0], [5.0, 5.0, 5.0 - h - 0.01]], np.float32)
    F = M.features(xp, xp, h, 0.5, 5.0, 40.0, (0.5,) * 3, (20.0,) * 3)
    assert F[0, 28] == pytest.approx(1 / 32)
    assert np.allclose(F[0, 4 + 21:7 + 21], [0.0, 1.0, 0.0])
    assert F[2, 28] == 0.0 and F[2, 29] == 0.0 and F[2, 30] == 0.0
    assert not F[2, 4:28].any()
