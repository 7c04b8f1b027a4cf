"""CPU oracle of the learned warm start: features and the 32-64-64-3 network.

TEST INFRASTRUCTURE ONLY.  Nothing in the library imports this module; the
tests use it to check the GPU kernels.  The learned warm start has no
counterpart in the reference (its fluid solve is one uncorrected pass,
src/Simulate.cpp:83-154), so this module restates the definitions the library
documents in include/lustrine_gpu.h and is pinned by closed forms and finite
differences in tests/test_oracle_mlp.py.

Network (float64 here): z1 = x W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2,
h2 = silu(z2), o = h2 W3 + b3; loss = mean_r sum_c (o_rc - y_rc * inv_scale)^2.
With bf16=True the operands are rounded to bfloat16 (round to nearest even)
where the GPU kernels round them: x, W1, W2, W3, h1, h2 in the forward pass;
dL/do, dL/dz2, dL/dz1 in the backward pass; the bias gradients are the column
sums of those rounded values.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 32, 64, 3
W1, B1 = 0, 32 * 64
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 64 * 3
N_PARAMS = B3 + 3
NET_MU, NET_INV_SIGMA, NET_PARAMS = 0, 32, 64
NET_OUT_SCALE = NET_PARAMS + N_PARAMS
NET_FLOATS = NET_OUT_SCALE + 1


def bf16(a) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    x = np.asarray(a, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    rounded = (b + np.uint64(0x7FFF) + ((b >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    out = rounded.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(x), x, out)
    return out.astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(N_IN, N_HID), p[B1:W2], p[W2:B2].reshape(N_HID, N_HID), p[B2:W3],
            p[W3:B3].reshape(N_HID, N_OUT), p[B3:N_PARAMS])


def pack(w1, b1, w2, b2, w3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(w1), b1, np.ravel(w2), b2, np.ravel(w3), b3]).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf16_round=True):
    """Network output o [rows][3] (without out_scale) and the intermediates."""
    r = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = r(X)
    z1 = x @ r(w1) + b1
    h1 = r(silu(z1))
    z2 = h1 @ r(w2) + b2
    h2 = r(silu(z2))
    o = h2 @ r(w3) + b3
    return o, (x, z1, h1, z2, h2)


def loss_grad(p, X, Y, inv_scale=1.0, bf16_round=True):
    """Loss and its gradient with respect to the parameter vector."""
    r = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    o, (x, z1, h1, z2, h2) = forward(p, X, bf16_round)
    rows = o.shape[0]
    diff = o - np.asarray(Y, np.float64) * inv_scale
    loss = float(np.sum(diff * diff) / rows)
    d3 = r(2.0 * diff / rows)
    gw3 = h2.T @ d3
    gb3 = d3.sum(axis=0)
    dh2 = d3 @ r(w3).T
    dz2 = r(dh2 * silu_grad(z2))
    gw2 = h1.T @ dz2
    gb2 = dz2.sum(axis=0)
    dh1 = dz2 @ r(w2).T
    dz1 = r(dh1 * silu_grad(z1))
    gw1 = x.T @ dz1
    gb1 = dz1.sum(axis=0)
    return loss, pack(gw1, gb1, gw2, gb2, gw3, gb3)


def infer(net, feat, bf16_round=True):
    """out_scale * network((feat - mu) * inv_sigma), as the inference kernel."""
    net = np.asarray(net, np.float64)
    mu = net[NET_MU:NET_MU + 32].astype(np.float32)
    inv = net[NET_INV_SIGMA:NET_INV_SIGMA + 32].astype(np.float32)
    xn = (np.asarray(feat, np.float32) - mu) * inv  # float32, as the kernel
    o, _ = forward(net[NET_PARAMS:NET_PARAMS + N_PARAMS], xn, bf16_round)
    return o * net[NET_OUT_SCALE]


def random_params(rng, scale=1.0) -> np.ndarray:
    """LeCun-uniform weights and small random biases (test inputs)."""
    def u(fan_in, shape):
        a = scale * np.sqrt(3.0 / fan_in)
        return rng.uniform(-a, a, shape)
    return pack(u(N_IN, (N_IN, N_HID)), rng.uniform(-0.1, 0.1, N_HID), u(N_HID, (N_HID, N_HID)),
                rng.uniform(-0.1, 0.1, N_HID), u(N_HID, (N_HID, N_OUT)), rng.uniform(-0.1, 0.1, N_OUT))


# ------------------------------------------------------------------ features
def pbf_feature_constants(h, f, mass, rest_density):
    """(c_rho, hf2, f2) in float32 as the library derives them: c_rho =
    m * P6 / rho0 with P6 = 315 / (64 * 3.14 * hf^9) rounded to float32."""
    hfd = float(np.float32(h)) * float(np.float32(f))
    p6 = np.float32(315.0 / (64.0 * 3.14 * hfd ** 9))
    c_rho = np.float32(float(mass) * float(p6) / float(rest_density))
    hf = np.float32(h) * np.float32(f)
    return c_rho, np.float32(hf * hf), np.float32(np.float32(f) * np.float32(f))


def features(xp, xold, h, f, mass, rest_density, lo, hi):
    """Raw features of every particle (float64 [n][32]) over N(i) = {j :
    |x*_i - x*_j|^2 <= h^2} by brute force.  Distances, bins and the density
    terms use float32 operations in the library's order (so the bin of every
    pair is the library's); the sums are float64."""
    xp = np.asarray(xp, np.float32)
    xold = np.asarray(xold, np.float32)
    n = xp.shape[0]
    h32 = np.float32(h)
    h2 = h32 * h32
    inv_h = np.float32(1.0) / np.float32(np.sqrt(h2))
    c_rho, hf2, f2 = pbf_feature_constants(h, f, mass, rest_density)
    out = np.zeros((n, 32), np.float64)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    for i in range(n):
        d = xp - xp[i]                                    # offsets x*_j - x*_i (float32)
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        nb = r2 <= h2
        t = hf2 - f2 * r2[nb]
        rho = np.sum(((t * t) * t).astype(np.float64))
        others = nb.copy()
        others[i] = False
        q = np.sqrt(r2[others]) * inv_h                   # float32
        b = np.minimum((q * np.float32(8.0)).astype(np.int32), 7)
        dd = d[others].astype(np.float64)
        out[i, 0] = rho * float(c_rho) - 1.0
        out[i, 1:4] = (xp[i] - xold[i]) * inv_h
        for k in range(8):
            sel = b == k
            if sel.any():
                out[i, 4 + 3 * k:7 + 3 * k] = dd[sel].sum(axis=0) / sel.sum() * float(inv_h)
        cnt = int(others.sum())
        out[i, 28] = cnt / 32.0
        if cnt:
            qd = q.astype(np.float64)
            out[i, 29] = qd.mean()
            out[i, 30] = (qd * qd).mean()
        wall = min(float(np.min(xp[i] - lo)), float(np.min(hi - xp[i])))
        out[i, 31] = min(1.0, wall * float(inv_h))
    return out
