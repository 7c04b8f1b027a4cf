"""Pins the magnitudes of the oracle's PBF substep (oracle/pbf_core.inc).

The expected values are written here, in float64, from the equations of
Macklin & Mueller, "Position Based Fluids" (2013), in the form the reference's
simulate_fluid gives them (src/Simulate.cpp:60-100):

    rho_i   = m * sum_{j in N(i)} W(|x_i - x_j|)                 (self once)
    C_i     = rho_i / rho0 - 1                                   (eq. 1)
    g_ij    = -(m / rho0) * gradW(x_i - x_j),  j != i            (eq. 8)
    lambda  = -C_i / (sum_j |g_ij|^2 + |sum_j g_ij|^2 + eps)     (eq. 9, 11)
    s_corr  = -k * (W(r_ij) / W(dq))^n                           (eq. 13)
    dp_i    = (1 / rho0) sum_{j != i} (l_i + l_j + s_corr) gradW (eq. 14)
    v_i    += c * sum_{j in N(i)} (v_j - v_i) W(|x_i - x_j|)     (eq. 17, XSPH)

with W = poly6 and gradW = -spiky in the shapes of src/Kernels.cpp:43-67
(kernelFactor f, pi written 3.14).  gradW is taken as the central finite
difference of the potential S(r) = 5 / (3.14 hf^6 f^2) (hf - f r)^3 and checked
against the oracle's orc_spiky_kernel, which test_oracle.py pins to the same
potential; W is checked against orc_poly6_kernel, pinned there by its integral.
None of the oracle's derived constants (P6, SPf, mr0, inv_tq3, cW) is used.

The mutation tests then compile variants of the oracle with one plausible
mistake each (a dropped 1/f in the spiky factor, (m/rho0) not squared, a wrong
epsilon, an s_corr exponent off by one, a doubled XSPH factor, lambda_i in
place of lambda_j, the s_corr sign, self left out of the density) and check
that the closed forms reject every one of them.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lustrine_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


# ------------------------------------------------------------------ closed form
def poly6(r, h, f):
    """src/Kernels.cpp:43-51 in float64."""
    hf = h * f
    r = np.asarray(r, np.float64)
    w = 315.0 / (64.0 * 3.14 * hf ** 9) * (hf ** 2 - (f * r) ** 2) ** 3
    return np.where(r <= h, w, 0.0)


def spiky_potential(r, h, f):
    hf = h * f
    return 5.0 / (3.14 * hf ** 6 * f * f) * (hf - f * r) ** 3


def grad_w(d, h, f):
    """-spiky(d) as the central difference of the potential along d."""
    r = float(np.linalg.norm(d))
    if r == 0.0 or r > h:
        return np.zeros(3)
    e = 1e-6 * h
    dS = (spiky_potential(r + e, h, f) - spiky_potential(r - e, h, f)) / (2 * e)
    return dS * d / r


def closed_form_step(x, v, dt, cfg, rho0):
    """One PBF substep with K = 1, no gravity, no wall contact, float64."""
    h, f, m = cfg.h, cfg.kernel_factor, cfg.mass
    x = x.astype(np.float64)
    v = v.astype(np.float64)
    xs = x + v * dt
    n = x.shape[0]
    d = xs[:, None, :] - xs[None, :, :]
    r = np.linalg.norm(d, axis=2)
    nbr = r * r <= h * h
    W = poly6(r, h, f) * nbr
    G = np.zeros((n, n, 3))
    for i in range(n):
        for j in range(n):
            if nbr[i, j] and i != j:
                G[i, j] = grad_w(d[i, j], h, f)
    rho = m * W.sum(axis=1)
    Ci = rho / rho0 - 1.0
    g = -(m / rho0) * G
    S = (g * g).sum(axis=(1, 2)) + (g.sum(axis=1) ** 2).sum(axis=1)
    lam = -Ci / (S + cfg.epsilon)
    wdq = poly6(cfg.scorr_dq, h, f)
    dp = np.zeros((n, 3))
    for i in range(n):
        for j in range(n):
            if nbr[i, j] and i != j:
                sc = -cfg.scorr_k * (W[i, j] / wdq) ** cfg.scorr_n
                dp[i] += (lam[i] + lam[j] + sc) * G[i, j]
    dp /= rho0
    x1 = xs + dp
    v1 = (x1 - x) / dt
    d1 = x1[:, None, :] - x1[None, :, :]
    W1 = poly6(np.linalg.norm(d1, axis=2), h, f) * nbr
    v2 = v1 + cfg.c_xsph * ((v1[None, :, :] - v1[:, None, :]) * W1[:, :, None]).sum(axis=1)
    return dict(xs=xs, x1=x1, v1=v1, v2=v2, lam=lam, C=Ci, dp=dp)


# ------------------------------------------------------------------ scenes
def _cfg(box=(12, 12, 12)):
    # h = 2, f = 1/2, m = 5 and the reference's Simulation defaults eps = 10,
    # s_corr k = 1, n = 4, dq = 0.5 (large terms, so each is far above the
    # tolerance); no gravity, one iteration, a large XSPH factor so that its
    # term is well above the rounding of v = (x* - x) / dt
    return scenes.small(2, 2, 2, box=box, gravity=(0.0, 0.0, 0.0), c_xsph=0.1, iterations=1,
                        dt=1.0 / 64.0, epsilon=10.0, scorr_k=1.0)


def _pair():
    x = np.array([[5.0, 6.0, 6.0], [6.2, 6.0, 6.0]], np.float32)
    v = np.array([[0.5, 0.0, -0.25], [-1.0, 0.75, 0.0]], np.float32)
    return x, v


def _triple():
    # collinear along x: 0-1 and 1-2 within h, 0-2 beyond it, so the middle
    # particle has two neighbours and the outer ones one: lambda differs
    x = np.array([[4.0, 6.0, 6.0], [4.9, 6.0, 6.0], [6.3, 6.0, 6.0]], np.float32)
    v = np.array([[0.0, 0.5, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.5]], np.float32)
    return x, v


def _cluster():
    rng = np.random.default_rng(11)
    u = rng.normal(size=(7, 3))
    u = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(0.4, 1.6, (7, 1))
    x = (u + 6.0).astype(np.float32)
    v = rng.uniform(-1.0, 1.0, (7, 3)).astype(np.float32)
    return x, v


def _block():
    # a 3x3x3 block at spacing 0.8 (compressed against the rest spacing 1): the
    # centre is above rest density and the corners below, so lambda changes sign
    g = np.arange(3, dtype=np.float32) * np.float32(0.8)
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.float32(5.3)
    rng = np.random.default_rng(5)
    x = (x + rng.uniform(-0.05, 0.05, x.shape)).astype(np.float32)
    v = rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)
    return x, v


SCENES = {"pair": _pair, "triple": _triple, "cluster": _cluster, "block": _block}


def run_oracle(oracle, x, v, cfg, L=None):
    p = oracle.pbf_params(cfg)
    xo, vo = x.copy(), v.copy()
    ids = np.arange(x.shape[0], dtype=np.int32)
    oracle.pbf_step(p, xo, vo, ids, np.float32(cfg.dt), L=L)
    xb, vb = np.zeros_like(xo), np.zeros_like(vo)
    xb[ids], vb[ids] = xo, vo
    return xb, vb, p.rest_density


def check_against_closed_form(oracle, name, L=None):
    """Largest deviation of the oracle from the closed form, in units of the
    float32 rounding each quantity carries (positions: 4 ulp of the largest
    coordinate; velocities: that over dt plus 1e-5 relative)."""
    x, v = SCENES[name]()
    cfg = _cfg()
    xb, vb, rho0 = run_oracle(oracle, x, v, cfg, L)
    cf = closed_form_step(x, v, cfg.dt, cfg, float(rho0))
    tol_x = 4 * float(np.spacing(np.float32(np.abs(cf["x1"]).max())))
    tol_v = tol_x / cfg.dt + 1e-5 * float(np.abs(cf["v2"]).max())
    ex = float(np.abs(xb - cf["x1"]).max()) / tol_x
    ev = float(np.abs(vb - cf["v2"]).max()) / tol_v
    return ex, ev, cf


@pytest.mark.parametrize("name", sorted(SCENES))
def test_pbf_step_matches_closed_form(oracle, name):
    ex, ev, cf = check_against_closed_form(oracle, name)
    # the scenes are built so that every term is far above the tolerance
    assert np.abs(cf["dp"]).max() > 200 * 4 * float(np.spacing(np.float32(8.0)))
    assert np.abs(cf["v2"] - cf["v1"]).max() > 0.01
    assert ex <= 1.0, f"positions off the closed form by {ex:.2f} x tolerance"
    assert ev <= 1.0, f"velocities off the closed form by {ev:.2f} x tolerance"


def test_closed_form_kernels_are_the_pinned_ones(oracle):
    # W and gradW of the closed form are the oracle's orc_poly6_kernel and
    # -orc_spiky_kernel (both pinned in test_oracle.py)
    s = oracle.init_constants(12, 12, 12, 0.5, 1.0, 4.0)
    h, f = float(s.kernelRadius), float(s.kernelFactor)
    for r in (0.0, 0.3, 0.9, 1.2, 1.7, 1.99):
        assert abs(poly6(r, h, f) - oracle.poly6_kernel(s, r)) <= 2e-7 * poly6(0.0, h, f)
        if r > 0:
            d = np.array([0.48, -0.6, 0.64]) * r
            want = -oracle.spiky_kernel(s, d.astype(np.float32)).astype(np.float64)
            assert np.allclose(grad_w(d, h, f), want, rtol=2e-6, atol=1e-7), r


def test_closed_form_lambda_terms_are_all_material(oracle):
    # the scenes are only useful if each term moves the result well beyond the
    # tolerance: epsilon is not negligible against the gradient sum, s_corr is
    # not negligible against the lambdas, lambda differs between neighbours
    cfg = _cfg()
    x, v = _block()
    _, _, rho0 = run_oracle(oracle, x, v, cfg)
    cf = closed_form_step(x, v, cfg.dt, cfg, float(rho0))
    assert cf["C"].max() > 0.1 and cf["C"].min() < -0.1
    assert cf["lam"].max() > 0 > cf["lam"].min()
    x, v = _triple()
    cf = closed_form_step(x, v, cfg.dt, cfg, float(rho0))
    h, f = cfg.h, cfg.kernel_factor
    g01 = (cfg.mass / rho0) * np.linalg.norm(grad_w(cf["xs"][0] - cf["xs"][1], h, f))
    assert 2 * g01 ** 2 > 1e-3 * cfg.epsilon  # the gradient sum is not lost in epsilon
    r01 = np.linalg.norm(cf["xs"][0] - cf["xs"][1])
    sc = cfg.scorr_k * (poly6(r01, h, f) / poly6(cfg.scorr_dq, h, f)) ** cfg.scorr_n
    assert sc > 0.02 * abs(cf["lam"][0] + cf["lam"][1])


# ------------------------------------------------------------------ mutants
MUTANTS = {
    "spiky factor without 1/f": ("pbf_core.inc", "(3.14 * pow(hfd, 6.0) * (double)p->f)",
                                 "(3.14 * pow(hfd, 6.0))"),
    "(m/rho0) not squared": ("pbf_core.inc", "(k.mr0 * k.mr0) * (a + gg)", "k.mr0 * (a + gg)"),
    "epsilon 1 instead of the parameter": ("pbf_core.inc", "k->eps = (REAL)p->epsilon;",
                                           "k->eps = (REAL)1;"),
    "s_corr exponent n - 1": ("pbf_core.inc", "for (int q = 1; q < k.n_scorr; q++)",
                              "for (int q = 2; q < k.n_scorr; q++)"),
    "XSPH factor doubled": ("pbf_core.inc", "k->cW = (REAL)p->c_xsph * k->P6;",
                            "k->cW = (REAL)2 * (REAL)p->c_xsph * k->P6;"),
    "lambda_i in place of lambda_j": ("pbf_core.inc", "(li + lam[j])", "(li + li)"),
    "s_corr sign": ("pbf_core.inc", "REAL sc = -k.scorr_k * pw;", "REAL sc = k.scorr_k * pw;"),
    "self left out of the density": ("pbf_core.inc", "                    rho = rho + (t * t) * t;\n"
                                     "                    if (r2 > 0) {",
                                     "                    if (r2 > 0) rho = rho + (t * t) * t;\n"
                                     "                    if (r2 > 0) {"),
    "s_corr dq not squared": ("pbf_core.inc", "k->f2 * (dq * dq)", "k->f2 * dq"),
}


def _build_mutant(tmp_path, fname, old, new):
    d = tmp_path / "src"
    d.mkdir()
    for f in ("lus_oracle.c", "lus_oracle.h", "pbf_core.inc", "mixed_core.inc"):
        shutil.copy(os.path.join(ORACLE_DIR, f), d / f)
    text = (d / fname).read_text()
    assert text.count(old) == 1, f"mutation site not unique in {fname}: {old!r}"
    (d / fname).write_text(text.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-fno-fast-math", "-ffp-contract=off", "-shared",
                    str(d / "lus_oracle.c"), "-o", str(so), "-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_closed_form_rejects_mutant(oracle, tmp_path, mutant):
    L = oracle.load(_build_mutant(tmp_path, *MUTANTS[mutant]))
    worst = 0.0
    for name in sorted(SCENES):
        ex, ev, _ = check_against_closed_form(oracle, name, L=L)
        worst = max(worst, ex, ev)
    assert worst > 10.0, f"mutant '{mutant}' passes the closed forms (worst {worst:.2f} x tolerance)"


def test_unmutated_copy_passes(oracle, tmp_path):
    # the mutation harness itself: an unmodified copy builds and passes
    L = oracle.load(_build_mutant(tmp_path, "pbf_core.inc", "typedef struct {", "typedef struct {"))
    for name in sorted(SCENES):
        ex, ev, _ = check_against_closed_form(oracle, name, L=L)
        assert ex <= 1.0 and ev <= 1.0
