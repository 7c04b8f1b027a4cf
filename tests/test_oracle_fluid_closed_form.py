"""Pins the oracle's restatement of the reference's simulate_fluid
(oracle/lus_oracle.c orc_simulate_fluid, src/Simulate.cpp:27-115).

The expected state is computed here in float64, straight from the reference's
loop structure: dt clamped to [0.001, 0.01] (:31); v += g * mass * dt and
x* = x + v dt (:48-51); neighbour lists in the order find_neighbors_uniform_grid
builds them (src/neighbors/Neighbors.cpp:366-450: cells filled with sand
indices then solid indices, ascending; a particle's list walks the 27 cells
in (dy, dx, dz) order); the density with self counted twice (:60-66); one
lambda per particle from the cubic kernel of src/Kernels.cpp:5-41 (:69-86);
then ONE sequential sweep in index order (:90-113) in which particle i's
correction uses lambdas[t] for the t-th entry of its list (the list position,
not the neighbour's index) and the positions x* of particles before it as
already corrected, followed by resolve_collision (:11-24: a coordinate at or
below r becomes 0.01, one above D - r becomes D - r - 0.01) and v = (x* - x)/dt.

The mutation tests compile oracle variants with one plausible mistake each and
check that the closed form rejects every one: self counted once, lambdas[j]
indexed by the neighbour, a Jacobi sweep (positions from before the sweep),
gravity not multiplied by the mass, no dt clamp, resolve_collision snapping
to r instead of 0.01.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


def cubic(r, h, f):
    """src/Kernels.cpp:5-19, float64."""
    k = 8.0 / (math.pi * h ** 3)
    q = r * f / h
    if q > 1.0:
        return 0.0
    if q <= 0.5:
        return k * (6 * q ** 3 - 6 * q ** 2 + 1)
    return k * 2 * (1 - q) ** 3


def cubic_grad(d, h, f):
    """src/Kernels.cpp:26-41, float64."""
    l = 48.0 / (math.pi * h ** 3)
    rl = float(np.linalg.norm(d)) * f
    q = rl / h
    if not (rl > 1e-5 and q <= 1.0):
        return np.zeros(3)
    gq = d / (rl * h)
    if q <= 0.5:
        return l * q * (3 * q - 2) * gq
    return l * (-(1 - q) ** 2) * gq


def resolve(value, lo, hi):
    if value <= lo:
        return 0.01
    if value > hi:
        return hi - 0.01
    return value


def reference_lists(ps, sand, solid, grid, cs, h):
    gx, gy, gz = grid
    cell = lambda p: int(p[1] / cs) * gx * gz + int(p[0] / cs) * gz + int(p[2] / cs)  # noqa: E731
    cells = {}
    for i in list(range(*sand)) + list(range(*solid)):
        cells.setdefault(cell(ps[i]), []).append(i)
    lists = {}
    for i in range(*sand):
        c = cell(ps[i])
        yy, rem = divmod(c, gx * gz)
        xx, zz = divmod(rem, gz)
        out = []
        for y in (-1, 0, 1):
            for x in (-1, 0, 1):
                for z in (-1, 0, 1):
                    if not (0 <= xx + x < gx and 0 <= yy + y < gy and 0 <= zz + z < gz):
                        continue
                    for j in cells.get((yy + y) * gx * gz + (xx + x) * gz + (zz + z), []):
                        d = ps[i] - ps[j]
                        if d @ d <= h * h:
                            out.append(j)
        lists[i] = out
    return lists


def closed_form(s, p, v, sand, solid, dt):
    h, f, m, rho0 = float(s.kernelRadius), float(s.kernelFactor), float(s.mass), float(s.rest_density)
    p = p.astype(np.float64)
    v = v.astype(np.float64)
    dt = min(max(dt, 0.001), 0.01)
    g = np.array(s.gravity[:], np.float64)
    for i in range(*sand):
        v[i] = v[i] + g * m * dt
    ps = p.copy()
    for i in range(*sand):
        ps[i] = p[i] + v[i] * dt
    lists = reference_lists(ps, sand, solid, (s.gridX, s.gridY, s.gridZ), float(s.cell_size), h)
    lam = np.zeros(p.shape[0])
    for i in range(*sand):
        rho = sum(m * cubic(np.linalg.norm(ps[i] - ps[j]), h, f) for j in lists[i]) + m * cubic(0.0, h, f)
        C = rho / rho0 - 1.0
        gs, gcp = 0.0, np.zeros(3)
        for j in lists[i]:
            ng = -(m / rho0) * cubic_grad(ps[i] - ps[j], h, f)
            gs += ng @ ng
            gcp -= ng
        gs += gcp @ gcp
        lam[i] = -C / (gs + float(s.relaxation_epsilon)) if gs > 0 else 0.0
    wdq = cubic(float(s.s_corr_dq), h, f)
    X, Y, Z = int(s.domainX), int(s.domainY), int(s.domainZ)
    r = float(s.particleRadius)
    for i in range(*sand):
        pf = np.zeros(3)
        for t, j in enumerate(lists[i]):
            ij = ps[i] - ps[j]
            sc = -float(s.s_corr_k) * (cubic(np.linalg.norm(ij), h, f) / wdq) ** float(s.s_corr_n)
            pf += (lam[i] + lam[t] + sc) * cubic_grad(ij, h, f)
        ps[i] = ps[i] + pf / rho0
        ps[i] = [resolve(ps[i][0], r, X - r), resolve(ps[i][1], r, Y - r), resolve(ps[i][2], r, Z - r)]
        v[i] = (ps[i] - p[i]) / dt
        p[i] = ps[i]
    return p, v, lam


def scene():
    # a 3x3x2 block at spacing 0.8 over three solid particles, one particle
    # driven through the floor (resolve_collision snaps it to 0.01), the rest
    # with seeded velocities; 18 sand slots, 3 solid slots at the tail
    g = np.arange(3, dtype=np.float32) * np.float32(0.8)
    blk = np.stack(np.meshgrid(g, g, np.arange(2, dtype=np.float32) * np.float32(0.8), indexing="ij"),
                   -1).reshape(-1, 3) + np.array([3.1, 0.9, 3.3], np.float32)
    rng = np.random.default_rng(2)
    blk = (blk + rng.uniform(-0.05, 0.05, blk.shape)).astype(np.float32)
    v = rng.uniform(-1.0, 1.0, blk.shape).astype(np.float32)
    v[0] = [0.0, -60.0, 0.0]
    sol = np.array([[3.5, 0.5, 3.5], [4.5, 0.5, 3.5], [3.5, 0.5, 4.5]], np.float32)
    n, ns = blk.shape[0], sol.shape[0]
    total = n + ns
    p = np.zeros((total, 3), np.float32)
    p[:n], p[n:] = blk, sol
    vv = np.zeros((total, 3), np.float32)
    vv[:n] = v
    return p, vv, (0, n), (n, total)


def run(oracle, L=None, dt=1.0 / 60.0):
    s = oracle.init_constants(12, 12, 12, 0.5, 1.0, 4.0)
    p, v, sand, solid = scene()
    po, pso, vo = p.copy(), p.copy(), v.copy()
    lam = np.zeros(p.shape[0], np.float32)
    lib = oracle.lib() if L is None else L
    import ctypes as C
    lib.orc_simulate_fluid(C.byref(s), oracle._fp(po), oracle._fp(pso), oracle._fp(vo), oracle._fp(lam),
                           sand[0], sand[1], solid[0], solid[1], dt)
    cp, cv, clam = closed_form(s, p, v, sand, solid, dt)
    return (po, vo, lam), (cp, cv, clam), sand


def deviation(oracle, L=None):
    (po, vo, lam), (cp, cv, clam), sand = run(oracle, L)
    n = sand[1]
    ex = float(np.abs(po[:n] - cp[:n]).max()) / (8 * float(np.spacing(np.float32(12.0))))
    ev = float(np.abs(vo[:n] - cv[:n]).max()) / (8 * float(np.spacing(np.float32(12.0))) / 0.01 +
                                                  1e-5 * float(np.abs(cv[:n]).max()))
    el = float(np.abs(lam[:n] - clam[:n]).max()) / (1e-5 * float(np.abs(clam[:n]).max()))
    return max(ex, ev, el), (po, vo, lam), (cp, cv, clam)


def test_simulate_fluid_matches_closed_form(oracle):
    worst, (po, vo, lam), (cp, cv, clam) = deviation(oracle)
    assert worst <= 1.0, worst
    # the scene exercises what the closed form is for: the floor snap to 0.01,
    # lambdas that differ (so list position vs neighbour index matters) and
    # corrections far above the tolerance
    assert po[0, 1] == np.float32(0.01)
    assert np.ptp(clam[:18]) > 0.1 * np.abs(clam[:18]).max()


MUTANTS = {
    "self counted once": ("density = density + s->mass * W0;\n", "\n"),
    "lambdas indexed by the neighbour": ("(lambdas[i] + lambdas[t - a])", "(lambdas[i] + lambdas[j])"),
    "gravity not multiplied by the mass": ("float g = s->gravity[c] * s->mass;", "float g = s->gravity[c];"),
    "no dt clamp": ("dt = gclamp(dt, 0.001f, 0.01f); /* :31 */", "/* no clamp */"),
    "resolve_collision snaps to r": ("if (value <= mn) return eps;", "if (value <= mn) return mn;"),
    "Jacobi sweep": ("float* q = ps + 3 * (size_t)i;\n        for (int c = 0; c < 3; c++) q[c] = q[c] + pf[c];",
                     "float* q = p + 3 * (size_t)i; float qq[3];\n"
                     "        for (int c = 0; c < 3; c++) qq[c] = ps[3 * (size_t)i + c] + pf[c];\n"
                     "        for (int c = 0; c < 3; c++) q[c] = qq[c];"),
}


def _build_mutant(tmp_path, old, new):
    d = tmp_path / "src"
    d.mkdir()
    for f in ("lus_oracle.c", "lus_oracle.h", "pbf_core.inc", "mixed_core.inc"):
        shutil.copy(os.path.join(ORACLE_DIR, f), d / f)
    text = (d / "lus_oracle.c").read_text()
    assert text.count(old) == 1, f"mutation site not unique: {old!r}"
    (d / "lus_oracle.c").write_text(text.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-fno-fast-math", "-ffp-contract=off", "-shared",
                    str(d / "lus_oracle.c"), "-o", str(so), "-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_closed_form_rejects_mutant(oracle, tmp_path, mutant):
    L = oracle.load(_build_mutant(tmp_path, *MUTANTS[mutant]))
    worst, _, _ = deviation(oracle, L)
    assert worst > 10.0, f"mutant '{mutant}' passes the closed form ({worst:.2f} x tolerance)"
