"""ctypes binding of the CPU oracle (oracle/liblus_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (lustrine_amd/) never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblus_oracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class OrcSim(C.Structure):
    _fields_ = [
        ("particleRadius", C.c_float), ("particleDiameter", C.c_float),
        ("kernelRadius", C.c_float), ("kernelFactor", C.c_float),
        ("cubic_kernel_k", C.c_float), ("cubic_kernel_l", C.c_float),
        ("domainX", C.c_float), ("domainY", C.c_float), ("domainZ", C.c_float),
        ("rest_density", C.c_float), ("mass", C.c_float), ("relaxation_epsilon", C.c_float),
        ("s_corr_dq", C.c_float), ("s_corr_k", C.c_float), ("s_corr_n", C.c_float),
        ("c_xsph", C.c_float), ("gravity", C.c_float * 3), ("cell_size", C.c_float),
        ("gridX", C.c_int), ("gridY", C.c_int), ("gridZ", C.c_int), ("num_grid_cells", C.c_int),
    ]


class OrcSandState(C.Structure):
    _fields_ = [
        ("total_allocated", C.c_size_t),
        ("ptr_sand_start", C.c_int), ("ptr_sand_end", C.c_int),
        ("ptr_solid_start", C.c_int), ("ptr_solid_end", C.c_int),
        ("num_sand_particles", C.c_int),
        ("positions", C.POINTER(C.c_float)), ("positions_star", C.POINTER(C.c_float)),
        ("positions_tmp", C.POINTER(C.c_float)), ("velocities", C.POINTER(C.c_float)),
        ("attracted", C.POINTER(C.c_int)),
        ("first_iteration", C.c_int), ("computed_static_particles", C.c_int),
        ("attract_flag", C.c_int), ("blow_flag", C.c_int), ("prev_attract_flag", C.c_int),
        ("player_position", C.c_float * 3),
        ("attract_radius", C.c_float), ("blow_radius", C.c_float),
        ("attract_coeff", C.c_float), ("blow_coeff", C.c_float),
        ("static_off", C.POINTER(C.c_long)), ("static_idx", C.POINTER(C.c_int)),
        ("sorted_indices", C.POINTER(C.c_int)),
        ("nbr_off", C.POINTER(C.c_long)), ("nbr_idx", C.POINTER(C.c_int)),
        ("nbr_cap", C.c_long),
        ("cell_sand_off", C.POINTER(C.c_long)),
    ]


class OrcPbfParams(C.Structure):
    _fields_ = [
        ("h", C.c_float), ("f", C.c_float), ("cell_size", C.c_float),
        ("gridX", C.c_int), ("gridY", C.c_int), ("gridZ", C.c_int),
        ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("gravity", C.c_float * 3),
        ("mass", C.c_float), ("rest_density", C.c_float), ("epsilon", C.c_float),
        ("scorr_k", C.c_float), ("scorr_dq", C.c_float), ("scorr_n", C.c_int),
        ("c_xsph", C.c_float), ("iterations", C.c_int),
    ]


class OrcMixParams(C.Structure):
    _fields_ = [("diameter", C.c_float), ("collision", C.c_float), ("friction", C.c_float),
                ("mu_s", C.c_float), ("mu_k", C.c_float)]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_long))


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load(build())
    return _lib


def load(path: str):
    """A configured handle on an oracle library at `path` (the built one, or a
    variant a test compiled from modified sources)."""
    L = C.CDLL(path)
    L.orc_init_constants.argtypes = [C.POINTER(OrcSim), C.c_int, C.c_int, C.c_int,
                                     C.c_float, C.c_float, C.c_float]
    L.orc_cubic_kernel.restype = C.c_float
    L.orc_cubic_kernel.argtypes = [C.POINTER(OrcSim), C.c_float]
    L.orc_cubic_kernel_grad.argtypes = [C.POINTER(OrcSim), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_poly6_kernel.restype = C.c_float
    L.orc_poly6_kernel.argtypes = [C.POINTER(OrcSim), C.c_float]
    L.orc_poly6_kernel_vec.restype = C.c_float
    L.orc_poly6_kernel_vec.argtypes = [C.POINTER(OrcSim), C.POINTER(C.c_float)]
    L.orc_spiky_kernel.argtypes = [C.POINTER(OrcSim), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_cell_ids.argtypes = [C.POINTER(OrcSim), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]
    L.orc_counting_sort.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_size_t,
                                    C.c_size_t, C.POINTER(C.c_int)]
    L.orc_neighbors_fluid.restype = C.c_long
    L.orc_neighbors_fluid.argtypes = [C.POINTER(OrcSim), C.POINTER(C.c_float), C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_int),
                                      C.c_long]
    L.orc_simulate_sand.argtypes = [C.POINTER(OrcSim), C.POINTER(OrcSandState), C.c_float]
    L.orc_simulate_sand_variant.argtypes = [C.POINTER(OrcSim), C.POINTER(OrcSandState), C.c_float, C.c_int]
    L.orc_sand_state_free.argtypes = [C.POINTER(OrcSandState)]
    L.orc_query_cell_num_particles.restype = C.c_int
    L.orc_query_cell_num_particles.argtypes = [C.POINTER(OrcSim), C.POINTER(OrcSandState),
                                               C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]
    L.orc_sink_evict.restype = C.c_int
    L.orc_sink_evict.argtypes = [C.POINTER(OrcSim), C.POINTER(OrcSandState), C.POINTER(C.c_int), C.c_int]
    L.orc_simulate_fluid.argtypes = [C.POINTER(OrcSim)] + [C.POINTER(C.c_float)] * 4 + \
        [C.c_int] * 4 + [C.c_float]
    L.orc_pbf_step_f32.argtypes = [C.POINTER(OrcPbfParams), C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_float,
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_pbf_step_f64.argtypes = [C.POINTER(OrcPbfParams), C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_mixed_step_f32.argtypes = [C.POINTER(OrcPbfParams), C.POINTER(OrcMixParams), C.c_int,
                                     C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.c_float, C.POINTER(C.c_float)]
    L.orc_mixed_step_f64.argtypes = [C.POINTER(OrcPbfParams), C.POINTER(OrcMixParams), C.c_int,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.c_double, C.POINTER(C.c_double)]
    L.orc_pbf_density_error_f64.restype = C.c_double
    L.orc_pbf_density_error_f64.argtypes = [C.POINTER(OrcPbfParams), C.c_int, C.POINTER(C.c_double)]
    return L


# ---------------------------------------------------------------- reference API
def init_constants(X, Y, Z, radius=0.5, diameter=1.0, kernel_radius_scale=4.0) -> OrcSim:
    s = OrcSim()
    lib().orc_init_constants(C.byref(s), X, Y, Z, radius, diameter, kernel_radius_scale)
    return s


def cubic_kernel(s, r):
    return lib().orc_cubic_kernel(C.byref(s), r)


def cubic_kernel_grad(s, r):
    a = np.asarray(r, np.float32).copy()
    out = np.zeros(3, np.float32)
    lib().orc_cubic_kernel_grad(C.byref(s), _fp(a), _fp(out))
    return out


def poly6_kernel(s, r):
    return lib().orc_poly6_kernel(C.byref(s), r)


def poly6_kernel_vec(s, r):
    a = np.asarray(r, np.float32).copy()
    return lib().orc_poly6_kernel_vec(C.byref(s), _fp(a))


def spiky_kernel(s, r):
    a = np.asarray(r, np.float32).copy()
    out = np.zeros(3, np.float32)
    lib().orc_spiky_kernel(C.byref(s), _fp(a), _fp(out))
    return out


def cell_ids(s, pos):
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(pos.shape[0], np.int32)
    lib().orc_cell_ids(C.byref(s), _fp(pos), pos.shape[0], _ip(out))
    return out


def counting_sort(keys, num_cells):
    keys = np.ascontiguousarray(keys, np.int32)
    counts = np.zeros(num_cells + 1, np.int32)
    out = np.zeros(keys.shape[0], np.int32)
    lib().orc_counting_sort(_ip(counts), _ip(keys), keys.shape[0], num_cells, _ip(out))
    return out, counts


def neighbors_fluid(s, positions_star, sand, solid):
    ps = np.ascontiguousarray(positions_star, np.float32)
    n = sand[1] - sand[0]
    off = np.zeros(n + 1, np.int64)
    cap = max(64 * n, 64)
    while True:
        idx = np.zeros(cap, np.int32)
        tot = lib().orc_neighbors_fluid(C.byref(s), _fp(ps), sand[0], sand[1], solid[0], solid[1],
                                        _lp(off), _ip(idx), cap)
        if tot >= 0:
            return off, idx[:tot]
        cap = -tot


class SandStream:
    """One reference simulate_sand stream over a Simulation-shaped state."""

    def __init__(self, s: OrcSim, positions, solids, total_allocated=None):
        n = positions.shape[0]
        ns = 0 if solids is None else solids.shape[0]
        # total_allocated = max(#particles, X*Y*Z) (src/Lustrine.cpp:387-407)
        total = total_allocated or max(n + ns, int(s.domainX) * int(s.domainY) * int(s.domainZ))
        self.s = s
        self.total = total
        self.p = np.zeros((total, 3), np.float32)
        self.ps = np.zeros((total, 3), np.float32)
        self.pt = np.zeros((total, 3), np.float32)
        self.v = np.zeros((total, 3), np.float32)
        self.att = np.zeros(total, np.int32)
        self.p[:n] = positions
        self.ps[:n] = positions
        # solids are written downward from the last slot (src/Lustrine.cpp:491-504)
        for k in range(ns):
            self.p[total - 1 - k] = solids[k]
            self.ps[total - 1 - k] = solids[k]
        st = OrcSandState()
        st.total_allocated = total
        st.ptr_sand_start, st.ptr_sand_end = 0, n
        st.ptr_solid_start, st.ptr_solid_end = (total - ns, total) if ns else (total - 1, total - 1)
        st.num_sand_particles = n
        st.positions, st.positions_star = _fp(self.p), _fp(self.ps)
        st.positions_tmp, st.velocities = _fp(self.pt), _fp(self.v)
        st.attracted = _ip(self.att)
        st.first_iteration = 1
        st.attract_radius, st.blow_radius = 1.5, 2.0
        st.attract_coeff, st.blow_coeff = 1000.0, 500.0
        self.st = st
        self.n = n

    def step(self, dt, variant=0):
        """variant 0 simulate_sand, 1 simulate_sand_credits, 2 simulate_sand_v3."""
        lib().orc_simulate_sand_variant(C.byref(self.s), C.byref(self.st), dt, variant)

    def perm(self):
        return np.ctypeslib.as_array(self.st.sorted_indices, shape=(self.n,)).copy()

    def query_cells(self, min_pos, max_pos, include_solid):
        """query_cell_num_particles (src/Lustrine.cpp:864-914) on the last step's grid."""
        lo = np.ascontiguousarray(min_pos, np.float32)
        hi = np.ascontiguousarray(max_pos, np.float32)
        return int(lib().orc_query_cell_num_particles(C.byref(self.s), C.byref(self.st), _fp(lo), _fp(hi),
                                                      int(bool(include_solid))))

    def sink_evict(self, cells):
        """Sink eviction (src/Lustrine.cpp:806-836) of the listed cells; returns the count."""
        c = np.ascontiguousarray(cells, np.int32)
        k = int(lib().orc_sink_evict(C.byref(self.s), C.byref(self.st), _ip(c), int(c.shape[0])))
        self.n = int(self.st.ptr_sand_end)
        return k

    def cell_sand_off(self):
        """The sand part of every cell's list in the last step's grid (CSR)."""
        return np.ctypeslib.as_array(self.st.cell_sand_off, shape=(self.s.num_grid_cells + 1,)).copy()

    def neighbors(self):
        off = np.ctypeslib.as_array(self.st.nbr_off, shape=(self.n + 1,)).copy()
        idx = np.ctypeslib.as_array(self.st.nbr_idx, shape=(int(off[-1]),)).copy() if off[-1] else \
            np.zeros(0, np.int32)
        return off, idx

    def __del__(self):
        try:
            lib().orc_sand_state_free(C.byref(self.st))
        except Exception:
            pass


def player_box(positions, player, half_extent, max_out=100):
    """set_particles_box_colliders_positions (src/BulletPhysics.cpp:615-646)
    over the sand range: the first max_out slots, in slot order, with
    glm::all(|x - player| <= half_extent) (particle_collide_with_player,
    :602-606; half_extent = (player_box_scale + 4 particle_radius) * 0.5),
    and their positions.  float32 arithmetic as there."""
    p = np.asarray(positions, np.float32)
    d = np.abs(p - np.asarray(player, np.float32))
    inside = np.all(d <= np.asarray(half_extent, np.float32), axis=1)
    slots = np.nonzero(inside)[0][:max_out].astype(np.int32)
    return slots, p[slots]


def chunk_from_grid(cells, X, Y, Z, position, cell_size, subdivision, mask_out=False, num_occupied=None,
                    colors=None):
    """init_chunk_from_grid (src/Lustrine.cpp:676-722): one particle per
    subdivided lattice point of an occupied cell (cells[x*Y*Z + y*Z + z] != 0,
    and != 1 with mask_out), in x-major order over the subdivided lattice, at
    ((x * cs) + position) + cs/2 per axis in float32; the positions array has
    num_occupied * subdivision^3 entries (num_occupied: the grid's
    num_occupied_grid_cells, which the random box overcounts) and entries past
    the generated ones stay 0.  colors: the per-cell colours, copied per
    particle.  Returns (positions, colors or None, generated count)."""
    cells = np.asarray(cells).reshape(X, Y, Z)
    s = int(subdivision)
    occ = (cells != 0) if not mask_out else ((cells != 0) & (cells != 1))
    fine = np.repeat(np.repeat(np.repeat(occ, s, 0), s, 1), s, 2)
    xf, yf, zf = np.nonzero(fine)  # x-major, as the loops
    cs = np.float32(cell_size)
    off = cs * np.float32(0.5)
    pos = np.asarray(position, np.float32)
    p = np.stack([(xf.astype(np.float32) * cs + pos[0]) + off, (yf.astype(np.float32) * cs + pos[1]) + off,
                  (zf.astype(np.float32) * cs + pos[2]) + off], axis=1).astype(np.float32)
    n_occ = int((cells != 0).sum()) if num_occupied is None else int(num_occupied)
    total = n_occ * s ** 3
    out = np.zeros((total, 3), np.float32)
    out[: p.shape[0]] = p[:total]
    col = None
    if colors is not None:
        c = np.asarray(colors, np.float32).reshape(X, Y, Z, 4)
        col = np.zeros((total, 4), np.float32)
        col[: p.shape[0]] = c[xf // s, yf // s, zf // s][:total]
    return out, col, min(p.shape[0], total)


def grid_from_voxels(voxels, sx, sy, sz, palette):
    """init_grid_from_magika_voxel (src/VoxelLoader.cpp:47-100) after the file
    is parsed: voxel (x, y, z) of a model at voxels[x + y*sx + z*sx*sy]
    (colour index, 0 = empty) goes to grid cell x*Y*Z + y*Z + z of a grid with
    X = sx, Y = sy, Z = sz; the cell keeps the colour index and its colour is
    the palette's rgba / 255.  Returns (cells, colors, occupied)."""
    v = np.asarray(voxels, np.uint8).reshape(sz, sy, sx)  # z slowest
    cells = np.transpose(v, (2, 1, 0)).astype(np.int32)   # [x][y][z]
    pal = np.asarray(palette, np.uint8).reshape(256, 4)
    colors = np.where((cells != 0)[..., None], pal[cells].astype(np.float32) / np.float32(255.0), np.float32(0.0))
    return cells.reshape(-1), colors.reshape(-1, 4).astype(np.float32), int((cells != 0).sum())


def simulate_fluid(s, p, ps, v, lambdas, sand, solid, dt):
    lib().orc_simulate_fluid(C.byref(s), _fp(p), _fp(ps), _fp(v), _fp(lambdas),
                             sand[0], sand[1], solid[0], solid[1], dt)


# ---------------------------------------------------------------- PBF substep
def pbf_params(cfg, iterations=None) -> OrcPbfParams:
    p = OrcPbfParams()
    r = np.float32(cfg.radius)
    p.h = cfg.h
    p.f = cfg.kernel_factor
    p.cell_size = cfg.cell_size
    p.gridX, p.gridY, p.gridZ = cfg.grid
    for a in range(3):
        p.lo[a] = r
        p.hi[a] = np.float32(cfg.box[a]) - r
        p.gravity[a] = cfg.gravity[a]
    p.mass = cfg.mass
    p.rest_density = cfg.rest_density()
    p.epsilon = cfg.epsilon
    p.scorr_k = cfg.scorr_k
    p.scorr_dq = cfg.scorr_dq
    p.scorr_n = cfg.scorr_n
    p.c_xsph = cfg.c_xsph
    p.iterations = cfg.iterations if iterations is None else iterations
    return p


def pbf_step(params: OrcPbfParams, x, v, ids, dt, fp64=False, final_err=False, L=None):
    """One substep in place; returns max density constraint of the last lambda
    pass, and with final_err=True also (max(0, max C), mean max(0, C)) at the
    final positions.  L: an oracle handle from load() (default: the built one)."""
    n = x.shape[0]
    L = lib() if L is None else L
    if fp64:
        assert x.dtype == np.float64 and v.dtype == np.float64
        mc, fe = C.c_double(0), (C.c_double * 2)()
        L.orc_pbf_step_f64(C.byref(params), n, _dp(x), _dp(v), _ip(ids), dt, C.byref(mc), fe)
    else:
        assert x.dtype == np.float32 and v.dtype == np.float32
        mc, fe = C.c_float(0), (C.c_float * 2)()
        L.orc_pbf_step_f32(C.byref(params), n, _fp(x), _fp(v), _ip(ids), dt, C.byref(mc), fe)
    return (mc.value, fe[0], fe[1]) if final_err else mc.value


def density_error(params, x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().orc_pbf_density_error_f64(C.byref(params), x.shape[0], _dp(x))


def mix_params(diameter=1.0, collision=0.8, friction=0.7, mu_s=0.6, mu_k=0.4) -> OrcMixParams:
    return OrcMixParams(diameter, collision, friction, mu_s, mu_k)


def mixed_step(params: OrcPbfParams, mp: OrcMixParams, x, v, ids, phase, dt, fp64=False):
    """One mixed sand + fluid substep in place (cell order on return); returns
    (max, mean) over fluid of max(0, rho/rho0 - 1) at the final positions."""
    n = x.shape[0]
    assert phase.dtype == np.int32 and ids.dtype == np.int32
    if fp64:
        assert x.dtype == np.float64 and v.dtype == np.float64
        fe = (C.c_double * 2)()
        lib().orc_mixed_step_f64(C.byref(params), C.byref(mp), n, _dp(x), _dp(v), _ip(ids), _ip(phase), dt, fe)
    else:
        assert x.dtype == np.float32 and v.dtype == np.float32
        fe = (C.c_float * 2)()
        lib().orc_mixed_step_f32(C.byref(params), C.byref(mp), n, _fp(x), _fp(v), _ip(ids), _ip(phase), dt, fe)
    return fe[0], fe[1]
