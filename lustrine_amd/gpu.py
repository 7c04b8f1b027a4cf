"""ctypes binding of liblustrine_gpu.so (include/lustrine_gpu.h).

The library is loaded from this package directory; if it is missing the import
fails with an error naming the build step.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblustrine_gpu.so")

NPHASES = 12
PHASES = ("predict_key", "cell_scan", "scatter", "stable_permute", "lambda", "deltap", "xsph",
          "lambda_build", "features", "mlp", "decide", "tiles")

# learned warm start: network file layout (include/lustrine_gpu.h)
FEATURES = 32
MLP_PARAMS = 6467
NET_MU, NET_INV_SIGMA, NET_PARAMS = 0, 32, 64
NET_OUT_SCALE = NET_PARAMS + MLP_PARAMS
NET_FLOATS = NET_OUT_SCALE + 1
DENSITY_HISTORY = 4096


class LusError(RuntimeError):
    pass


class LusGrid(C.Structure):
    _fields_ = [("cell_size", C.c_float), ("gridX", C.c_int), ("gridY", C.c_int), ("gridZ", C.c_int)]


class LusPbfParams(C.Structure):
    _fields_ = [
        ("h", C.c_float), ("f", C.c_float), ("cell_size", C.c_float),
        ("gridX", C.c_int), ("gridY", C.c_int), ("gridZ", C.c_int),
        ("lo", C.c_float * 3), ("hi", C.c_float * 3), ("gravity", C.c_float * 3),
        ("mass", C.c_float), ("rest_density", C.c_float), ("epsilon", C.c_float),
        ("scorr_k", C.c_float), ("scorr_dq", C.c_float), ("scorr_n", C.c_int),
        ("c_xsph", C.c_float), ("iterations", C.c_int), ("exact", C.c_int),
    ]


class LusSandParams(C.Structure):
    _fields_ = [
        ("particleRadius", C.c_float), ("particleDiameter", C.c_float),
        ("kernelRadius", C.c_float), ("mass", C.c_float),
        ("domain", C.c_float * 3), ("grid", LusGrid), ("gravity", C.c_float * 3),
        ("collision_coeff", C.c_float), ("friction_coeff", C.c_float),
        ("mu_s", C.c_float), ("mu_k", C.c_float), ("substeps", C.c_int),
        ("attract_radius", C.c_float), ("blow_radius", C.c_float),
        ("attract_coeff", C.c_float), ("blow_coeff", C.c_float),
        ("ptr_solid_start", C.c_int),
    ]


class LusFluidParams(C.Structure):
    _fields_ = [
        ("particleRadius", C.c_float), ("kernelRadius", C.c_float), ("kernelFactor", C.c_float),
        ("cubic_kernel_k", C.c_float), ("cubic_kernel_l", C.c_float),
        ("mass", C.c_float), ("rest_density", C.c_float), ("relaxation_epsilon", C.c_float),
        ("s_corr_dq", C.c_float), ("s_corr_k", C.c_float), ("s_corr_n", C.c_float),
        ("gravity", C.c_float * 3), ("domain", C.c_float * 3), ("grid", LusGrid),
    ]


_P = C.POINTER
_f = _P(C.c_float)
_i = _P(C.c_int)
_l = _P(C.c_long)

class LusMixParams(C.Structure):
    _fields_ = [("diameter", C.c_float), ("collision", C.c_float), ("friction", C.c_float),
                ("mu_s", C.c_float), ("mu_k", C.c_float)]


class LusMlpFitParams(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("lr_final", C.c_float),
                ("val_fraction", C.c_float), ("seed", C.c_ulonglong)]


_SIGS = {
    "lus_gpu_last_error": (C.c_int, [C.c_char_p, C.c_int]),
    "lus_gpu_device_count": (C.c_int, []),
    "lus_cell_ids": (C.c_int, [_P(LusGrid), C.c_int, _f, _i]),
    "lus_counting_sort": (C.c_int, [_i, _i, C.c_size_t, C.c_size_t, _i]),
    "lus_counting_sort_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lus_neighbors_fluid": (C.c_int, [_P(LusGrid), C.c_float, C.c_int, _f, C.c_int, C.c_int, _f, _i,
                                      _l, _i, C.c_long, _l]),
    "lus_gpu_stream_copy": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, _f]),
    "lus_gpu_create": (C.c_int, [_P(C.c_void_p), C.c_int]),
    "lus_gpu_destroy": (None, [C.c_void_p]),
    "lus_gpu_sync": (C.c_int, [C.c_void_p]),
    "lus_pbf_init": (C.c_int, [C.c_void_p, _P(LusPbfParams), C.c_int, _f, _f]),
    "lus_pbf_batch_init": (C.c_int, [C.c_void_p, _P(LusPbfParams), C.c_int, _i, _f, _f, _f]),
    "lus_mix_init": (C.c_int, [C.c_void_p, _P(LusPbfParams), _P(LusMixParams), C.c_int, _f, _f, _i]),
    "lus_pbf_step": (C.c_int, [C.c_void_p, C.c_float, C.c_int]),
    "lus_pbf_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_set_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "lus_pbf_set_build": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_set_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_set_sort_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_set_key_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_set_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_pbf_step_latency": (C.c_int, [C.c_void_p, C.c_float, C.c_int, _f]),
    "lus_pbf_download": (C.c_int, [C.c_void_p, C.c_int, _f, _f, _i]),
    "lus_pbf_last_max_constraint": (C.c_int, [C.c_void_p, _f]),
    "lus_pbf_bench": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.c_int, _f, _f, _i]),
    "lus_pbf_time_kernels": (C.c_int, [C.c_void_p, C.c_int, _f]),
    "lus_pbf_set_network": (C.c_int, [C.c_void_p, _f, C.c_int]),
    "lus_pbf_set_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int]),
    "lus_pbf_learned_stats": (C.c_int, [C.c_void_p, _P(C.c_longlong), _P(C.c_longlong), _f, _f, C.c_int, _i]),
    "lus_pbf_retest_count": (C.c_int, [C.c_void_p, _P(C.c_longlong)]),
    "lus_mlp_trainer_create": (C.c_int, [C.c_void_p, C.c_long, _P(C.c_void_p)]),
    "lus_mlp_trainer_destroy": (None, [C.c_void_p]),
    "lus_mlp_trainer_add": (C.c_int, [C.c_void_p, _f, _f, C.c_long]),
    "lus_mlp_trainer_rows": (C.c_int, [C.c_void_p, _l]),
    "lus_mlp_trainer_download": (C.c_int, [C.c_void_p, C.c_long, C.c_long, _f, _f]),
    "lus_pbf_record_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int, _l]),
    "lus_mlp_fit": (C.c_int, [C.c_void_p, _P(LusMlpFitParams), _f, _f]),
    "lus_mlp_trainer_network": (C.c_int, [C.c_void_p, _f]),
    "lus_mlp_loss_grad": (C.c_int, [C.c_void_p, _f, C.c_int, _f, _f, C.c_float, _f, _f]),
    "lus_mlp_infer": (C.c_int, [C.c_void_p, _f, C.c_int, _f, _f]),
    "lus_sand_init": (C.c_int, [C.c_void_p, _P(LusSandParams), C.c_int, _f, _f, _i, C.c_int, _f, _i]),
    "lus_sand_step": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.c_int, _f]),
    "lus_sand_set_list_capacity": (C.c_int, [C.c_void_p, C.c_long]),
    "lus_sand_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_sand_download": (C.c_int, [C.c_void_p, _f, _f, _f, _i, _i]),
    "lus_sand_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_sand_neighbors": (C.c_int, [C.c_void_p, _l, _i, C.c_long, _l]),
    "lus_sand_query_cells": (C.c_int, [C.c_void_p, _f, _f, C.c_int, _i]),
    "lus_sand_sink_evict": (C.c_int, [C.c_void_p, _i, C.c_int, _i]),
    "lus_sand_player_box": (C.c_int, [C.c_void_p, _f, _f, C.c_int, _i, _f, _i]),
    "lus_profiling_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "lus_chunk_from_grid": (C.c_int, [C.c_void_p, _i, C.c_int, C.c_int, C.c_int, _f, C.c_float, C.c_int, C.c_int,
                                      C.c_int, _f, _f, _f, _i]),
    "lus_grid_from_voxels": (C.c_int, [C.c_void_p, _P(C.c_ubyte), C.c_int, C.c_int, C.c_int, _P(C.c_ubyte), _i, _f,
                                       _i]),
    "lus_fluid_init": (C.c_int, [C.c_void_p, _P(LusFluidParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _f, _f, _f, C.c_int, _f]),
    "lus_fluid_step": (C.c_int, [C.c_void_p, C.c_float]),
    "lus_fluid_download": (C.c_int, [C.c_void_p, _f, _f, _f, _f]),
    "lus_fluid_stats": (C.c_int, [C.c_void_p, _i, _l]),
    "lus_get_num_observation": (C.c_long, []),
    "lus_get_cycles": (C.c_longlong, [C.c_int]),
    "lus_get_duration": (C.c_double, [C.c_int]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LusError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                           " or `python lustrine_amd/build.py`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(512)
    load().lus_gpu_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise LusError(f"{what} failed (code {rc}): {last_error()}")


def device_count() -> int:
    return load().lus_gpu_device_count()


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr_f(a):
    return None if a is None else a.ctypes.data_as(_f)


def _ptr_i(a):
    return None if a is None else a.ctypes.data_as(_i)


def _ptr_l(a):
    return None if a is None else a.ctypes.data_as(_l)


def grid(cell_size, gx, gy, gz) -> LusGrid:
    return LusGrid(cell_size, gx, gy, gz)


# ------------------------------------------------------------------ building blocks
def cell_ids(g: LusGrid, pos) -> np.ndarray:
    pos = _f32(pos, (-1, 3))
    out = np.zeros(pos.shape[0], np.int32)
    _check(load().lus_cell_ids(C.byref(g), pos.shape[0], _ptr_f(pos), _ptr_i(out)), "lus_cell_ids")
    return out


def fluid_params(sim) -> LusFluidParams:
    """LusFluidParams from any object with the reference Simulation's field
    names (particleRadius, kernelRadius, kernelFactor, cubic_kernel_k/l, mass,
    rest_density, relaxation_epsilon, s_corr_dq/k/n, gravity, domainX/Y/Z,
    cell_size, gridX/Y/Z)."""
    p = LusFluidParams()
    for f in ("particleRadius", "kernelRadius", "kernelFactor", "cubic_kernel_k", "cubic_kernel_l", "mass",
              "rest_density", "relaxation_epsilon", "s_corr_dq", "s_corr_k", "s_corr_n"):
        setattr(p, f, getattr(sim, f))
    for a in range(3):
        p.gravity[a] = sim.gravity[a]
    p.domain[0], p.domain[1], p.domain[2] = sim.domainX, sim.domainY, sim.domainZ
    p.grid = grid(sim.cell_size, sim.gridX, sim.gridY, sim.gridZ)
    return p


def counting_sort_noalloc(keys_dev, n, num_cells, counts_dev, sorted_dev, stream=None):
    """lus_counting_sort_noalloc on device pointers (ints): enqueues the sort on
    `stream` (None: the null stream) and returns without waiting."""
    _check(load().lus_counting_sort_noalloc(keys_dev, n, num_cells, counts_dev, sorted_dev, stream),
           "lus_counting_sort_noalloc")


def counting_sort(keys, num_cells):
    keys = np.ascontiguousarray(keys, np.int32)
    counts = np.zeros(num_cells + 1, np.int32)
    out = np.zeros(keys.shape[0], np.int32)
    _check(load().lus_counting_sort(_ptr_i(counts), _ptr_i(keys), keys.shape[0], num_cells, _ptr_i(out)),
           "lus_counting_sort")
    return out, counts


def neighbors_fluid(g: LusGrid, kernel_radius, sand_pos, sand_start=0, solid_pos=None, solid_index=None):
    sand_pos = _f32(sand_pos, (-1, 3))
    ns = 0 if solid_pos is None else int(np.asarray(solid_pos).shape[0])
    solid_pos = _f32(solid_pos, (-1, 3)) if ns else None
    solid_index = np.ascontiguousarray(solid_index, np.int32) if ns else None
    n = sand_pos.shape[0]
    off = np.zeros(n + 1, np.int64)
    total = C.c_long(0)
    cap = 0
    lib = load()
    rc = lib.lus_neighbors_fluid(C.byref(g), kernel_radius, n, _ptr_f(sand_pos), sand_start, ns,
                                 _ptr_f(solid_pos), _ptr_i(solid_index), _ptr_l(off), None, cap,
                                 C.byref(total))
    if rc != 0 and total.value <= cap:
        _check(rc, "lus_neighbors_fluid")
    idx = np.zeros(max(total.value, 1), np.int32)
    _check(lib.lus_neighbors_fluid(C.byref(g), kernel_radius, n, _ptr_f(sand_pos), sand_start, ns,
                                   _ptr_f(solid_pos), _ptr_i(solid_index), _ptr_l(off), _ptr_i(idx),
                                   idx.shape[0], C.byref(total)), "lus_neighbors_fluid")
    return off, idx[: total.value]


# ------------------------------------------------------------------ params
def pbf_params(cfg, iterations=None, exact=True) -> LusPbfParams:
    """LusPbfParams of a scenes.PbfConfig.  exact=True: operation order of
    the CPU oracle (bit-identical results); False: fused fast path."""
    p = LusPbfParams()
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
    p.exact = 1 if exact else 0
    return p


def mix_params(diameter=1.0, collision=0.8, friction=0.7, mu_s=0.6, mu_k=0.4) -> LusMixParams:
    """Sand constants of the mixed scene: contact distance 2r, the collision
    and friction coefficients of src/Simulate.cpp:159-161, friction limits."""
    return LusMixParams(diameter, collision, friction, mu_s, mu_k)


def sand_params(X, Y, Z, radius=0.5, diameter=1.0, kernel_radius_scale=3.1, substeps=4,
                gravity=(0.0, -10.0, 0.0), mass=5.0, collision=0.8, friction=0.7, mu_s=0.95,
                mu_k=0.9) -> LusSandParams:
    """simulate_sand constants of a scene built like init_simulation_extra_parameters
    (src/Lustrine.cpp:383-386, 543-558); coefficients of src/Simulate.cpp:159-163."""
    p = LusSandParams()
    p.particleRadius = radius
    p.particleDiameter = diameter
    h = np.float32(kernel_radius_scale) * np.float32(radius)
    p.kernelRadius = h
    p.mass = mass
    p.domain[:] = [float(X), float(Y), float(Z)]
    cs = np.float32(1.0) * h
    p.grid = LusGrid(cs, int(np.float32(X) / cs) + 1, int(np.float32(Y) / cs) + 1, int(np.float32(Z) / cs) + 1)
    p.gravity[:] = list(gravity)
    p.collision_coeff, p.friction_coeff, p.mu_s, p.mu_k = collision, friction, mu_s, mu_k
    p.substeps = substeps
    p.attract_radius, p.blow_radius, p.attract_coeff, p.blow_coeff = 1.5, 2.0, 1000.0, 500.0
    p.ptr_solid_start = 1 << 30
    return p


# ------------------------------------------------------------------ context
class Gpu:
    """One device context (one HIP stream) holding a PBF and/or a sand scene."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = C.c_void_p()
        _check(self._lib.lus_gpu_create(C.byref(h), device), "lus_gpu_create")
        self._h = h
        self.n = 0
        self.n_sand = 0
        self._trainers = weakref.WeakSet()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for t in list(getattr(self, "_trainers", ())):  # a trainer lives inside its context
                t.close()
            self._lib.lus_gpu_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self._lib.lus_gpu_sync(self._h), "lus_gpu_sync")

    def stream_copy_gbps(self, nbytes=1 << 30, reps=20):
        """Measured HBM ceiling: GB/s of read + write of a float4 copy."""
        us = C.c_float(0)
        _check(self._lib.lus_gpu_stream_copy(self._h, nbytes, reps, C.byref(us)), "lus_gpu_stream_copy")
        return 2.0 * nbytes / (us.value * 1e-6) / 1e9, us.value

    # --- PBF
    def pbf_init(self, params: LusPbfParams, pos, vel=None):
        pos = _f32(pos, (-1, 3))
        vel = None if vel is None else _f32(vel, (-1, 3))
        _check(self._lib.lus_pbf_init(self._h, C.byref(params), pos.shape[0], _ptr_f(pos), _ptr_f(vel)),
               "lus_pbf_init")
        self.n = pos.shape[0]

    def pbf_batch_init(self, params: LusPbfParams, counts, boxes, pos, vel=None):
        counts = np.ascontiguousarray(counts, np.int32)
        boxes = _f32(boxes, (-1, 3))
        pos = _f32(pos, (-1, 3))
        vel = None if vel is None else _f32(vel, (-1, 3))
        assert pos.shape[0] == int(counts.sum()) and boxes.shape[0] == counts.shape[0]
        _check(self._lib.lus_pbf_batch_init(self._h, C.byref(params), counts.shape[0], _ptr_i(counts),
                                            _ptr_f(boxes), _ptr_f(pos), _ptr_f(vel)), "lus_pbf_batch_init")
        self.n = pos.shape[0]

    def mix_init(self, params: LusPbfParams, mix: LusMixParams, pos, phase, vel=None):
        pos = _f32(pos, (-1, 3))
        phase = np.ascontiguousarray(phase, np.int32)
        vel = None if vel is None else _f32(vel, (-1, 3))
        assert phase.shape[0] == pos.shape[0]
        _check(self._lib.lus_mix_init(self._h, C.byref(params), C.byref(mix), pos.shape[0], _ptr_f(pos),
                                      _ptr_f(vel), _ptr_i(phase)), "lus_mix_init")
        self.n = pos.shape[0]

    def pbf_step(self, dt, steps=1):
        _check(self._lib.lus_pbf_step(self._h, dt, steps), "lus_pbf_step")

    def pbf_set_graph(self, on=True):
        _check(self._lib.lus_pbf_set_graph(self._h, int(bool(on))), "lus_pbf_set_graph")

    def pbf_set_launch(self, block=256):
        _check(self._lib.lus_pbf_set_launch(self._h, block), "lus_pbf_set_launch")

    def pbf_set_tiling(self, tiled=True, zt=16):
        _check(self._lib.lus_pbf_set_tiling(self._h, int(bool(tiled)), zt), "lus_pbf_set_tiling")

    def pbf_set_build(self, staged=True):
        """staged: the list build tests wave-staged candidates in LDS; else
        per-lane gathers (same results)."""
        _check(self._lib.lus_pbf_set_build(self._h, int(bool(staged))), "lus_pbf_set_build")

    def pbf_set_lanes(self, lanes=0):
        """Lanes per particle in fused list walks: 0 automatic, 1 or 4."""
        _check(self._lib.lus_pbf_set_lanes(self._h, int(lanes)), "lus_pbf_set_lanes")

    def pbf_set_sort_tile(self, cells=0):
        """Cells per block of the counting sort's prefix scan: 1024, 2048 or 4096;
        0: by the grid's cell count (4096 from 2**19 cells, else 2048)."""
        _check(self._lib.lus_pbf_set_sort_tile(self._h, int(cells)), "lus_pbf_set_sort_tile")

    def pbf_set_key_fusion(self, on=True):
        """The XSPH pass writes the next step's cell keys (True) or every step runs the key kernel."""
        _check(self._lib.lus_pbf_set_key_fusion(self._h, int(bool(on))), "lus_pbf_set_key_fusion")

    def pbf_step_latency(self, dt, steps):
        ms = np.zeros(steps, np.float32)
        _check(self._lib.lus_pbf_step_latency(self._h, dt, steps, _ptr_f(ms)), "lus_pbf_step_latency")
        return ms

    def pbf_download(self, order=0):
        pos = np.zeros((self.n, 3), np.float32)
        vel = np.zeros((self.n, 3), np.float32)
        ids = np.zeros(self.n, np.int32)
        _check(self._lib.lus_pbf_download(self._h, order, _ptr_f(pos), _ptr_f(vel), _ptr_i(ids)),
               "lus_pbf_download")
        return pos, vel, ids

    def pbf_last_max_constraint(self):
        v = C.c_float(0)
        _check(self._lib.lus_pbf_last_max_constraint(self._h, C.byref(v)), "lus_pbf_last_max_constraint")
        return v.value

    def pbf_bench(self, dt, warmup, steps, phases=True):
        ms = C.c_float(0)
        ph = np.zeros(NPHASES, np.float32)
        nl = np.zeros(NPHASES, np.int32)
        _check(self._lib.lus_pbf_bench(self._h, dt, warmup, steps, C.byref(ms),
                                       _ptr_f(ph) if phases else None, _ptr_i(nl) if phases else None),
               "lus_pbf_bench")
        return ms.value, (dict(zip(PHASES, ph.tolist())) if phases else None), \
            (dict(zip(PHASES, nl.tolist())) if phases else None)

    def pbf_time_kernels(self, reps=20):
        """Average us per launch of the lambda, delta-p and XSPH kernels."""
        us = np.zeros(4, np.float32)
        _check(self._lib.lus_pbf_time_kernels(self._h, reps, _ptr_f(us)), "lus_pbf_time_kernels")
        return {"lambda": float(us[0]), "deltap": float(us[1]), "xsph": float(us[2]),
                "lambda_build": float(us[3])}

    # --- learned warm start
    def pbf_set_network(self, net):
        net = _f32(net, (-1,))
        _check(self._lib.lus_pbf_set_network(self._h, _ptr_f(net), net.shape[0]), "lus_pbf_set_network")

    def pbf_set_mode(self, learned=False, learned_iterations=1, bound=0.01, check=False):
        _check(self._lib.lus_pbf_set_mode(self._h, int(bool(learned)), learned_iterations, bound,
                                          int(bool(check))), "lus_pbf_set_mode")

    def pbf_retest_count(self):
        """Particles of the last step that took the re-test path (a list past
        128 entries or an entry beyond 16 bits)."""
        c = C.c_longlong(0)
        _check(self._lib.lus_pbf_retest_count(self._h, C.byref(c)), "lus_pbf_retest_count")
        return c.value

    def pbf_learned_stats(self, with_mean=False):
        """(measured steps, plain reruns, per-step max density error e, oldest
        first[, per-step mean compression mean_i max(0, rho_i/rho0 - 1)])."""
        steps, reruns = C.c_longlong(0), C.c_longlong(0)
        errs = np.zeros(DENSITY_HISTORY, np.float32)
        errm = np.zeros(DENSITY_HISTORY, np.float32)
        ne = C.c_int(0)
        _check(self._lib.lus_pbf_learned_stats(self._h, C.byref(steps), C.byref(reruns), _ptr_f(errs),
                                               _ptr_f(errm), DENSITY_HISTORY, C.byref(ne)),
               "lus_pbf_learned_stats")
        out = (steps.value, reruns.value, errs[: ne.value].copy())
        return out + (errm[: ne.value].copy(),) if with_mean else out

    def mlp_loss_grad(self, params, X, Y, inv_scale=1.0):
        params = _f32(params, (MLP_PARAMS,))
        X = _f32(X, (-1, FEATURES))
        Y = _f32(Y, (-1, 3))
        loss = C.c_float(0)
        grad = np.zeros(MLP_PARAMS, np.float32)
        _check(self._lib.lus_mlp_loss_grad(self._h, _ptr_f(params), X.shape[0], _ptr_f(X), _ptr_f(Y),
                                           inv_scale, C.byref(loss), _ptr_f(grad)), "lus_mlp_loss_grad")
        return loss.value, grad

    def mlp_infer(self, net, feat):
        net = _f32(net, (NET_FLOATS,))
        feat = _f32(feat, (-1, FEATURES))
        out = np.zeros((feat.shape[0], 3), np.float32)
        _check(self._lib.lus_mlp_infer(self._h, _ptr_f(net), feat.shape[0], _ptr_f(feat), _ptr_f(out)),
               "lus_mlp_infer")
        return out

    def trainer(self, capacity):
        return Trainer(self, capacity)

    # --- sand
    def sand_init(self, params: LusSandParams, pos, vel=None, attracted=None, solid_pos=None,
                  solid_index=None):
        pos = _f32(pos, (-1, 3))
        vel = None if vel is None else _f32(vel, (-1, 3))
        att = None if attracted is None else np.ascontiguousarray(attracted, np.int32)
        ns = 0 if solid_pos is None else int(np.asarray(solid_pos).shape[0])
        sp = _f32(solid_pos, (-1, 3)) if ns else None
        si = np.ascontiguousarray(solid_index, np.int32) if ns else None
        _check(self._lib.lus_sand_init(self._h, C.byref(params), pos.shape[0], _ptr_f(pos), _ptr_f(vel),
                                       _ptr_i(att), ns, _ptr_f(sp), _ptr_i(si)), "lus_sand_init")
        self.n_sand = pos.shape[0]

    def sand_step(self, dt, attract=False, blow=False, player=None):
        pl = None if player is None else _f32(player, (3,))
        _check(self._lib.lus_sand_step(self._h, dt, int(attract), int(blow), _ptr_f(pl)), "lus_sand_step")

    def sand_set_list_capacity(self, entries):
        """A fixed neighbour-list buffer of `entries` (steps then do not wait); 0: the growing buffer."""
        _check(self._lib.lus_sand_set_list_capacity(self._h, int(entries)), "lus_sand_set_list_capacity")

    def sand_set_graph(self, on=True):
        """Replay each sand step as a hipGraph (needs a fixed list capacity)."""
        _check(self._lib.lus_sand_set_graph(self._h, int(bool(on))), "lus_sand_set_graph")

    def sand_set_variant(self, variant):
        """0 simulate_sand, 1 simulate_sand_credits, 2 simulate_sand_v3."""
        _check(self._lib.lus_sand_set_variant(self._h, int(variant)), "lus_sand_set_variant")

    def sand_download(self):
        n = self.n_sand
        pos = np.zeros((n, 3), np.float32)
        ps = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        att = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        _check(self._lib.lus_sand_download(self._h, _ptr_f(pos), _ptr_f(ps), _ptr_f(vel), _ptr_i(att),
                                           _ptr_i(perm)), "lus_sand_download")
        return pos, ps, vel, att, perm

    # --- simulate_fluid (src/Simulate.cpp:27-115)
    def fluid_init(self, params: LusFluidParams, positions, positions_star, velocities, sand, solid,
                   lambdas=None):
        """The reference's slot layout: arrays of `total` rows; sand = (start,
        end) of the moving slots, solid = (start, end) of the solid slots."""
        p = _f32(positions, (-1, 3))
        ps = _f32(positions_star, (-1, 3))
        v = _f32(velocities, (-1, 3))
        total = p.shape[0]
        assert ps.shape[0] == total and v.shape[0] >= sand[1]
        lam = np.zeros(total, np.float32) if lambdas is None else np.ascontiguousarray(lambdas, np.float32)
        self.fluid_total, self.fluid_nl = total, lam.shape[0]
        _check(self._lib.lus_fluid_init(self._h, C.byref(params), total, sand[0], sand[1], solid[0], solid[1],
                                        _ptr_f(p), _ptr_f(ps), _ptr_f(v), lam.shape[0], _ptr_f(lam)),
               "lus_fluid_init")

    def fluid_step(self, dt):
        _check(self._lib.lus_fluid_step(self._h, float(dt)), "lus_fluid_step")

    def fluid_download(self):
        """(positions, positions_star, velocities, lambdas) in the slot layout;
        velocities hold only the moving slots' values."""
        n = self.fluid_total
        p, ps, v = (np.zeros((n, 3), np.float32) for _ in range(3))
        lam = np.zeros(self.fluid_nl, np.float32)
        _check(self._lib.lus_fluid_download(self._h, _ptr_f(p), _ptr_f(ps), _ptr_f(v), _ptr_f(lam)),
               "lus_fluid_download")
        return p, ps, v, lam

    def fluid_stats(self):
        lv, ent = C.c_int(0), C.c_long(0)
        _check(self._lib.lus_fluid_stats(self._h, C.byref(lv), C.byref(ent)), "lus_fluid_stats")
        return lv.value, ent.value

    def sand_neighbors(self):
        n = self.n_sand
        off = np.zeros(n + 1, np.int64)
        total = C.c_long(0)
        self._lib.lus_sand_neighbors(self._h, _ptr_l(off), None, 0, C.byref(total))
        idx = np.zeros(max(total.value, 1), np.int32)
        _check(self._lib.lus_sand_neighbors(self._h, _ptr_l(off), _ptr_i(idx), idx.shape[0], C.byref(total)),
               "lus_sand_neighbors")
        return off, idx[: total.value]

    def sand_query_cells(self, min_pos, max_pos, include_solid):
        """query_cell_num_particles over the last step's grid."""
        lo, hi = _f32(min_pos, (3,)), _f32(max_pos, (3,))
        out = C.c_int(0)
        _check(self._lib.lus_sand_query_cells(self._h, _ptr_f(lo), _ptr_f(hi), int(bool(include_solid)),
                                              C.byref(out)), "lus_sand_query_cells")
        return out.value

    def sand_sink_evict(self, cells):
        """Sink eviction of the listed cells; returns the new sand count."""
        c = np.ascontiguousarray(cells, np.int32)
        out = C.c_int(0)
        _check(self._lib.lus_sand_sink_evict(self._h, _ptr_i(c) if c.size else None, int(c.size),
                                             C.byref(out)), "lus_sand_sink_evict")
        self.n_sand = out.value
        return out.value


    def chunk_from_grid(self, cells, X, Y, Z, position, cell_size, subdivision, mask_out=False, num_occupied=None,
                        colors=None):
        """init_chunk_from_grid on the device: (positions, colors or None, generated)."""
        c = np.ascontiguousarray(cells, np.int32).reshape(-1)
        n_occ = int(np.count_nonzero(c)) if num_occupied is None else int(num_occupied)
        cap = n_occ * int(subdivision) ** 3
        pos = np.zeros((max(cap, 1), 3), np.float32)
        cc = None if colors is None else _f32(colors, (-1, 4))
        col = None if colors is None else np.zeros((max(cap, 1), 4), np.float32)
        gen = C.c_int(0)
        _check(self._lib.lus_chunk_from_grid(self._h, _ptr_i(c), X, Y, Z, _ptr_f(_f32(position, (3,))),
                                             float(cell_size), int(subdivision), int(bool(mask_out)), n_occ,
                                             _ptr_f(cc), _ptr_f(pos), _ptr_f(col), C.byref(gen)),
               "lus_chunk_from_grid")
        return pos[:cap], (None if col is None else col[:cap]), gen.value

    def grid_from_voxels(self, voxels, sx, sy, sz, palette):
        """init_grid_from_magika_voxel after parsing: (cells, colors, occupied)."""
        v = np.ascontiguousarray(voxels, np.uint8).reshape(-1)
        pal = np.ascontiguousarray(palette, np.uint8).reshape(-1)
        n = sx * sy * sz
        cells = np.zeros(n, np.int32)
        colors = np.zeros((n, 4), np.float32)
        occ = C.c_int(0)
        u8 = C.POINTER(C.c_ubyte)
        _check(self._lib.lus_grid_from_voxels(self._h, v.ctypes.data_as(u8), sx, sy, sz, pal.ctypes.data_as(u8),
                                              _ptr_i(cells), _ptr_f(colors), C.byref(occ)), "lus_grid_from_voxels")
        return cells, colors, occ.value

    def profiling(self, on=True):
        """Phase events into the reference's 7 profiling slots (see profile_slots)."""
        _check(self._lib.lus_profiling_enable(self._h, int(bool(on))), "lus_profiling_enable")

    def sand_player_box(self, player, half_extent, max_out=100):
        """Slots and positions of the first max_out sand particles in the player's box."""
        pl, he = _f32(player, (3,)), _f32(half_extent, (3,))
        slots = np.zeros(max(max_out, 1), np.int32)
        pos = np.zeros((max(max_out, 1), 3), np.float32)
        cnt = C.c_int(0)
        _check(self._lib.lus_sand_player_box(self._h, _ptr_f(pl), _ptr_f(he), int(max_out), _ptr_i(slots),
                                             _ptr_f(pos), C.byref(cnt)), "lus_sand_player_box")
        return slots[: cnt.value], pos[: cnt.value]


def profile_slots():
    """(cycles, seconds) of the last profiled step in the reference's slots."""
    lib = load()
    n = int(lib.lus_get_num_observation())
    return ([int(lib.lus_get_cycles(i)) for i in range(n)], [float(lib.lus_get_duration(i)) for i in range(n)])


class Trainer:
    """Device-resident training set of the warm-start network (LusMlpTrainer)."""

    def __init__(self, gpu: Gpu, capacity: int):
        self._lib = gpu._lib
        self.gpu = gpu
        h = C.c_void_p()
        _check(self._lib.lus_mlp_trainer_create(gpu._h, capacity, C.byref(h)), "lus_mlp_trainer_create")
        self._h = h
        gpu._trainers.add(self)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lus_mlp_trainer_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def rows(self) -> int:
        r = C.c_long(0)
        _check(self._lib.lus_mlp_trainer_rows(self._h, C.byref(r)), "lus_mlp_trainer_rows")
        return r.value

    def add(self, feat, target):
        feat = _f32(feat, (-1, FEATURES))
        target = _f32(target, (-1, 3))
        assert feat.shape[0] == target.shape[0]
        _check(self._lib.lus_mlp_trainer_add(self._h, _ptr_f(feat), _ptr_f(target), feat.shape[0]),
               "lus_mlp_trainer_add")

    def download(self, first=0, rows=None):
        rows = self.rows - first if rows is None else rows
        feat = np.zeros((rows, FEATURES), np.float32)
        tgt = np.zeros((rows, 3), np.float32)
        _check(self._lib.lus_mlp_trainer_download(self._h, first, rows, _ptr_f(feat), _ptr_f(tgt)),
               "lus_mlp_trainer_download")
        return feat, tgt

    def record_step(self, dt, stride=1) -> int:
        added = C.c_long(0)
        _check(self._lib.lus_pbf_record_step(self.gpu._h, self._h, dt, stride, C.byref(added)),
               "lus_pbf_record_step")
        return added.value

    def fit(self, epochs=20, batch=4096, lr=2e-3, lr_final=1e-4, val_fraction=0.05, seed=1):
        fp = LusMlpFitParams(epochs, batch, lr, lr_final, val_fraction, seed)
        tl = np.zeros(epochs, np.float32)
        vl = np.zeros(epochs, np.float32)
        _check(self._lib.lus_mlp_fit(self._h, C.byref(fp), _ptr_f(tl), _ptr_f(vl)), "lus_mlp_fit")
        return tl, vl

    def network(self) -> np.ndarray:
        net = np.zeros(NET_FLOATS, np.float32)
        _check(self._lib.lus_mlp_trainer_network(self._h, _ptr_f(net)), "lus_mlp_trainer_network")
        return net
