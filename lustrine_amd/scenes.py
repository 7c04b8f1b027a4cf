"""Scene construction for the particle substep: jittered lattices, heightfield
solids and the five benchmark configurations.

Lattice order follows ``init_chunk_from_grid`` (src/Lustrine.cpp:676-722):
loops x (outer), y, z (inner); site = (x, y, z) * cell + grid.position +
cell / 2, evaluated in float32 in that order.  Jitter is drawn from
Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
SC'11) with key = (seed, 0) and counter = (particle index, axis, 0, 0), so any
implementation reproduces the same draws.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, asdict

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(counter: np.ndarray, key) -> np.ndarray:
    """Philox4x32 with 10 rounds.  counter: (n, 4) uint32; key: 2 uint32."""
    c = np.asarray(counter, dtype=np.uint32).reshape(-1, 4)
    c0 = c[:, 0].astype(np.uint64)
    c1 = c[:, 1].astype(np.uint64)
    c2 = c[:, 2].astype(np.uint64)
    c3 = c[:, 3].astype(np.uint64)
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    for r in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        n0 = hi1 ^ c1 ^ np.uint64(k0)
        n1 = lo1
        n2 = hi0 ^ c3 ^ np.uint64(k1)
        n3 = lo0
        c0, c1, c2, c3 = n0, n1, n2, n3
        if r < 9:
            k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return np.stack([c0, c1, c2, c3], axis=1).astype(np.uint32)


def uniform01(n: int, axis_count: int, seed: int) -> np.ndarray:
    """u[i, a] in [0, 1) from counter (i, a, 0, 0), 24-bit resolution."""
    idx = np.repeat(np.arange(n, dtype=np.uint64), axis_count)
    ax = np.tile(np.arange(axis_count, dtype=np.uint64), n)
    ctr = np.zeros((n * axis_count, 4), dtype=np.uint32)
    ctr[:, 0] = (idx & _MASK).astype(np.uint32)
    ctr[:, 1] = ax.astype(np.uint32)
    out = philox4x32_10(ctr, (seed & 0xFFFFFFFF, 0))[:, 0]
    return ((out >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)).reshape(n, axis_count)


def lattice(nx: int, ny: int, nz: int, cell: float = 1.0, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Sites of a full nx*ny*nz grid in init_chunk_from_grid order, float32 (n, 3)."""
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    c = np.float32(cell)
    pos = ijk.astype(np.float32) * c
    pos = pos + np.asarray(origin, dtype=np.float32)
    pos = pos + np.float32(cell * 0.5)
    return pos.astype(np.float32)


def jittered_lattice(nx, ny, nz, cell=1.0, origin=(0.0, 0.0, 0.0), amplitude=0.005, seed=0):
    pos = lattice(nx, ny, nz, cell, origin)
    u = uniform01(pos.shape[0], 3, seed)
    j = ((2.0 * u - 1.0) * amplitude).astype(np.float32)
    return (pos + j).astype(np.float32)


@dataclass
class PbfConfig:
    """All constants of one plain-PBF scene (fluid in a closed box)."""
    name: str
    lattice: tuple                 # (nx, ny, nz) particles
    box: tuple                     # (X, Y, Z) domain in particle diameters
    origin: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.5
    kernel_radius_scale: float = 4.0
    kernel_factor: float = 0.5
    mass: float = 5.0              # src/Simulation.hpp:158
    # Relaxation epsilon and the s_corr strength.  The reference's Simulation
    # defaults (epsilon 10, k 1, src/Simulation.hpp:163-168) go with its cubic
    # kernel; with these kernels the rest-state gradient sum
    # (m/rho0)^2 sum |grad W|^2 is about 0.65, so epsilon 10 shrinks every
    # Jacobi lambda 16-fold and k = 1 makes s_corr an attraction as large as the
    # pressure terms.  epsilon = 0.1 (~0.15 of the gradient sum: it only guards
    # the sparse surface) and k = 0.1 (Macklin & Mueller's value, with
    # dq = 0.25 h) halve the per-step density error at every iteration count.
    epsilon: float = 0.1
    scorr_k: float = 0.1
    scorr_n: int = 4
    scorr_dq: float = 0.5
    c_xsph: float = 0.01
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1.0 / 60.0
    iterations: int = 4
    steps: int = 100
    jitter: float = 0.005          # +-0.01 r
    seed: int = 0

    @property
    def h(self) -> float:
        return float(np.float32(self.kernel_radius_scale) * np.float32(self.radius))

    @property
    def cell_size(self) -> float:
        return self.h

    @property
    def grid(self) -> tuple:
        cs = np.float32(self.h)
        return tuple(int(np.float32(d) / cs) + 1 for d in self.box)

    @property
    def n(self) -> int:
        return int(np.prod(self.lattice))

    def rest_density(self) -> float:
        """m * sum_j W(r_ij) at an interior site of the unjittered lattice
        (spacing 2r), self term once, W of the kernel the solver uses."""
        h = self.h
        f = self.kernel_factor
        hf = h * f
        p6 = 315.0 / (64.0 * 3.14 * hf ** 9)
        d = 2.0 * self.radius
        k = int(math.ceil(h / d)) + 1
        s = 0.0
        for a in range(-k, k + 1):
            for b in range(-k, k + 1):
                for c in range(-k, k + 1):
                    r2 = (a * d) ** 2 + (b * d) ** 2 + (c * d) ** 2
                    if r2 <= h * h:
                        s += (hf * hf - f * f * r2) ** 3
        return float(np.float32(self.mass * p6 * s))

    def positions(self) -> np.ndarray:
        return jittered_lattice(*self.lattice, cell=2.0 * self.radius, origin=self.origin,
                                amplitude=self.jitter, seed=self.seed)

    def as_dict(self):
        d = asdict(self)
        d.update(h=self.h, grid=self.grid, rest_density=self.rest_density(), n=self.n)
        return d


def cfg0(**kw) -> PbfConfig:
    """Plain PBF dam break, 32,768 particles (32^3) in a 64^3 box, 200 steps."""
    return PbfConfig(name="pbf_dam_break_32k", lattice=(32, 32, 32), box=(64, 64, 64),
                     steps=200, **kw)


def cfg2(**kw) -> PbfConfig:
    """Plain PBF dam break at scale: 1,048,576 particles (128x64x128) in a
    256x128x256 box, 4 iterations, 100 steps."""
    return PbfConfig(name="pbf_dam_break_1m", lattice=(128, 64, 128), box=(256, 128, 256),
                     steps=100, **kw)


def cfg2_tail(**kw) -> PbfConfig:
    """4,194,304 particles (256x64x256) in a 512x128x512 box."""
    return PbfConfig(name="pbf_dam_break_4m", lattice=(256, 64, 256), box=(512, 128, 512),
                     steps=100, **kw)


def small(nx=8, ny=8, nz=8, box=(16, 16, 16), **kw) -> PbfConfig:
    return PbfConfig(name=f"pbf_small_{nx}x{ny}x{nz}", lattice=(nx, ny, nz), box=box, **kw)


def dam_break_of_size(n_target: int, seed: int, **kw) -> PbfConfig:
    """A jittered-lattice dam break of about n_target particles: a block twice
    as wide as tall (ny = (n/4)^(1/3)), in a box twice the block extent per axis."""
    ny = max(2, int(round((n_target / 4.0) ** (1.0 / 3.0))))
    nx = max(1, int(round(math.sqrt(n_target / ny))))
    nz = max(1, int(round(n_target / (nx * ny))))
    return PbfConfig(name=f"dam_{nx}x{ny}x{nz}", lattice=(nx, ny, nz), box=(2 * nx, 2 * ny, 2 * nz),
                     seed=seed, **kw)


def dam_break_aspect(n_target: int, aspect: float, seed: int, **kw) -> PbfConfig:
    """A dam break of about n_target particles whose block (and box, twice the
    block per axis) has width / depth = aspect; height (n/4)^(1/3) as in
    dam_break_of_size."""
    ny = max(2, int(round((n_target / 4.0) ** (1.0 / 3.0))))
    area = n_target / ny
    nx = max(1, int(round(math.sqrt(area * aspect))))
    nz = max(1, int(round(area / nx)))
    return PbfConfig(name=f"dam_{nx}x{ny}x{nz}", lattice=(nx, ny, nz), box=(2 * nx, 2 * ny, 2 * nz),
                     seed=seed, **kw)


@dataclass
class MixedConfig:
    """Sand + fluid block over a voxel heightfield of static solid particles.

    The moving block is `fluid`'s jittered lattice; each moving particle is
    sand with probability 1/2 (Philox, seed `phase_seed`, counter (i, 0)).
    Column (x, z) of the field has height hmin + floor((hmax - hmin + 1) u)
    (Philox, seed `field_seed`, counter (x * nz + z, 0)) and one solid particle
    at each voxel centre (x + 0.5, y + 0.5, z + 0.5), y < height."""
    name: str
    fluid: PbfConfig
    field: tuple = (128, 128)
    hmin: int = 1
    hmax: int = 16
    field_seed: int = 11
    phase_seed: int = 12
    mu_s: float = 0.6
    mu_k: float = 0.4
    collision: float = 0.8          # src/Simulate.cpp:159-160
    friction: float = 0.7           # src/Simulate.cpp:161

    def heights(self) -> np.ndarray:
        fx, fz = self.field
        u = uniform01(fx * fz, 1, self.field_seed)[:, 0]
        return (self.hmin + np.floor((self.hmax - self.hmin + 1) * u)).astype(np.int32).reshape(fx, fz)

    def solids(self) -> np.ndarray:
        h = self.heights()
        fx, fz = self.field
        out = []
        for x in range(fx):
            for z in range(fz):
                y = np.arange(h[x, z], dtype=np.float32)
                col = np.empty((y.shape[0], 3), np.float32)
                col[:, 0] = np.float32(x) + np.float32(0.5)
                col[:, 1] = y + np.float32(0.5)
                col[:, 2] = np.float32(z) + np.float32(0.5)
                out.append(col)
        return np.concatenate(out).astype(np.float32)

    def moving_phase(self) -> np.ndarray:
        u = uniform01(self.fluid.n, 1, self.phase_seed)[:, 0]
        return np.where(u < 0.5, 1, 0).astype(np.int32)

    def arrays(self):
        """(positions, phases): the moving block followed by the solids."""
        sol = self.solids()
        pos = np.concatenate([self.fluid.positions(), sol]).astype(np.float32)
        ph = np.concatenate([self.moving_phase(), np.full(sol.shape[0], 2, np.int32)])
        return pos, ph

    @property
    def n_moving(self) -> int:
        return self.fluid.n


def cfg1(**kw) -> MixedConfig:
    """Mixed sand + fluid: 262,144 moving particles (64^3 lattice, block
    origin (32, 17, 32)) over a 128 x 128 heightfield with heights 1..16
    (about 139 k solids) in a 128^3 box, 3 iterations, 200 steps."""
    fl = PbfConfig(name="mixed_sand_fluid_262k", lattice=(64, 64, 64), box=(128, 128, 128),
                   origin=(32.0, 17.0, 32.0), iterations=3, steps=200, seed=1)
    return MixedConfig(name="mixed_sand_fluid_262k", fluid=fl, **kw)


def mixed_small(nx=10, ny=8, nz=10, box=(24, 24, 24), field=(24, 24), hmax=4, seed=2, **kw) -> MixedConfig:
    fl = PbfConfig(name=f"mixed_small_{nx}x{ny}x{nz}", lattice=(nx, ny, nz), box=box,
                   origin=(float((box[0] - nx) // 2), float(hmax + 1), float((box[2] - nz) // 2)),
                   iterations=3, seed=seed)
    return MixedConfig(name=fl.name, fluid=fl, field=field, hmax=hmax, field_seed=seed + 100,
                       phase_seed=seed + 200, **kw)


@dataclass
class BatchConfig:
    """Independent dam-break scenes stepped together."""
    name: str
    scenes: list
    steps: int = 100

    @property
    def n(self) -> int:
        return sum(s.n for s in self.scenes)

    def counts(self) -> np.ndarray:
        return np.array([s.n for s in self.scenes], np.int32)

    def boxes(self) -> np.ndarray:
        return np.array([s.box for s in self.scenes], np.float32)

    def positions(self) -> np.ndarray:
        return np.concatenate([s.positions() for s in self.scenes]).astype(np.float32)


def cfg4(n_scenes: int = 64, lo: int = 4096, hi: int = 200000, seed: int = 4, **kw) -> BatchConfig:
    """64 dam breaks with particle counts log-uniform in [4096, 200000]
    (numpy PCG64 seeded with `seed`), jitter seed = seed * 1000 + scene."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), n_scenes))
    scenes = [dam_break_of_size(int(round(s)), seed * 1000 + k, **kw) for k, s in enumerate(sizes)]
    return BatchConfig(name=f"pbf_batch_{n_scenes}", scenes=scenes)


def training_batch(n_scenes: int = 64, n_per_scene: int = 8192, seed: int = 3, **kw) -> BatchConfig:
    """The warm-start network's training scenes: dam breaks of n_per_scene
    particles with box aspect ratio width / depth uniform in [0.5, 2] (numpy
    PCG64 seeded with `seed`), jitter seed = seed * 1000 + scene, 200 steps."""
    rng = np.random.default_rng(seed)
    aspects = rng.uniform(0.5, 2.0, n_scenes)
    scenes = [dam_break_aspect(n_per_scene, float(a), seed * 1000 + k, **kw) for k, a in enumerate(aspects)]
    return BatchConfig(name=f"pbf_train_{n_scenes}", scenes=scenes, steps=200)
