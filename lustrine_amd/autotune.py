"""Autotuner of the PBF launch configuration, per size bucket.

Candidates: threads per block of the neighbour passes (64, 128, 256), cell
size (h or 1.125 h; any cell size >= h gives the same neighbour sets), and the
neighbour-pass tiling (off, or on with chunks of `zt` cells); then, for the
fastest of those, the sort tile (cells per block of the counting sort's prefix
scan: 1024, 2048 or 4096).  Each candidate is timed on the scene itself: the
scene is re-initialised from the same state for every candidate, stepped
`warmup` steps untimed and `steps` steps timed (hipEvents on the library's
stream), so all candidates see the same simulation window; the fastest wins.
Choices are kept per size bucket (floor(log2 of the particle count)) in memory
and, when a path is given, in a JSON file.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import asdict, dataclass, replace

import numpy as np

from . import gpu as G


@dataclass(frozen=True)
class Launch:
    block: int = 256          # threads per block of the per-slot neighbour passes
    cell_scale: float = 1.0   # cell size / kernel radius
    tiled: bool = False       # LDS-staged tiles instead of one thread per slot
    zt: int = 32              # cells per tile chunk along z (tiled only)
    sort_tile: int = 2048     # cells per block of the counting sort's prefix scan


CANDIDATES = tuple(Launch(b, c, t) for b in (64, 128, 256) for c in (1.0, 1.125) for t in (False, True))
SORT_TILES = (1024, 2048, 4096)


def bucket(n: int) -> int:
    return int(math.floor(math.log2(max(int(n), 1))))


def cell_params(base: "G.LusPbfParams", box, cell_scale: float) -> "G.LusPbfParams":
    """A copy of `base` with cell size cell_scale * h and the grid of `box`
    recomputed (int(D / cs) + 1 per axis, as the reference builds it)."""
    p = G.LusPbfParams.from_buffer_copy(base)
    cs = np.float32(np.float32(base.h) * np.float32(cell_scale))
    p.cell_size = float(cs)
    if box is not None:
        p.gridX, p.gridY, p.gridZ = (int(np.float32(d) / cs) + 1 for d in box)
    return p


def apply(g: "G.Gpu", launch: Launch) -> None:
    g.pbf_set_launch(launch.block)
    g.pbf_set_tiling(launch.tiled, launch.zt)
    g.pbf_set_sort_tile(launch.sort_tile)


class Tuner:
    """Chooses a Launch per size bucket by timing the candidates."""

    def __init__(self, path: str | None = None, candidates=CANDIDATES, sort_tiles=SORT_TILES):
        self.path = path
        self.candidates = tuple(candidates)
        self.sort_tiles = tuple(sort_tiles)
        self.choices: dict[int, Launch] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                for k, v in json.load(f).items():
                    self.choices[int(k)] = Launch(**v)

    def save(self) -> None:
        if self.path:
            with open(self.path, "w") as f:
                json.dump({str(k): asdict(v) for k, v in sorted(self.choices.items())}, f, indent=1)

    def choice(self, n: int) -> Launch | None:
        return self.choices.get(bucket(n))

    def tune(self, g: "G.Gpu", init, n: int, dt, warmup: int = 3, steps: int = 8, force: bool = False):
        """init(launch) (re)initialises the scene for a candidate (cell size);
        returns (chosen Launch, [(launch, ms per step)])."""
        b = bucket(n)
        if not force and b in self.choices:
            return self.choices[b], []
        table = []

        def time(cand):
            init(cand)
            apply(g, cand)
            g.pbf_set_graph(True)
            ms, _, _ = g.pbf_bench(dt, warmup, steps, phases=False)
            table.append((cand, ms / steps))

        for cand in self.candidates:
            time(cand)
        best = min(table, key=lambda t: t[1])[0]
        for st in self.sort_tiles:
            if st != best.sort_tile:
                time(replace(best, sort_tile=st))
        best = min(table, key=lambda t: t[1])[0]
        self.choices[b] = best
        self.save()
        return best, table
