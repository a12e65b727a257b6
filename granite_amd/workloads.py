"""Synthetic block layouts of the benchmark configurations (host-side geometry only).

uneven_amr_levels(): the batched-AMR configuration — two Brill-Lindquist punctures of mass
0.5 at (+-3, 0, 0), 4 refinement levels with ratio 2 and finest spacing M/32, about 250
blocks whose per-axis edges are drawn uniformly from {16, 24, 32, 48, 64} cells with a fixed
seed. GRANITE's own AMR merges each level into one block (src/amr/amr.cpp:147-300); this
layout is a batching benchmark, so blocks of a level may overlap and are evolved
independently, as the reference evolves each block with block-local ghost fills.
"""
from __future__ import annotations

import numpy as np

EDGES = (16, 24, 32, 48, 64)
PUNCTURES = ([0.5, 0.5], [[-3.0, 0.0, 0.0], [3.0, 0.0, 0.0]])


def uneven_amr_levels(nblocks: int = 250, nlevels: int = 4, finest_dx: float = 1.0 / 32, seed: int = 20240611):
    """Returns [(dx, [(n[3], lo[3], hi[3]), ...]) for each level], coarsest first."""
    rng = np.random.default_rng(seed)
    per = [nblocks // nlevels + (1 if l < nblocks % nlevels else 0) for l in range(nlevels)]
    levels = []
    for l in range(nlevels):
        dx = finest_dx * 2 ** (nlevels - 1 - l)
        blocks = []
        # blocks of every level are placed within +-96 cells of one of the punctures
        reach = 96
        for _ in range(per[l]):
            n = [int(rng.choice(EDGES)) for _ in range(3)]
            centre = PUNCTURES[1][int(rng.integers(0, 2))]
            lo = []
            for d in range(3):
                c = int(round(centre[d] / dx))
                off = int(rng.integers(-reach, reach - n[d] + 1))
                lo.append((c + off) * dx)
            hi = [lo[d] + n[d] * dx for d in range(3)]
            blocks.append((n, lo, hi))
        levels.append((dx, blocks))
    return levels


def cells(levels) -> int:
    return sum(int(np.prod(n)) for _, bl in levels for n, _, _ in bl)
