"""Level geometry of a Berger-Oliger hierarchy built from tracking spheres (host side).

tracking_levels() returns the blocks that amr::AMRHierarchy::initialize() builds when only
tracking spheres ask for refinement (src/amr/amr.cpp:34-78): a level-0 block over the
domain, then regrid(l) for each level a sphere's min_level asks for
(src/amr/amr.cpp:80-301). Per level and sphere, a candidate box of ncells fine cells is
centred on the sphere, its lower corner snapped down onto the parent grid and the box
clamped to the domain; overlapping (or touching, within one fine cell) candidates are merged
into their union until none overlap; each box becomes one block whose cell count is the
rounded extent over the fine spacing. The arithmetic is the reference's, operation for
operation, so lo, hi and dx = (hi - lo) / n come out bit-identical.

binary_puncture_levels(): the 4-level binary-puncture hierarchy benchmark: 96^3 on
[-64, 64]^3 (dx 4/3), Brill-Lindquist punctures of mass 0.5 at (+-3, 0, 0) with the
tracking spheres src/main.cpp:1119-1126 registers for them (radius max(2m, 0.5), min_level
min(max_levels - 1, 3)), max_levels 4: levels at dx 4/3, 2/3, 1/3, 1/6 of
96x96x96, 106x96x96, 114x96x96 and 132x96x96 cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class Sphere:
    center: tuple
    radius: float
    min_level: int


def _cround(x: float) -> int:
    """std::round: halves away from zero."""
    a = abs(x)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1
    return int(math.copysign(r, x))


def _block(n, lo, hi):
    return ([int(v) for v in n], [float(v) for v in lo], [float(v) for v in hi])


def _dx(block):
    n, lo, hi = block
    return [(hi[d] - lo[d]) / float(n[d]) for d in range(3)]


def _regrid(levels, level, ncells, dlo, dhi, spheres, max_levels, ratio):
    """AMRHierarchy::regrid(level) with a tagger that tags nothing (src/amr/amr.cpp:80-301)."""
    if level < 0 or level >= len(levels) or max_levels <= 1:
        return
    nxt = level + 1
    if nxt >= max_levels or not levels[level]:
        return
    parent_dx = _dx(levels[level][0])
    child_dx = [parent_dx[d] / float(ratio) for d in range(3)]
    half = [0.5 * float(ncells[d]) * child_dx[d] for d in range(3)]
    if not any(nxt <= s.min_level for s in spheres):
        return
    if nxt >= len(levels):
        levels.append([])
    cands = []
    for s in spheres:
        if nxt > s.min_level:
            continue
        lo, hi = [0.0] * 3, [0.0] * 3
        for d in range(3):
            origin = dlo[d]
            ideal_lo = s.center[d] - half[d]
            snapped = math.floor((ideal_lo - origin) / parent_dx[d]) * parent_dx[d] + origin
            lo[d] = snapped
            hi[d] = snapped + float(ncells[d]) * child_dx[d]
            if lo[d] < dlo[d]:
                lo[d] = dlo[d]
                hi[d] = lo[d] + float(ncells[d]) * child_dx[d]
            if hi[d] > dhi[d]:
                hi[d] = dhi[d]
                lo[d] = hi[d] - float(ncells[d]) * child_dx[d]
                if lo[d] < dlo[d]:
                    lo[d] = dlo[d]
        cands.append([lo, hi])
    if not cands:
        return
    tol = child_dx[0]
    merged = True
    while merged:
        merged = False
        a = 0
        while a < len(cands) and not merged:
            for b in range(a + 1, len(cands)):
                ov = all(not (cands[a][0][d] > cands[b][1][d] + tol or cands[b][0][d] > cands[a][1][d] + tol)
                         for d in range(3))
                if not ov:
                    continue
                for d in range(3):
                    cands[a][0][d] = min(cands[a][0][d], cands[b][0][d])
                    cands[a][1][d] = max(cands[a][1][d], cands[b][1][d])
                    cands[a][0][d] = max(cands[a][0][d], dlo[d])
                    cands[a][1][d] = min(cands[a][1][d], dhi[d])
                cands[b] = cands[-1]
                cands.pop()
                merged = True
                break
            a += 1
    for lo, hi in cands:
        n = [max(_cround((hi[d] - lo[d]) / child_dx[d]), 1) for d in range(3)]
        dup = False
        for blk in levels[nxt]:
            bdx = _dx(blk)
            same = True
            for d in range(3):
                exp_first = lo[d] + 0.5 * child_dx[d]
                got_first = blk[1][d] + (0.0 + 0.5) * bdx[d]
                if abs(bdx[d] - child_dx[d]) > 1e-10 * child_dx[d] or abs(got_first - exp_first) > 0.1 * child_dx[d]:
                    same = False
            if same:
                dup = True
                break
        if not dup:
            levels[nxt].append(_block(n, lo, hi))


def tracking_levels(ncells, domain_lo, domain_hi, spheres, max_levels, ratio=2):
    """[[(n[3], lo[3], hi[3]), ...] per level], coarsest first, as initialize() builds them."""
    levels = [[_block(ncells, domain_lo, domain_hi)]]
    deepest = min(max((s.min_level for s in spheres), default=0), max_levels - 1)
    for l in range(deepest):
        _regrid(levels, l, ncells, domain_lo, domain_hi, spheres, max_levels, ratio)
        if len(levels) <= l + 1:
            break
    return levels


PUNCTURES = ([0.5, 0.5], [[-3.0, 0.0, 0.0], [3.0, 0.0, 0.0]])


def puncture_spheres(masses, positions, max_levels):
    """Tracking spheres of src/main.cpp:1119-1126: radius max(2m, 0.5), min_level min(max_levels-1, 3)."""
    return [Sphere(tuple(float(c) for c in p), max(2.0 * m, 0.5), min(max_levels - 1, 3))
            for m, p in zip(masses, positions)]


def binary_puncture_levels(ncells=96, extent=64.0, max_levels=4):
    spheres = puncture_spheres(*PUNCTURES, max_levels)
    return tracking_levels([ncells] * 3, [-extent] * 3, [extent] * 3, spheres, max_levels)


def level_cells(levels):
    return [sum(n[0] * n[1] * n[2] for n, _, _ in lv) for lv in levels]


def cell_steps_per_coarse_step(levels, ratio=2):
    """Interior cell updates of one subcycled coarse step: level l steps ratio^l times."""
    return sum(c * ratio ** l for l, c in enumerate(level_cells(levels)))
