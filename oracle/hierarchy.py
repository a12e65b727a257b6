"""CPU restatement of the reference's Berger-Oliger subcycling (test infrastructure only).

A hierarchy is a list of levels, coarsest first; a level is a list of (Geom, u) blocks in
the reference's block order. Every per-block operation is a call into the strict-IEEE oracle
(oracle/ccz4_oracle.cpp); this module only restates the order of operations of
amr::AMRHierarchy (src/amr/amr.cpp:303-350 subcycle, :389-398 setLevelDt, :400-482
fillGhostZones) and the constraint telemetry of src/main.cpp:1216-1250.

Pinned by: tests/test_oracle_hierarchy.py runs the compiled reference's own
AMRHierarchy::subcycle (oracle/_ref, regrid_interval 0) with the oracle's SSP-RK3 step as
its per-level callback and requires bit-identical states on every level after two coarse
steps; the geometry of the levels is pinned against the reference's regrid the same way.

step(..., skip_overwritten=True) leaves out the work whose results the reference overwrites
before anything reads them, and gives bit-identical states (also tested): on a level >= 1,
fillGhostZones prolongates the parent into every cell of every block, so the SSP-RK3 step
that precedes it, and everything done on that level and below during all but the last of
its parent's substeps, is overwritten. This is what makes long runs of the oracle cheap.
"""
from __future__ import annotations

import numpy as np

from .oracle import Geom, NV


def uncovered_faces(fine: Geom, parents) -> int:
    """Faces of `fine` that fillGhostZones leaves to the outflow copy (src/amr/amr.cpp:412-428).

    Bit 2d: low side of axis d, bit 2d+1: high side."""
    lo = [False] * 3
    hi = [False] * 3
    for c in parents:
        for d in range(3):
            fine_lo = fine.lo[d] + (0.0 + 0.5) * fine.dx[d]
            fine_hi = fine.lo[d] + (float(fine.n[d] - 1) + 0.5) * fine.dx[d]
            c_lo = c.lo[d] + (0.0 + 0.5) * c.dx[d]
            c_hi = c.lo[d] + (float(c.n[d] - 1) + 0.5) * c.dx[d]
            if c_lo <= fine_lo + 0.5 * fine.dx[d]:
                lo[d] = True
            if c_hi >= fine_hi - 0.5 * fine.dx[d]:
                hi[d] = True
    m = 0
    for d in range(3):
        m |= (0 if lo[d] else 1 << (2 * d)) | (0 if hi[d] else 1 << (2 * d + 1))
    return m


class OracleHierarchy:
    def __init__(self, oracle, levels, ratio: int = 2):
        """levels: [[(Geom, u), ...] per level]; the arrays are updated in place."""
        self.o = oracle
        self.levels = levels
        self.ratio = ratio
        self._scratch = {}

    def _tmp(self, g: Geom):
        key = (g.n, g.ng)
        if key not in self._scratch:
            self._scratch[key] = (g.zeros(), g.zeros())
        return self._scratch[key]

    def prolongate(self, l: int):
        """for fine in level l: for coarse in level l-1: prolongate (src/amr/amr.cpp:313-318)."""
        for gf, uf in self.levels[l]:
            for gc, uc in self.levels[l - 1]:
                self.o.prolongate(gc, uc, gf, uf)

    def restrict(self, l: int):
        """for fine in level l: for coarse in level l-1: restrict_data (src/amr/amr.cpp:339-341)."""
        for gf, uf in self.levels[l]:
            for gc, uc in self.levels[l - 1]:
                self.o.restrict(gf, uf, gc, uc)

    def fill_ghosts(self, l: int):
        """fillGhostZones(l) (src/amr/amr.cpp:400-482)."""
        parents = [g for g, _ in self.levels[l - 1]] if l > 0 else []
        for gf, uf in self.levels[l]:
            if l > 0:
                for gc, uc in self.levels[l - 1]:
                    self.o.prolongate(gc, uc, gf, uf)
            faces = uncovered_faces(gf, parents) if l > 0 else 0x3F
            if faces:
                self.o.fill_outflow_faces(gf, uf, faces)

    def evolve(self, l: int, dt: float, prm):
        """The per-level callback: SSP-RK3 with 1/r ghosts after every stage, block by block."""
        for g, u in self.levels[l]:
            stage, rhs = self._tmp(g)
            self.o.rk3_step(g, u, dt, 0, stage, rhs, **prm)

    def subcycle(self, l: int, dt: float, prm, skip_overwritten=False, last=True):
        """subcycle(l) (src/amr/amr.cpp:303-350) without its regrid step."""
        if l >= len(self.levels):
            return
        dead = skip_overwritten and l > 0
        if dead and not last:
            return  # overwritten by the parent's next substep
        if l > 0 and not dead:
            self.prolongate(l)
        self.fill_ghosts(l)
        if not dead:
            self.evolve(l, dt, prm)
            self.fill_ghosts(l)
        if l + 1 < len(self.levels):
            dt_fine = dt / float(self.ratio)
            for r in range(self.ratio):
                self.subcycle(l + 1, dt_fine, prm, skip_overwritten, last=(r == self.ratio - 1))
            self.restrict(l + 1)

    def step(self, dt0: float, prm, skip_overwritten=False):
        """One coarse step: setLevelDt(0, dt0); subcycle(0) (src/main.cpp:1184-1187)."""
        self.subcycle(0, dt0, prm, skip_overwritten)

    def constraints_l2(self):
        """(sum H^2, sum |M|^2, count) over the interiors of all blocks (src/main.cpp:1216-1250)."""
        sh = sm = 0.0
        n = 0
        for lv in self.levels:
            for g, u in lv:
                ham, mom = self.o.constraints(g, u)
                s = g.ng
                hi = ham[s:s + g.n[2], s:s + g.n[1], s:s + g.n[0]]
                mi = mom[s:s + g.n[2], s:s + g.n[1], s:s + g.n[0], :]
                sh += float(np.sum(hi * hi))
                sm += float(np.sum(mi * mi))
                n += g.n[0] * g.n[1] * g.n[2]
        return sh, sm, n


def build_levels(oracle, level_geoms, ng=4, masses=None, positions=None):
    """Blocks for level_geoms ([[(n, lo, hi), ...] per level]); level 0 (and, when given,
    every level) gets Brill-Lindquist data, finer levels are prolongated from their parent
    as regrid does (src/amr/amr.cpp:290-293)."""
    levels = []
    for l, lv in enumerate(level_geoms):
        blocks = []
        for n, lo, hi in lv:
            g = Geom(n, ng, lo, hi)
            u = oracle.init_bl(g, masses, positions) if l == 0 else g.zeros()
            blocks.append((g, u))
        levels.append(blocks)
        if l > 0:
            for gf, uf in blocks:
                for gc, uc in levels[l - 1]:
                    oracle.prolongate(gc, uc, gf, uf)
    return levels


class ReferenceHierarchy:
    """The compiled reference's own AMRHierarchy (oracle/_ref), levels from tracking spheres."""

    def __init__(self, ref, ncells, domain_lo, domain_hi, spheres, max_levels, ng=4):
        import ctypes as C

        L = ref.lib
        self.C = C
        self.lib = L
        L.ref_hier_build.restype = C.c_void_p
        L.ref_hier_build.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                     C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int)]
        L.ref_hier_free.argtypes = [C.c_void_p]
        L.ref_hier_nblocks.argtypes = [C.c_void_p]
        L.ref_hier_block.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ref_hier_data.restype = C.POINTER(C.c_double)
        L.ref_hier_data.argtypes = [C.c_void_p, C.c_int]
        L.ref_hier_fill_ghosts.argtypes = [C.c_void_p, C.c_int]
        self.fn_t = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double)
        L.ref_hier_step.argtypes = [C.c_void_p, C.c_double, self.fn_t]
        ns = len(spheres)
        cen = (C.c_double * (3 * ns))(*[c for s in spheres for c in s.center])
        rad = (C.c_double * ns)(*[s.radius for s in spheres])
        mlv = (C.c_int * ns)(*[s.min_level for s in spheres])
        self.h = L.ref_hier_build((C.c_int * 3)(*ncells), (C.c_double * 3)(*domain_lo), (C.c_double * 3)(*domain_hi),
                                  ng, max_levels, ns, cen, rad, mlv)
        self.ng = ng

    def close(self):
        if self.h:
            self.lib.ref_hier_free(self.h)
            self.h = None

    def blocks(self):
        """[(level, n, lo, hi, dx)] in getAllBlocks() order."""
        C = self.C
        out = []
        for i in range(self.lib.ref_hier_nblocks(self.h)):
            lv = C.c_int()
            n = (C.c_int * 3)()
            lo, hi, dx = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
            self.lib.ref_hier_block(self.h, i, C.byref(lv), n, lo, hi, dx)
            out.append((lv.value, list(n), list(lo), list(hi), list(dx)))
        return out

    def view(self, i):
        """numpy view (22, nz, ny, nx) of block i's storage."""
        lv, n, lo, hi, dx = self.blocks()[i]
        nt = [v + 2 * self.ng for v in n]
        ptr = self.lib.ref_hier_data(self.h, i)
        return np.ctypeslib.as_array(ptr, shape=(NV, nt[2], nt[1], nt[0]))

    def fill_ghosts(self, level):
        self.lib.ref_hier_fill_ghosts(self.h, level)

    def step(self, dt0, evolve):
        """evolve(level, n, ng, lo, dx, u_view, dt) is called for every block the reference evolves."""
        C = self.C

        def cb(level, n, ng, lo, dx, u, dt):
            nn = [n[0], n[1], n[2]]
            nt = [v + 2 * ng for v in nn]
            arr = np.ctypeslib.as_array(u, shape=(NV, nt[2], nt[1], nt[0]))
            evolve(level, nn, ng, [lo[0], lo[1], lo[2]], [dx[0], dx[1], dx[2]], arr, dt)

        f = self.fn_t(cb)
        self.lib.ref_hier_step(self.h, C.c_double(dt0), f)
