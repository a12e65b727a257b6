"""ctypes bindings for the CPU oracle and the compiled reference (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product (granite_amd/) never does.

Arrays follow GridBlock storage (include/granite/core/grid.hpp:28-42): a float64
array of shape (22, nz+2ng, ny+2ng, nx+2ng), C-contiguous, so that element
(v, k, j, i) sits at v*stride + i + nx_t*(j + ny_t*k).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NV = 22

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference sources exist)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Geom:
    """Block geometry in the reference's terms: n interior cells, ng ghosts, [lo, hi]."""

    def __init__(self, n, ng, lo, hi):
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        # same IEEE operation as GridBlock's constructor (src/core/grid.cpp:22-24)
        self.dx = (self.hi - self.lo) / np.array(self.n, dtype=np.float64)
        self.nt = tuple(v + 2 * self.ng for v in self.n)

    @property
    def shape(self):
        return (NV, self.nt[2], self.nt[1], self.nt[0])

    @property
    def stride(self):
        return self.nt[0] * self.nt[1] * self.nt[2]

    def zeros(self, nv=NV):
        return np.zeros((nv, self.nt[2], self.nt[1], self.nt[0]), dtype=np.float64)

    def interior(self):
        g = self.ng
        return (slice(None), slice(g, g + self.n[2]), slice(g, g + self.n[1]), slice(g, g + self.n[0]))

    def coords(self, d):
        idx = np.arange(self.nt[d], dtype=np.float64)
        return self.lo[d] + ((idx - self.ng) + 0.5) * self.dx[d]

    def cargs(self):
        n = (C.c_int * 3)(*self.n)
        lo = (C.c_double * 3)(*self.lo)
        dx = (C.c_double * 3)(*self.dx)
        return n, C.c_int(self.ng), lo, dx

    def cargs_ref(self):
        n = (C.c_int * 3)(*self.n)
        lo = (C.c_double * 3)(*self.lo)
        hi = (C.c_double * 3)(*self.hi)
        return n, C.c_int(self.ng), lo, hi


DEFAULT_PARAMS = dict(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1, chi_blend_center=0.05, chi_blend_width=0.02)


def param_array(**kw):
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    return (C.c_double * 6)(p["kappa1"], p["kappa2"], p["eta"], p["ko_sigma"], p["chi_blend_center"],
                            p["chi_blend_width"])


class Oracle:
    """The strict-IEEE restatement (oracle/ccz4_oracle.cpp)."""

    def __init__(self, fast: bool = False):
        name = "liboracle_ccz4_fast.so" if fast else "liboracle_ccz4.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_gauss.restype = C.c_double
        L.orc_gauss.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_add_noise.argtypes = [_ip, C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_double, C.c_uint]
        L.orc_num_threads.restype = C.c_int

    def num_threads(self):
        return self.lib.orc_num_threads()

    def rhs(self, g: Geom, u, rhs=None, sources=None, **prm):
        if rhs is None:
            rhs = g.zeros()
        rho = Si = Sij = S = None
        if sources is not None:
            rho, Si, Sij, S = sources
        self.lib.orc_rhs(*g.cargs(), param_array(**prm), _p(u), _p(rhs), _p(rho), _p(Si), _p(Sij), _p(S))
        return rhs

    def ko(self, g: Geom, u, rhs, sigma):
        self.lib.orc_ko(*g.cargs(), C.c_double(sigma), _p(u), _p(rhs))
        return rhs

    def constraints(self, g: Geom, u):
        ham = np.zeros((g.nt[2], g.nt[1], g.nt[0]))
        mom = np.zeros((g.nt[2], g.nt[1], g.nt[0], 3))
        self.lib.orc_constraints(*g.cargs(), _p(u), _p(ham), _p(mom))
        return ham, mom

    def floors(self, g: Geom, u):
        self.lib.orc_floors(*g.cargs(), _p(u))
        return u

    def algebraic(self, g: Geom, u):
        self.lib.orc_algebraic(*g.cargs(), _p(u))
        return u

    def fill_ghosts(self, g: Geom, u, mode: int):
        self.lib.orc_fill_ghosts(*g.cargs(), _p(u), C.c_int(mode))
        return u

    def fill_outflow_faces(self, g: Geom, u, faces: int, nvars=NV):
        self.lib.orc_fill_outflow_faces(*g.cargs(), _p(u), C.c_int(nvars), C.c_int(faces))
        return u

    def combine(self, g: Geom, dst, a1, s1, a2, s2, a3, dt, r):
        self.lib.orc_combine(*g.cargs(), _p(dst), C.c_double(a1), _p(s1), C.c_double(a2), _p(s2), C.c_double(a3),
                             C.c_double(dt), _p(r))
        return dst

    def rk3_step(self, g: Geom, u, dt, bc_mode=0, stage=None, rhs=None, sources=None, **prm):
        stage = g.zeros() if stage is None else stage
        rhs = g.zeros() if rhs is None else rhs
        rho = Si = Sij = S = None
        if sources is not None:
            rho, Si, Sij, S = sources
        self.lib.orc_rk3_step_src(*g.cargs(), param_array(**prm), _p(u), _p(stage), _p(rhs), C.c_double(dt),
                                  C.c_int(bc_mode), _p(rho), _p(Si), _p(Sij), _p(S))
        return u

    def init_flat(self, g: Geom):
        u = g.zeros()
        self.lib.orc_init_flat(*g.cargs(), _p(u))
        return u

    def init_bl(self, g: Geom, masses, positions):
        u = g.zeros()
        m = np.ascontiguousarray(masses, dtype=np.float64)
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        self.lib.orc_init_bl(*g.cargs(), C.c_int(len(m)), _p(m), _p(p), _p(u))
        return u

    def init_gauge_wave(self, g: Geom, A, L):
        u = g.zeros()
        self.lib.orc_init_gauge_wave(*g.cargs(), C.c_double(A), C.c_double(L), _p(u))
        return u

    def gauss(self, seed, q):
        return self.lib.orc_gauss(C.c_uint64(seed), C.c_uint64(q))

    def add_noise(self, g: Geom, u, seed, amp, fields):
        mask = 0
        for f in fields:
            mask |= 1 << int(f)
        self.lib.orc_add_noise(*g.cargs(), _p(u), C.c_uint64(seed), C.c_double(amp), C.c_uint(mask))
        return u

    def prolongate(self, gc: Geom, uc, gf: Geom, uf, nvars=NV):
        self.lib.orc_prolongate(*gc.cargs(), _p(uc), *gf.cargs(), _p(uf), C.c_int(nvars))
        return uf

    def restrict(self, gf: Geom, uf, gc: Geom, uc, nvars=NV):
        self.lib.orc_restrict(*gf.cargs(), _p(uf), *gc.cargs(), _p(uc), C.c_int(nvars))
        return uc


class Reference:
    """The compiled GRANITE sources (oracle/_ref), available where /root/reference was present."""

    def __init__(self, fast: bool = False):
        name = "libgranite_ref_fast.so" if fast else "libgranite_ref.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.lib.ref_bl_psi.restype = C.c_double
        self.lib.ref_bl_psi.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double]

    @staticmethod
    def available(fast: bool = False) -> bool:
        name = "libgranite_ref_fast.so" if fast else "libgranite_ref.so"
        return os.path.exists(os.path.join(HERE, "_ref", name))

    def rhs(self, g: Geom, u, rhs=None, sources=None, **prm):
        if rhs is None:
            rhs = g.zeros()
        rho = Si = Sij = S = None
        if sources is not None:
            rho, Si, Sij, S = sources
        self.lib.ref_rhs(*g.cargs_ref(), param_array(**prm), _p(u), _p(rhs), _p(rho), _p(Si), _p(Sij), _p(S))
        return rhs

    def ko(self, g: Geom, u, rhs, sigma):
        self.lib.ref_ko(*g.cargs_ref(), param_array(ko_sigma=sigma), _p(u), _p(rhs))
        return rhs

    def constraints(self, g: Geom, u):
        ham = np.zeros((g.nt[2], g.nt[1], g.nt[0]))
        mom = np.zeros((g.nt[2], g.nt[1], g.nt[0], 3))
        self.lib.ref_constraints(*g.cargs_ref(), _p(u), _p(ham), _p(mom))
        return ham, mom

    def init_bl(self, g: Geom, masses, positions):
        u = g.zeros()
        m = np.ascontiguousarray(masses, dtype=np.float64)
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        self.lib.ref_init_bl(*g.cargs_ref(), C.c_int(len(m)), _p(m), _p(p), _p(u))
        return u

    def bl_psi(self, masses, positions, x, y, z):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        return self.lib.ref_bl_psi(len(m), _p(m), _p(p), x, y, z)

    def init_gauge_wave(self, g: Geom, A, L):
        u = g.zeros()
        self.lib.ref_init_gauge_wave(*g.cargs_ref(), C.c_double(A), C.c_double(L), _p(u))
        return u

    def prolongate(self, gc: Geom, uc, gf: Geom, uf, nvars=NV):
        self.lib.ref_prolongate(*gc.cargs_ref(), _p(uc), *gf.cargs_ref(), _p(uf), C.c_int(nvars))
        return uf

    def restrict(self, gf: Geom, uf, gc: Geom, uc, nvars=NV):
        self.lib.ref_restrict(*gf.cargs_ref(), _p(uf), *gc.cargs_ref(), _p(uc), C.c_int(nvars))
        return uc


def normwise_rel(a, b, floor=1e-300):
    """Per-field max|a-b| / max(max|b|, floor) over the given arrays (field axis 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for v in range(a.shape[0]):
        d = np.max(np.abs(a[v] - b[v])) if a[v].size else 0.0
        s = max(np.max(np.abs(b[v])) if b[v].size else 0.0, floor)
        out.append(d / s)
    return np.array(out)


# tensor groups of the 22 fields: chi | metric | A~ | K | Gamma-hat | Theta | lapse | shift
FIELD_GROUPS = [[0], list(range(1, 7)), list(range(7, 13)), [13], [14, 15, 16], [17], [18], [19, 20, 21]]


def group_rel(a, b):
    """Per-field max|a-b| divided by the largest |b| over the field's tensor group.

    A component that vanishes by symmetry (gt_xy of a plane wave) carries only round-off in
    any implementation; measuring it against its own size would compare two round-off
    patterns. It is measured against the size of the tensor it belongs to instead."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.empty(b.shape[0])
    for grp in FIELD_GROUPS:
        s = max(float(np.max(np.abs(b[grp]))), 1e-300)
        scale[grp] = s
    return np.array([np.max(np.abs(a[v] - b[v])) / scale[v] for v in range(a.shape[0])])


def ulp_spread(fn, u, trials=2, seed=0, norm=group_rel):
    """Per-field spread of the oracle map `fn` (input state -> output state) under 1-ulp
    relative perturbations of its input: the largest norm(fn(u * (1 +- 2^-52)), fn(u)) over
    `trials` random sign patterns. Non-finite outputs are left out of the comparison.

    An RK3 step feeds its stage states back through second-derivative stencils, so two
    correct implementations whose stage states differ in the last bit (any change of
    rounding order does that) differ at the end by about this spread. Where it exceeds
    the FP64 parity bar, the bar cannot be met by any implementation that is not bitwise
    equal to the oracle, and the spread is the tolerance instead."""
    rng = np.random.default_rng(seed)
    ref = fn(u.copy())
    bad = ~np.isfinite(ref)
    out = np.zeros(ref.shape[0])
    for _ in range(trials):
        up = u * (1.0 + rng.choice([-1.0, 1.0], u.shape) * 2.0**-52)
        got = fn(up)
        out = np.maximum(out, norm(np.where(bad | ~np.isfinite(got), 0.0, got), np.where(bad, 0.0, ref)))
    return out
