"""Host mirror of GRANITE's CCZ4 interface over the HIP engine's C ABI.

GridBlock keeps the reference storage (include/granite/core/grid.hpp:28-42,157-161):
one float64 buffer of num_vars x stride, element (v, i, j, k) at v*stride + i + nx*(j + ny*k),
exposed as a numpy array of shape (num_vars, nz, ny, nx). CCZ4Evolution keeps the
reference method names, arguments and effects (include/granite/spacetime/ccz4.hpp:49-111);
C ABI failures raise EngineError (a RuntimeError), as the reference's C++ wrapper
would turn them into std::runtime_error.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass

import numpy as np

NUM_SPACETIME_VARS = 22
BC_NONE, BC_SOMMERFELD, BC_PERIODIC, BC_OUTFLOW = -1, 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
# tuned choices (enum gr_tune_kind)
TUNE_STAGE, TUNE_GHOST, TUNE_TRANSFER = 0, 1, 2
GROUP_CHOICES = (1, 2, 4, 11)
STAGE_VARIANTS = 3  # FP64 staging: 0 registers, 1 direct LDS (8x4x4 parts), 2 direct LDS (8x4x8 parts)

_HERE = os.path.dirname(os.path.abspath(__file__))


class SpacetimeVar(enum.IntEnum):
    """Field order of include/granite/core/types.hpp:104-127."""

    CHI = 0
    GAMMA_XX = 1
    GAMMA_XY = 2
    GAMMA_XZ = 3
    GAMMA_YY = 4
    GAMMA_YZ = 5
    GAMMA_ZZ = 6
    A_XX = 7
    A_XY = 8
    A_XZ = 9
    A_YY = 10
    A_YZ = 11
    A_ZZ = 12
    K_TRACE = 13
    GAMMA_HAT_X = 14
    GAMMA_HAT_Y = 15
    GAMMA_HAT_Z = 16
    THETA = 17
    LAPSE = 18
    SHIFT_X = 19
    SHIFT_Y = 20
    SHIFT_Z = 21


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"granite engine error {code}: {msg}")
        self.code = code


class _Geom(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("ng", C.c_int), ("lo", C.c_double * 3), ("dx", C.c_double * 3)]


class _Params(C.Structure):
    _fields_ = [
        ("kappa1", C.c_double),
        ("kappa2", C.c_double),
        ("eta", C.c_double),
        ("ko_sigma", C.c_double),
        ("chi_blend_center", C.c_double),
        ("chi_blend_width", C.c_double),
    ]


_dp = C.POINTER(C.c_double)
_lib = None


def library_path() -> str:
    """In-tree library; GRANITE_CCZ4_LIB may name another build of the same C ABI."""
    return os.environ.get("GRANITE_CCZ4_LIB") or os.path.join(_HERE, "libgranite_ccz4.so")


def library():
    """The loaded C ABI library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise EngineError(-4, f"{path} not built (run granite_amd.build.build())")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "gr_version": (C.c_char_p, []),
        "gr_stage_tile": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
        "gr_engine_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "gr_engine_destroy": (C.c_int, [vp]),
        "gr_last_error": (C.c_char_p, [vp]),
        "gr_engine_sync": (C.c_int, [vp]),
        "gr_ccz4_rhs": (C.c_int, [vp, C.POINTER(_Geom), C.POINTER(_Params), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "gr_ccz4_dissipation": (C.c_int, [vp, C.POINTER(_Geom), C.POINTER(_Params), C.c_int, _dp, _dp]),
        "gr_ccz4_constraints": (C.c_int, [vp, C.POINTER(_Geom), _dp, _dp, _dp]),
        "gr_ccz4_rhs_device": (C.c_int, [vp, C.POINTER(_Geom), C.POINTER(_Params), vp, vp, vp, vp, vp, vp]),
        "gr_level_create": (C.c_int, [vp, C.c_int, C.POINTER(_Geom), C.POINTER(vp)]),
        "gr_level_destroy": (C.c_int, [vp]),
        "gr_level_upload": (C.c_int, [vp, C.c_int, _dp]),
        "gr_level_download": (C.c_int, [vp, C.c_int, _dp]),
        "gr_level_state": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "gr_level_init_brill_lindquist": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "gr_level_add_noise": (C.c_int, [vp, C.c_uint64, C.c_double, C.c_uint32]),
        "gr_level_fill_ghosts": (C.c_int, [vp, C.c_int]),
        "gr_level_rk3_step": (C.c_int, [vp, C.POINTER(_Params), C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]),
        "gr_level_sync": (C.c_int, [vp]),
        "gr_level_constraints_l2": (C.c_int, [vp, _dp, C.POINTER(C.c_longlong)]),
        "gr_ccz4_constraints_l2_noalloc": (C.c_int, [vp, C.POINTER(_Geom), vp, vp, C.POINTER(C.c_longlong)]),
        "gr_level_nonfinite": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "gr_level_read_cell": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
        "gr_level_profile": (C.c_int, [vp, C.c_int]),
        "gr_level_set_mixed_policy": (C.c_int, [vp, C.c_double]),
        "gr_level_rhs": (C.c_int, [vp, C.c_int, C.POINTER(_Params), C.c_int, _dp]),
        "gr_level_profile_read": (C.c_int, [vp, C.c_int, _dp, C.POINTER(C.c_int)]),
        "gr_level_init_gauge_wave": (C.c_int, [vp, C.c_double, C.c_double]),
        "gr_level_set_sources": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
        "gr_level_rk3_stage": (C.c_int, [vp, C.POINTER(_Params), C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]),
        "gr_level_first_nonfinite": (C.c_int, [vp] + [C.POINTER(C.c_int)] * 6 + [_dp]),
        "gr_level_cfl": (C.c_int, [vp, C.c_double, _dp]),
        "gr_hierarchy_set_cfl_monitor": (C.c_int, [vp, C.c_int]),
        "gr_hierarchy_cfl": (C.c_int, [vp, _dp, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
        "gr_hierarchy_create": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(vp)]),
        "gr_hierarchy_destroy": (C.c_int, [vp]),
        "gr_hierarchy_prolongate": (C.c_int, [vp, C.c_int]),
        "gr_hierarchy_restrict": (C.c_int, [vp, C.c_int]),
        "gr_hierarchy_fill_ghosts": (C.c_int, [vp, C.c_int]),
        "gr_hierarchy_step": (C.c_int, [vp, C.POINTER(_Params), C.c_double, C.c_int, C.c_int]),
        "gr_hierarchy_constraints_l2": (C.c_int, [vp, _dp, _dp, C.POINTER(C.c_longlong)]),
        "gr_checkpoint_create": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(vp)]),
        "gr_checkpoint_destroy": (C.c_int, [vp]),
        "gr_checkpoint_begin": (C.c_int, [vp]),
        "gr_checkpoint_query": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "gr_checkpoint_wait": (C.c_int, [vp]),
        "gr_checkpoint_nblocks": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "gr_checkpoint_block": (C.c_int, [vp, C.c_int, C.POINTER(_dp), C.POINTER(C.c_long)]),
        "gr_level_autotune": (C.c_int, [vp, C.POINTER(_Params), C.c_double, _dp, C.c_int, C.POINTER(C.c_int)]),
        "gr_hierarchy_autotune": (C.c_int, [vp, C.POINTER(_Params), C.c_double, _dp, C.c_int, C.POINTER(C.c_int)]),
        "gr_level_get_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "gr_level_set_tuning": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "gr_hierarchy_get_tuning": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
        "gr_hierarchy_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
        "gr_tune_save": (C.c_int, [C.c_char_p]),
        "gr_tune_load": (C.c_int, [C.c_char_p]),
        "gr_tune_clear": (C.c_int, []),
        "gr_levels_rk3_step": (C.c_int, [C.c_int, C.POINTER(vp), _dp, C.POINTER(_Params), C.c_int, C.c_int, C.c_int,
                                         C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_dp)


def levels_rk3_step(levels, dts, params: "CCZ4Params", precision=0, bc_mode=0, pre_fill=-1, use_graph=False):
    """gr_levels_rk3_step: one RK3 step of independent levels (level i with dts[i]), each
    level's ghost fills overlapping the next levels' stage kernels."""
    levels = list(levels)
    if not levels:
        return
    eng = levels[0]._e
    arr = (C.c_void_p * len(levels))(*[lv._h for lv in levels])
    d = (C.c_double * len(levels))(*[float(x) for x in dts])
    eng.check(eng.lib.gr_levels_rk3_step(len(levels), arr, d, C.byref(params.c()), int(precision), int(bc_mode),
                                         int(pre_fill), 1 if use_graph else 0))


def tune_save(path):
    """Write the process-wide tuning table (gr_tune_save)."""
    if library().gr_tune_save(os.fsencode(path)) != 0:
        raise EngineError(-1, f"cannot write tuning table {path}")


def tune_load(path):
    """Merge a tuning table file into the process-wide table (gr_tune_load); levels created
    afterwards use its entries."""
    if library().gr_tune_load(os.fsencode(path)) != 0:
        raise EngineError(-1, f"cannot read tuning table {path} (missing, or a malformed / unknown entry)")


def tune_clear():
    library().gr_tune_clear()


# ----------------------------------------------------------------------------------------
# GridBlock
# ----------------------------------------------------------------------------------------
class GridBlock:
    """Uniform block with ghost zones and field-major storage (reference GridBlock)."""

    def __init__(self, id, level, ncells, lo, hi, nghost, num_vars=NUM_SPACETIME_VARS):
        self._id = int(id)
        self._level = int(level)
        self._n = tuple(int(v) for v in ncells)
        self._lo = np.array(lo, dtype=np.float64)
        self._hi = np.array(hi, dtype=np.float64)
        self._ng = int(nghost)
        self._nv = int(num_vars)
        # dx = (hi - lo) / n, src/core/grid.cpp:22-24
        self._dx = (self._hi - self._lo) / np.array(self._n, dtype=np.float64)
        nt = [v + 2 * self._ng for v in self._n]
        self.array = np.zeros((self._nv, nt[2], nt[1], nt[0]), dtype=np.float64)

    # accessors of include/granite/core/grid.hpp
    def getId(self):
        return self._id

    def getLevel(self):
        return self._level

    def getNumGhost(self):
        return self._ng

    def getNumVars(self):
        return self._nv

    def totalCells(self, d):
        return self._n[d] + 2 * self._ng

    def dx(self, d):
        return float(self._dx[d])

    def x(self, d, idx):
        return float(self._lo[d] + ((float(idx) - self._ng) + 0.5) * self._dx[d])

    def istart(self, d=0):
        return self._ng

    def iend(self, d):
        return self._ng + self._n[d]

    def totalSize(self):
        return self.totalCells(0) * self.totalCells(1) * self.totalCells(2)

    def numCells(self):
        return self._n

    def lowerCorner(self):
        return tuple(self._lo)

    def upperCorner(self):
        return tuple(self._hi)

    def cellSpacing(self):
        return tuple(self._dx)

    def data(self, var, i, j, k):
        return float(self.array[var, k, j, i])

    def set(self, var, i, j, k, value):
        self.array[var, k, j, i] = value

    def rawData(self, var):
        return self.array[var]

    def flatBuffer(self):
        return self.array

    def interior(self):
        g = self._ng
        n = self._n
        return (slice(None), slice(g, g + n[2]), slice(g, g + n[1]), slice(g, g + n[0]))

    def geom(self) -> _Geom:
        g = _Geom()
        for d in range(3):
            g.n[d] = self._n[d]
            g.lo[d] = self._lo[d]
            g.dx[d] = self._dx[d]
        g.ng = self._ng
        return g


@dataclass
class CCZ4Params:
    """include/granite/spacetime/ccz4.hpp:25-34 (fd_order and use_moving_puncture are not
    read by the reference RHS)."""

    kappa1: float = 0.02
    kappa2: float = 0.0
    eta: float = 2.0
    ko_sigma: float = 0.1
    fd_order: int = 4
    use_moving_puncture: bool = True
    chi_blend_center: float = 0.05
    chi_blend_width: float = 0.02

    def c(self) -> _Params:
        return _Params(self.kappa1, self.kappa2, self.eta, self.ko_sigma, self.chi_blend_center, self.chi_blend_width)


# ----------------------------------------------------------------------------------------
# engine
# ----------------------------------------------------------------------------------------
class Engine:
    """One HIP device + stream (gr_engine)."""

    _default = None

    def __init__(self, device: int = 0):
        L = library()
        h = C.c_void_p()
        rc = L.gr_engine_create(int(device), C.byref(h))
        if rc != 0:
            raise EngineError(rc, f"no usable gfx950 device {device}")
        self._h = h
        self.lib = L

    @classmethod
    def default(cls) -> "Engine":
        if cls._default is None:
            cls._default = Engine(0)
        return cls._default

    def check(self, rc):
        if rc != 0:
            raise EngineError(rc, self.lib.gr_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def sync(self):
        self.check(self.lib.gr_engine_sync(self._h))

    def close(self):
        if self._h:
            self.lib.gr_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CCZ4Evolution:
    """Drop-in for spacetime::CCZ4Evolution; every method runs on the MI355X engine."""

    def __init__(self, params: CCZ4Params, engine: Engine | None = None):
        self._p = params
        self._e = engine if engine is not None else Engine.default()

    def params(self) -> CCZ4Params:
        return self._p

    def computeRHS(self, grid: GridBlock, rhs: GridBlock, rho_matter, Si_matter, Sij_matter, S_trace):
        """src/spacetime/ccz4.cpp:251-865; sources are arrays of totalSize (Si: (N,3), Sij: (N,6))."""
        N = grid.totalSize()
        rho = np.ascontiguousarray(rho_matter, dtype=np.float64).reshape(N)
        Si = np.ascontiguousarray(Si_matter, dtype=np.float64).reshape(3 * N)
        Sij = np.ascontiguousarray(Sij_matter, dtype=np.float64).reshape(6 * N)
        S = np.ascontiguousarray(S_trace, dtype=np.float64).reshape(N)
        self._rhs(grid, rhs, rho, Si, Sij, S)

    def computeRHSVacuum(self, grid: GridBlock, rhs: GridBlock):
        """src/spacetime/ccz4.cpp:229-245"""
        self._rhs(grid, rhs, None, None, None, None)

    def _rhs(self, grid, rhs, rho, Si, Sij, S):
        g = grid.geom()
        self._e.check(
            self._e.lib.gr_ccz4_rhs(self._e.handle, C.byref(g), C.byref(self._p.c()), grid.getNumVars(),
                                    _ptr(grid.array), _ptr(rhs.array), _ptr(rho), _ptr(Si), _ptr(Sij), _ptr(S)))

    def applyDissipation(self, grid: GridBlock, rhs: GridBlock):
        """src/spacetime/ccz4.cpp:871-935"""
        g = grid.geom()
        self._e.check(self._e.lib.gr_ccz4_dissipation(self._e.handle, C.byref(g), C.byref(self._p.c()),
                                                      grid.getNumVars(), _ptr(grid.array), _ptr(rhs.array)))

    def computeConstraints(self, grid: GridBlock):
        """src/spacetime/ccz4.cpp:941-1159; returns (ham[totalSize], mom[totalSize, 3])."""
        N = grid.totalSize()
        ham = np.zeros(N)
        mom = np.zeros((N, 3))
        g = grid.geom()
        self._e.check(self._e.lib.gr_ccz4_constraints(self._e.handle, C.byref(g), _ptr(grid.array), _ptr(ham),
                                                      _ptr(mom)))
        return ham, mom


class Level:
    """Device-resident blocks of one refinement level (gr_level)."""

    def __init__(self, blocks, engine: Engine | None = None):
        self._e = engine if engine is not None else Engine.default()
        self.blocks = list(blocks)
        geoms = (_Geom * len(self.blocks))(*[b.geom() for b in self.blocks])
        h = C.c_void_p()
        self._e.check(self._e.lib.gr_level_create(self._e.handle, len(self.blocks), geoms, C.byref(h)))
        self._h = h

    def upload(self, i=None):
        for b in range(len(self.blocks)) if i is None else [i]:
            self._e.check(self._e.lib.gr_level_upload(self._h, b, _ptr(self.blocks[b].array)))

    def download(self, i=None):
        for b in range(len(self.blocks)) if i is None else [i]:
            self._e.check(self._e.lib.gr_level_download(self._h, b, _ptr(self.blocks[b].array)))

    def init_brill_lindquist(self, masses, positions):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        self._e.check(self._e.lib.gr_level_init_brill_lindquist(self._h, len(m), _ptr(m), _ptr(p)))

    def add_noise(self, seed, amp, fields):
        mask = 0
        for f in fields:
            mask |= 1 << int(f)
        self._e.check(self._e.lib.gr_level_add_noise(self._h, int(seed), float(amp), mask))

    def fill_ghosts(self, mode):
        self._e.check(self._e.lib.gr_level_fill_ghosts(self._h, int(mode)))

    def init_gauge_wave(self, amplitude, wavelength):
        self._e.check(self._e.lib.gr_level_init_gauge_wave(self._h, float(amplitude), float(wavelength)))

    def set_sources(self, i, d_rho=None, d_Si=None, d_Sij=None, d_S=None):
        """Device source arrays (addresses as ints, e.g. torch tensors' data_ptr()) of block i;
        all None = vacuum."""
        self._e.check(self._e.lib.gr_level_set_sources(self._h, int(i), d_rho, d_Si, d_Sij, d_S))

    def rk3_stage(self, params: CCZ4Params, dt, stage, precision=PREC_FP64, bc_mode=BC_SOMMERFELD,
                  pre_fill=BC_NONE):
        self._e.check(self._e.lib.gr_level_rk3_stage(self._h, C.byref(params.c()), float(dt), int(stage),
                                                     int(precision), int(bc_mode), int(pre_fill)))

    def first_nonfinite(self):
        """None, or (block, var, i, j, k, value) of the first non-finite interior value in the
        reference's scan order (cells counted with ghosts)."""
        out = [C.c_int() for _ in range(6)]
        v = C.c_double()
        self._e.check(self._e.lib.gr_level_first_nonfinite(self._h, *[C.byref(o) for o in out], C.byref(v)))
        if not out[0].value:
            return None
        return tuple(o.value for o in out[1:]) + (v.value,)

    def cfl(self, dt):
        v = C.c_double()
        self._e.check(self._e.lib.gr_level_cfl(self._h, float(dt), C.byref(v)))
        return v.value

    def rk3_step(self, params: CCZ4Params, dt, precision=PREC_FP64, bc_mode=BC_SOMMERFELD, pre_fill=BC_NONE,
                 use_graph=False):
        self._e.check(self._e.lib.gr_level_rk3_step(self._h, C.byref(params.c()), float(dt), int(precision),
                                                    int(bc_mode), int(pre_fill), 1 if use_graph else 0))

    def sync(self):
        self._e.check(self._e.lib.gr_level_sync(self._h))

    def constraints_l2(self):
        s = C.c_double()
        n = C.c_longlong()
        self._e.check(self._e.lib.gr_level_constraints_l2(self._h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def nonfinite(self):
        n = C.c_longlong()
        self._e.check(self._e.lib.gr_level_nonfinite(self._h, C.byref(n)))
        return n.value

    def read_cell(self, i, var, x, y, z):
        """Field var of block i at cell (x, y, z) counted with ghosts (GridBlock::data)."""
        v = C.c_double()
        self._e.check(self._e.lib.gr_level_read_cell(self._h, int(i), int(var), int(x), int(y), int(z), C.byref(v)))
        return v.value

    def set_mixed_policy(self, chi_fp64_below):
        self._e.check(self._e.lib.gr_level_set_mixed_policy(self._h, float(chi_fp64_below)))

    def autotune(self, params: CCZ4Params, dt):
        """Time the stage-kernel staging variants (per group of equal spacing) and the ghost
        fill's field groups on this level, keep the fastest (gr_level_autotune). Returns
        {"stage": [[ms per variant, None where not applicable] per group], "ghost": {groups: ms}}."""
        cap = 256
        t = (C.c_double * cap)()
        n = C.c_int()
        self._e.check(self._e.lib.gr_level_autotune(self._h, C.byref(params.c()), float(dt), t, cap, C.byref(n)))
        vals = [t[i] for i in range(n.value)]
        ng = len(vals) - len(GROUP_CHOICES)
        nv = STAGE_VARIANTS
        stage = [[None if v < 0 else v for v in vals[i:i + nv]] for i in range(0, ng, nv)]
        return {"stage": stage, "ghost": dict(zip(GROUP_CHOICES, vals[ng:]))}

    def tuning(self, kind=TUNE_STAGE, group=0):
        v = C.c_int()
        self._e.check(self._e.lib.gr_level_get_tuning(self._h, int(kind), int(group), C.byref(v)))
        return v.value

    def set_tuning(self, kind, group, value):
        self._e.check(self._e.lib.gr_level_set_tuning(self._h, int(kind), int(group), int(value)))

    def rhs(self, i, params: CCZ4Params, precision=PREC_FP64):
        """RHS of block i's device state (host array shaped like the block)."""
        out = np.zeros_like(self.blocks[i].array)
        self._e.check(self._e.lib.gr_level_rhs(self._h, int(i), C.byref(params.c()), int(precision), _ptr(out)))
        return out

    def profile(self, enable=True):
        self._e.check(self._e.lib.gr_level_profile(self._h, 1 if enable else 0))

    def profile_read(self, kind):
        """(total_ms, launches) of kernel kind 0 (stages 1-2), 1 (last stage), 2 (ghost fills)."""
        t = C.c_double()
        n = C.c_int()
        self._e.check(self._e.lib.gr_level_profile_read(self._h, int(kind), C.byref(t), C.byref(n)))
        return t.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            self._e.lib.gr_level_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Hierarchy:
    """Levels of a Berger-Oliger hierarchy on the device (gr_hierarchy), coarsest first.

    Plays amr::AMRHierarchy for a fixed geometry (include/granite/amr/amr.hpp:49-117):
    fill_ghosts = fillGhostZones, prolongate / restrict = the inter-level transfers of a
    whole level, step = setLevelDt(0, dt0) + subcycle(0) with the SSP-RK3 level step."""

    def __init__(self, levels, engine: Engine | None = None):
        self._e = engine if engine is not None else Engine.default()
        self.levels = list(levels)
        arr = (C.c_void_p * len(self.levels))(*[lv._h for lv in self.levels])
        h = C.c_void_p()
        self._e.check(self._e.lib.gr_hierarchy_create(self._e.handle, len(self.levels), arr, C.byref(h)))
        self._h = h

    def prolongate(self, level):
        self._e.check(self._e.lib.gr_hierarchy_prolongate(self._h, int(level)))

    def restrict(self, level):
        self._e.check(self._e.lib.gr_hierarchy_restrict(self._h, int(level)))

    def fill_ghosts(self, level):
        self._e.check(self._e.lib.gr_hierarchy_fill_ghosts(self._h, int(level)))

    def step(self, params: CCZ4Params, dt0, precision=PREC_FP64, use_graph=False):
        self._e.check(self._e.lib.gr_hierarchy_step(self._h, C.byref(params.c()), float(dt0), int(precision),
                                                    1 if use_graph else 0))

    def constraints_l2(self):
        """(sum H^2, sum |M|^2, count) over the interiors of every block of every level."""
        h = C.c_double()
        m = C.c_double()
        n = C.c_longlong()
        self._e.check(self._e.lib.gr_hierarchy_constraints_l2(self._h, C.byref(h), C.byref(m), C.byref(n)))
        return h.value, m.value, n.value

    def set_cfl_monitor(self, enable=True):
        self._e.check(self._e.lib.gr_hierarchy_set_cfl_monitor(self._h, 1 if enable else 0))

    def autotune(self, params: CCZ4Params, dt0):
        """gr_hierarchy_autotune: each level's variants (dt0 / 2^l), then the transfer field
        groups of every level l > 0. Returns the raw per-launch times (ms) in the C ABI's order."""
        cap = 1024
        t = (C.c_double * cap)()
        n = C.c_int()
        self._e.check(self._e.lib.gr_hierarchy_autotune(self._h, C.byref(params.c()), float(dt0), t, cap, C.byref(n)))
        return [t[i] for i in range(n.value)]

    def tuning(self, level):
        v = C.c_int()
        self._e.check(self._e.lib.gr_hierarchy_get_tuning(self._h, int(level), C.byref(v)))
        return v.value

    def set_tuning(self, level, groups):
        self._e.check(self._e.lib.gr_hierarchy_set_tuning(self._h, int(level), int(groups)))

    def cfl(self):
        """[(level, advective CFL)] of the last coarse step, one per level step in subcycle order."""
        n = C.c_int()
        cap = 1 << 12
        vals = (C.c_double * cap)()
        lev = (C.c_int * cap)()
        self._e.check(self._e.lib.gr_hierarchy_cfl(self._h, vals, lev, cap, C.byref(n)))
        return [(lev[i], vals[i]) for i in range(n.value)]

    def sync(self):
        self._e.check(self._e.lib.gr_engine_sync(self._e.handle))

    def close(self):
        if getattr(self, "_h", None):
            self._e.lib.gr_hierarchy_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Checkpoint:
    """Checkpoint staging (gr_checkpoint) for every block of `levels`, coarsest level first:
    the state HDF5Writer::writeCheckpoint stores (src/io/hdf5_io.cpp:166-232). begin()
    snapshots the device state in engine-stream order and moves it to pinned host memory on
    a copy stream while the evolution continues; wait(), then block(i) reads a copy."""

    def __init__(self, levels, engine: Engine | None = None):
        self._e = engine if engine is not None else Engine.default()
        self.levels = list(levels)
        arr = (C.c_void_p * len(self.levels))(*[lv._h for lv in self.levels])
        h = C.c_void_p()
        self._e.check(self._e.lib.gr_checkpoint_create(self._e.handle, len(self.levels), arr, C.byref(h)))
        self._h = h
        self._geoms = [(lv, b) for lv in self.levels for b in lv.blocks]

    def begin(self):
        self._e.check(self._e.lib.gr_checkpoint_begin(self._h))

    def done(self) -> bool:
        d = C.c_int()
        self._e.check(self._e.lib.gr_checkpoint_query(self._h, C.byref(d)))
        return bool(d.value)

    def wait(self):
        self._e.check(self._e.lib.gr_checkpoint_wait(self._h))

    def __len__(self):
        n = C.c_int()
        self._e.check(self._e.lib.gr_checkpoint_nblocks(self._h, C.byref(n)))
        return n.value

    def block(self, index) -> np.ndarray:
        """Copy of block `index` of the last completed snapshot, shape (22, nz+2g, ny+2g, nx+2g)."""
        p = _dp()
        stride = C.c_long()
        self._e.check(self._e.lib.gr_checkpoint_block(self._h, int(index), C.byref(p), C.byref(stride)))
        g = self._geoms[index][1]
        flat = np.ctypeslib.as_array(p, shape=(22 * stride.value,))
        return flat.reshape(22, g.totalCells(2), g.totalCells(1), g.totalCells(0)).copy()

    def close(self):
        if getattr(self, "_h", None):
            self._e.lib.gr_checkpoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
