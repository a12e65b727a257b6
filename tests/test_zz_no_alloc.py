"""The allocation-free constraint monitor (gr_ccz4_constraints_l2_noalloc) against the level
monitor it varies (gr_level_constraints_l2): the same sum of H^2 over the interior, to the FP64
parity bar, on two calls in a row whose second block is smaller (the device scratch keeps the
first call's partials; the second must not read them). Called through the C ABI."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle.oracle import Geom

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _level(ga, engine, oracle, n, lo, hi, seed):
    g = Geom(n, 4, lo, hi)
    u = oracle.init_bl(g, [1.0], [[0.05, -0.02, 0.01]])
    oracle.add_noise(g, u, seed, 1e-3, range(22))
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    return blk, lvl


def _noalloc(ga, engine, blk, lvl, out):
    lib = engine.lib
    d_u = C.c_void_p()
    engine.check(lib.gr_level_state(lvl._h, 0, C.byref(d_u)))
    count = C.c_longlong(-1)
    engine.check(lib.gr_ccz4_constraints_l2_noalloc(engine.handle, C.byref(blk.geom()), d_u,
                                                    C.c_void_p(out.data_ptr()), C.byref(count)))
    engine.sync()
    return float(out.to_host()[0]), count.value


def test_noalloc_monitor_matches_level_monitor(engine, oracle):
    import granite_amd as ga
    from hipmem import DeviceArray

    out = DeviceArray(np.full(1, np.nan))
    cases = [((40, 36, 28), (-2.5, -2.25, -1.75), (2.5, 2.25, 1.75), 5),
             ((13, 9, 7), (-1.0, -0.75, -0.5), (1.0, 0.75, 0.5), 6)]
    for n, lo, hi, seed in cases:
        blk, lvl = _level(ga, engine, oracle, n, lo, hi, seed)
        say(f"no-alloc monitor on {n}")
        got, cnt = _noalloc(ga, engine, blk, lvl, out)
        want, wcnt = lvl.constraints_l2()
        assert cnt == wcnt == int(np.prod(n))
        assert np.isfinite(want) and want > 0.0
        assert abs(got - want) <= TOL * want, (n, got, want)
        lvl.close()
    out.close()


def test_noalloc_monitor_rejects_oversized_block(engine):
    from granite_amd.ccz4 import _Geom

    big = _Geom()  # 1024^3 interior cells: over 2^29 with ghosts (geometry only, nothing allocated)
    for d in range(3):
        big.n[d] = 1024
        big.lo[d] = 0.0
        big.dx[d] = 1.0 / 1024
    big.ng = 4
    count = C.c_longlong(-1)
    rc = engine.lib.gr_ccz4_constraints_l2_noalloc(engine.handle, C.byref(big), C.c_void_p(8), C.c_void_p(8),
                                                   C.byref(count))
    assert rc == -2  # GR_ERR_GEOM, before any launch or device access
    assert b"2^29" in engine.lib.gr_last_error(engine.handle)
