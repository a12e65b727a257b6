"""Parity of the HIP engine (through the C ABI) against the CPU oracle.

Tolerances: FP64 forms are held to 1e-12 normwise relative error per field
(max|gpu - oracle| / max|oracle| over interior cells) for the RHS and the constraints; the
state after RK3 steps is held to 1e-12 of the largest value of the field's tensor group
(components that vanish by symmetry carry only round-off); the Kreiss-Oliger-only path, the
ghost fills and the initial data are bitwise.
The oracle itself agrees bitwise with the compiled reference (test_oracle_golden.py)."""
import sys

import numpy as np
import pytest

from golden_cases import NAMES, load, make_input, sources
from oracle.oracle import Geom, group_rel, normwise_rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def to_block(g: Geom, u):
    import granite_amd as ga

    b = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    assert np.array_equal(b._dx, g.dx)
    b.array[...] = u
    return b


def gpu_rhs(engine, g, u, sources=None, **prm):
    import granite_amd as ga

    p = ga.CCZ4Params(**prm)
    ev = ga.CCZ4Evolution(p, engine)
    blk = to_block(g, u)
    rhs = ga.GridBlock(1, 0, g.n, g.lo, g.hi, g.ng)
    if sources is None:
        ev.computeRHSVacuum(blk, rhs)
    else:
        rho, Si, Sij, S = sources
        ev.computeRHS(blk, rhs, rho, Si.reshape(-1, 3), Sij.reshape(-1, 6), S)
    return rhs.array


@pytest.mark.parametrize("name", NAMES)
def test_rhs_golden_cases(engine, oracle, name):
    rec, z = load(name)
    g, u = make_input(oracle, rec)
    I = g.interior()
    say(f"rhs {name} n={g.n}")
    got = gpu_rhs(engine, g, u)
    e = normwise_rel(got[I], z["rhs"])
    assert e.max() < TOL, e
    # ghost cells of rhs are left untouched (zeros in, zeros out)
    mask = np.ones(got.shape[1:], bool)
    mask[I[1:]] = False
    assert np.all(got[:, mask] == 0.0)


def test_rhs_ko_dominated_fields_bitwise(engine, oracle):
    # zero shift: the metric RHS is the dissipation term alone, accumulated in the reference order
    rec, z = load("bl_binary_noise8")
    g, u = make_input(oracle, rec)
    I = g.interior()
    say("rhs bl_binary_noise8 bitwise metric fields")
    got = gpu_rhs(engine, g, u)
    assert np.array_equal(got[I][1:7], z["rhs"][1:7])


@pytest.mark.parametrize("n", [(32, 32, 32), (20, 12, 9), (5, 7, 3)])
def test_rhs_bl_partial_tiles(engine, oracle, n):
    # C2 spacing (dx = 0.25) around the puncture; odd extents exercise partially filled tiles
    h = 0.25
    lo = tuple(-h * v / 2 for v in n)
    hi = tuple(h * v / 2 for v in n)
    g = Geom(n, 4, lo, hi)
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    oracle.rk3_step(g, u, 0.0625, 0)  # non-trivial values in every field
    I = g.interior()
    say(f"rhs bl evolved n={n}")
    got = gpu_rhs(engine, g, u)
    e = normwise_rel(got[I], oracle.rhs(g, u)[I])
    assert e.max() < TOL, e


def test_rhs_with_sources_and_vacuum_identity(engine, oracle):
    rec, z = load("bl_puncture_noise")
    g, u = make_input(oracle, rec)
    I = g.interior()
    say("rhs with matter sources")
    got = gpu_rhs(engine, g, u, sources=sources(g, 99))
    assert normwise_rel(got[I], z["rhs_sources"]).max() < TOL
    N = g.stride
    zero = (np.zeros(N), np.zeros(3 * N), np.zeros(6 * N), np.zeros(N))
    say("rhs zero sources vs vacuum")
    assert np.array_equal(gpu_rhs(engine, g, u, sources=zero), gpu_rhs(engine, g, u))


def test_rhs_without_dissipation(engine, oracle):
    rec, z = load("bl_puncture_noise")
    g, u = make_input(oracle, rec)
    say("rhs sigma = 0")
    got = gpu_rhs(engine, g, u, ko_sigma=0.0)
    assert normwise_rel(got[g.interior()], z["rhs_nodiss"]).max() < TOL


def test_dissipation_bitwise(engine, oracle):
    import granite_amd as ga

    rec, z = load("bl_puncture_noise")
    g, u = make_input(oracle, rec)
    ev = ga.CCZ4Evolution(ga.CCZ4Params(ko_sigma=0.1), engine)
    rhs = ga.GridBlock(1, 0, g.n, g.lo, g.hi, g.ng)
    say("applyDissipation")
    ev.applyDissipation(to_block(g, u), rhs)
    assert np.array_equal(rhs.array[g.interior()], z["ko"])


@pytest.mark.parametrize("name", ["bl_puncture", "bl_puncture_noise", "gauge_wave"])
def test_constraints(engine, oracle, name):
    import granite_amd as ga

    rec, z = load(name)
    g, u = make_input(oracle, rec)
    I = g.interior()
    ev = ga.CCZ4Evolution(ga.CCZ4Params(), engine)
    say(f"constraints {name}")
    ham, mom = ev.computeConstraints(to_block(g, u))
    ham = ham.reshape(g.nt[2], g.nt[1], g.nt[0])
    mom = mom.reshape(g.nt[2], g.nt[1], g.nt[0], 3)
    scale_h = max(np.max(np.abs(z["ham"])), 1e-300)
    assert np.max(np.abs(ham[I[1:]] - z["ham"])) / scale_h < TOL
    for a in range(3):
        s = max(np.max(np.abs(z["mom"][..., a])), 1e-300)
        assert np.max(np.abs(mom[I[1:]][..., a] - z["mom"][..., a])) / s < TOL


def make_level(engine, g, u):
    import granite_amd as ga

    blk = to_block(g, u)
    lvl = ga.Level([blk], engine)
    lvl.upload()
    return lvl, blk


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ghost_fill_bitwise(engine, oracle, mode):
    g = Geom((13, 9, 11), 4, (-1.3, -0.4, 0.2), (1.1, 0.6, 1.6))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.3, 0.1, 0.7]]), 11, 0.1, range(22))
    lvl, blk = make_level(engine, g, u)
    say(f"ghost fill mode {mode}")
    lvl.fill_ghosts(mode)
    lvl.download()
    assert np.array_equal(blk.array, oracle.fill_ghosts(g, u.copy(), mode))
    lvl.close()


def test_initial_data_bitwise(engine, oracle):
    g = Geom((17, 12, 10), 4, (-2.1, -1.0, -0.9), (2.15, 1.4, 1.1))
    lvl, blk = make_level(engine, g, g.zeros())
    say("brill-lindquist initial data")
    lvl.init_brill_lindquist([0.5, 0.3], [[-1.0, 0.1, 0.0], [1.2, -0.1, 0.05]])
    lvl.download()
    assert np.array_equal(blk.array, oracle.init_bl(g, [0.5, 0.3], [[-1.0, 0.1, 0.0], [1.2, -0.1, 0.05]]))
    lvl.close()


def test_noise_matches_portable_generator(engine, oracle):
    g = Geom((6, 5, 4), 4, (0, 0, 0), (1, 1, 1))
    lvl, blk = make_level(engine, g, g.zeros())
    say("seeded noise")
    lvl.add_noise(123, 1e-8, [1, 2, 13])
    lvl.download()
    ref = oracle.add_noise(g, g.zeros(), 123, 1e-8, [1, 2, 13])
    assert np.max(np.abs(blk.array - ref)) <= 1e-8 * 1e-12
    lvl.close()


@pytest.mark.parametrize("case", ["bl_sommerfeld", "bl_outflow_prefill", "gauge_periodic", "binary_noise_outflow"])
def test_rk3_step_parity(engine, oracle, case):
    import granite_amd as ga

    if case == "gauge_periodic":
        g = Geom((16, 8, 8), 4, (0, 0, 0), (1, 0.5, 0.5))
        u = oracle.init_gauge_wave(g, 0.1, 1.0)
        bc, pre, dt = 1, -1, 0.25 / 16
    elif case == "binary_noise_outflow":
        # two punctures with 1e-8 noise on the conformal metric and K: the metric's RHS starts
        # as pure dissipation of the noise, the regime the exact-order KO sum is for
        g = Geom((20, 16, 12), 4, (-2.5, -2.0, -1.5), (2.5, 2.0, 1.5))
        u = oracle.add_noise(g, oracle.init_bl(g, [0.5, 0.5], [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]]),
                             20240611, 1e-8, [1, 2, 3, 4, 5, 6, 13])
        bc, pre, dt = 0, 2, 0.0625
    else:
        g = Geom((24, 20, 16), 4, (-3.0, -2.5, -2.0), (3.0, 2.5, 2.0))
        u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
        bc, pre, dt = 0, (2 if case == "bl_outflow_prefill" else -1), 0.0625
    lvl, blk = make_level(engine, g, u)
    ref = u.copy()
    p = ga.CCZ4Params()
    for step in range(3):
        say(f"rk3 {case} step {step}")
        lvl.rk3_step(p, dt, bc_mode=bc, pre_fill=pre)
        if pre >= 0:
            oracle.fill_ghosts(g, ref, pre)
        oracle.rk3_step(g, ref, dt, bc_mode=bc)
    lvl.download()
    e = group_rel(blk.array, ref)
    assert e.max() < TOL, e
    lvl.close()


def test_c1_gauge_wave_to_t1(engine, oracle, oracle_release):
    """C1 at its full size: the gauge wave (setGaugeWaveData, A = 0.1, L = 1) on 32^3 cells of
    [0, 1)^3 with periodic ghosts, 128 RK3 steps of dt = dx / 4 to t = 1, graph-replayed,
    against the oracle's 128 steps: every field within max(1e-12, 2 x the departure of the
    reference's Release build from the strict oracle over the same 128 steps). The Release
    build itself departs by up to 2.5e-11 (Theta) by t = 1: round-off, accumulated by a
    long evolution whose K and Theta stay near zero."""
    import granite_amd as ga

    g = Geom((32, 32, 32), 4, (0, 0, 0), (1, 1, 1))
    u = oracle.init_gauge_wave(g, 0.1, 1.0)
    lvl, blk = make_level(engine, g, u)
    p = ga.CCZ4Params()
    dt = 1.0 / 128
    say("C1 gauge wave: 128 steps on the device")
    for _ in range(128):
        lvl.rk3_step(p, dt, bc_mode=1, pre_fill=-1, use_graph=True)
    lvl.download()
    say("C1 gauge wave: 128 steps of the oracle, strict and Release")
    ref, rel = u.copy(), u.copy()
    for _ in range(128):
        oracle.rk3_step(g, ref, dt, bc_mode=1)
        oracle_release.rk3_step(g, rel, dt, bc_mode=1)
    e = group_rel(blk.array, ref)
    release = group_rel(rel, ref)
    assert (e <= np.maximum(1e-12, 2.0 * release)).all(), (e, release)
    lvl.close()


def test_graph_replay_equals_eager_and_is_deterministic(engine, oracle):
    import granite_amd as ga

    g = Geom((16, 16, 16), 4, (-2, -2, -2), (2, 2, 2))
    u = oracle.init_bl(g, [1.0], [[0.1, 0.0, -0.1]])
    p = ga.CCZ4Params()
    outs = []
    for use_graph in (False, True, True):
        lvl, blk = make_level(engine, g, u)
        say(f"determinism use_graph={use_graph}")
        for _ in range(4):
            lvl.rk3_step(p, 0.0625, pre_fill=2, use_graph=use_graph)
            lvl.sync()
        lvl.download()
        outs.append(blk.array.copy())
        lvl.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


@pytest.mark.parametrize("precision", [0, 1])
def test_direct_staging_equals_register_staging_and_repeats(engine, oracle, precision):
    """The direct-to-LDS staging (variant 2: three field buffers, a barrier that leaves the
    next field's loads in flight) against the register staging (variant 0) on a 64x48x40
    block with noise on every field (shift signs mixed cell by cell), four graph-replayed
    steps: the two evaluate the same arithmetic per cell, so they agree bit for bit, and
    variant 2 repeats itself bit for bit. Precision 1 runs the mixed form, whose FP64 tiles
    take the chosen staging."""
    import granite_amd as ga

    g = Geom((64, 48, 40), 4, (-4.0, -3.0, -2.5), (4.0, 3.0, 2.5))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.05, -0.02, 0.01]]), 23, 1e-3, range(22))
    p = ga.CCZ4Params()
    outs = []
    for variant in (2, 0, 2):
        lvl, blk = make_level(engine, g, u)
        lvl.set_tuning(ga.TUNE_STAGE, 0, variant)
        say(f"staging variant {variant}, precision {precision}: four steps")
        for _ in range(4):
            lvl.rk3_step(p, 0.03, precision=precision, pre_fill=ga.BC_OUTFLOW, use_graph=True)
        lvl.download()
        outs.append(blk.array.copy())
        lvl.close()
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    assert np.array_equal(outs[0].view(np.uint64), outs[2].view(np.uint64))


def test_level_monitor_and_scan(engine, oracle):
    import granite_amd as ga

    g = Geom((16, 12, 12), 4, (-2, -1.5, -1.5), (2, 1.5, 1.5))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.05, 0.0, 0.0]]), 2, 1e-4, [1, 4, 6, 13])
    lvl, blk = make_level(engine, g, u)
    say("level constraint L2 and non-finite scan")
    s, n = lvl.constraints_l2()
    h, _ = oracle.constraints(g, u)
    ref = float(np.sum(h[g.interior()[1:]] ** 2))
    assert n == 16 * 12 * 12
    assert abs(s - ref) <= 1e-12 * ref
    assert lvl.nonfinite() == 0
    u2 = u.copy()
    u2[5, 6, 7, 8] = np.nan
    u2[0, 4, 4, 4] = np.inf
    blk.array[...] = u2
    lvl.upload()
    assert lvl.nonfinite() == 2
    lvl.close()


def test_bad_arguments_raise(engine):
    import granite_amd as ga

    ev = ga.CCZ4Evolution(ga.CCZ4Params(), engine)
    b = ga.GridBlock(0, 0, (4, 4, 4), (0, 0, 0), (1, 1, 1), 3)  # ng = 3 is too few ghosts
    r = ga.GridBlock(1, 0, (4, 4, 4), (0, 0, 0), (1, 1, 1), 3)
    say("error path ng=3")
    with pytest.raises(ga.EngineError) as ei:
        ev.computeRHSVacuum(b, r)
    assert ei.value.code == -2
    b = ga.GridBlock(0, 0, (4, 4, 4), (0, 0, 0), (1, 1, 1), 4, num_vars=21)
    r = ga.GridBlock(1, 0, (4, 4, 4), (0, 0, 0), (1, 1, 1), 4, num_vars=21)
    with pytest.raises(ga.EngineError) as ei:
        ev.computeRHSVacuum(b, r)
    assert ei.value.code == -3

    # level geometry limits, checked before any allocation: a block over 2^29 cells with
    # ghosts (32-bit byte offsets per field), an empty extent, no blocks at all
    import ctypes as C

    from granite_amd.ccz4 import _Geom

    lib = engine.lib
    out = C.c_void_p()
    for n in ((1024, 1024, 1024), (8, 0, 8)):
        say(f"error path level of {n}")
        g = _Geom((C.c_int * 3)(*n), 4, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(0.1, 0.1, 0.1))
        assert lib.gr_level_create(engine.handle, 1, C.byref(g), C.byref(out)) == -2
        assert not out.value
    assert lib.gr_level_create(engine.handle, 0, None, C.byref(out)) == -1


def test_full_size_step_against_oracle(engine, oracle):
    """The headline workload itself: the 128^3 Brill-Lindquist block (M = 1 at the origin, dx =
    0.25 on [-16,16]^3), one graph-replayed RK3 step with the level-0 outflow fill before it
    and 1/r ghosts after every stage, untuned (the direct-to-LDS staging a level of this size
    starts on), against the oracle's step: every field within 1e-12 of its tensor group. The
    oracle's step takes about half a minute on one core."""
    import granite_amd as ga

    n = 128
    g = Geom((n, n, n), 4, (-16.0,) * 3, (16.0,) * 3)
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    lvl, blk = make_level(engine, g, u)
    assert lvl.tuning(ga.TUNE_STAGE, 0) == 2
    say("full size 128^3: one step on the device")
    lvl.rk3_step(ga.CCZ4Params(), 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.download()
    lvl.close()
    say("full size 128^3: one step of the oracle")
    oracle.fill_ghosts(g, u, 2)
    oracle.rk3_step(g, u, 0.0625, 0)
    e = group_rel(blk.array, u)
    assert e.max() < TOL, e


def test_full_size_block_properties(engine):
    """The 128^3 Brill-Lindquist block (dx = 0.25 on [-16,16]^3): after two graph-replayed
    steps the state is finite and mirror-symmetric about x = 0 (scalars even, x-odd tensor
    components odd), and the constraint monitor is finite."""
    import granite_amd as ga

    n = 128
    blk = ga.GridBlock(0, 0, (n, n, n), (-16.0,) * 3, (16.0,) * 3, 4)
    lvl = ga.Level([blk], engine)
    say("full size 128^3 init + 2 steps")
    lvl.init_brill_lindquist([1.0], [[0.0, 0.0, 0.0]])
    p = ga.CCZ4Params()
    for _ in range(2):
        lvl.rk3_step(p, 0.0625, pre_fill=2, use_graph=True)
    lvl.sync()
    assert lvl.nonfinite() == 0
    s, cnt = lvl.constraints_l2()
    assert np.isfinite(s) and cnt == n**3
    lvl.download()
    a = blk.array[blk.interior()]
    odd_x = {2, 3, 8, 9, 14, 19}  # gt_xy, gt_xz, A_xy, A_xz, Gamma-hat^x, beta^x
    for v in range(22):
        m = a[v][:, :, ::-1]
        sign = -1.0 if v in odd_x else 1.0
        scale = max(np.max(np.abs(a[v])), 1e-300)
        assert np.max(np.abs(a[v] - sign * m)) / scale < 1e-12, v
    lvl.close()
