"""The pointwise repairs the stage kernels fuse, on states that reach them.

TimeIntegrator::sspRK3Step repairs chi and the lapse after every stage (src/main.cpp:588-608:
chi non-finite or < 1e-4 -> 1e-4, chi > 1.5 -> 1.5, alpha non-finite or < 1e-6 -> 1e-6),
resets a conformal metric whose determinant is not finite or <= 1e-10 to the identity after
the last stage (src/main.cpp:622-645), and computeRHS floors a non-finite or small chi before
using it (src/spacetime/ccz4.cpp:278-279). A Brill-Lindquist block is seeded with cells and
cubes that take every one of these branches; the GPU result is compared with the oracle:
the sets of non-finite values must be equal (a NaN or an infinity poisons the stencils
around it in any implementation), and every finite value is held to the FP64 tolerance.
"""
import sys

import numpy as np
import pytest

from oracle.oracle import Geom, group_rel, normwise_rel, ulp_spread

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


N = 40
H = 0.25


def seeded_state(oracle):
    """BL puncture (M = 1, at a cell corner) on a 40^3 block at C2 spacing, then:
    chi cube at 5e-5 (floor), chi cube at 2.0 (ceiling), lapse cube at 1e-7 (floor), single
    cells with chi = NaN, +Inf, -1 and alpha = NaN, -1, a cube whose conformal metric is
    1e-4 * identity (det 1e-12: reset), and one cell with gt_xx = NaN (non-finite det)."""
    half = H * N / 2
    g = Geom((N, N, N), 4, (-half,) * 3, (half,) * 3)
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    s = 4  # first interior index

    def cube(v, c, r, val):
        k, j, i = (s + x for x in c)
        u[v, k - r:k + r + 1, j - r:j + r + 1, i - r:i + r + 1] = val

    cube(0, (8, 8, 8), 4, 5e-5)
    cube(0, (8, 31, 8), 4, 2.0)
    cube(18, (31, 8, 8), 4, 1e-7)
    for v in (1, 4, 6):
        cube(v, (31, 31, 31), 4, 1e-4)
    for v, c, val in ((0, (20, 8, 30), np.nan), (0, (8, 20, 30), np.inf), (0, (30, 20, 20), -1.0),
                      (18, (20, 30, 8), np.nan), (18, (14, 30, 30), -1.0), (1, (30, 30, 14), np.nan)):
        k, j, i = (s + x for x in c)
        u[v, k, j, i] = val
    return g, u


def compare(got, ref, tol, norm=group_rel):
    """Equal non-finite sets; finite values within tol (a number, or one per field)."""
    bad_g = ~np.isfinite(got)
    bad_r = ~np.isfinite(ref)
    assert np.array_equal(bad_g, bad_r), (int(bad_g.sum()), int(bad_r.sum()))
    a = np.where(bad_r, 0.0, got)
    b = np.where(bad_r, 0.0, ref)
    e = norm(a, b)
    assert (e <= tol).all(), (e, tol)
    return bad_r


def test_rhs_on_poisoned_state(engine, oracle):
    import granite_amd as ga

    g, u = seeded_state(oracle)
    I = g.interior()
    say("rhs of the seeded state (NaN, Inf, negative and tiny chi and lapse, degenerate metric)")
    ev = ga.CCZ4Evolution(ga.CCZ4Params(), engine)
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    rhs = ga.GridBlock(1, 0, g.n, g.lo, g.hi, g.ng)
    ev.computeRHSVacuum(blk, rhs)
    bad = compare(rhs.array[I], oracle.rhs(g, u)[I], TOL, norm=normwise_rel)
    assert bad.any()  # the poisoned cells are part of the comparison


def test_rk3_step_takes_every_repair_branch(engine, oracle):
    import granite_amd as ga

    g, u = seeded_state(oracle)
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    say("one FP64 RK3 step of the seeded state")
    lvl.rk3_step(ga.CCZ4Params(), 0.0625, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
    lvl.download()

    def step(x):
        oracle.fill_ghosts(g, x, 2)
        return oracle.rk3_step(g, x, 0.0625, 0)

    ref = step(u.copy())
    # the cubes at chi = 5e-5 and det = 1e-12 make the step ill-conditioned: a 1-ulp
    # perturbation of the input moves the oracle's own A~ by ~4e-5 of its size, while every
    # other field stays below 1e-12; each field is held to max(1e-12, 2 x that spread)
    spread = ulp_spread(step, u)
    compare(blk.array, ref, np.maximum(TOL, 2.0 * spread))
    s = 4
    at = lambda v, c: ref[v, s + c[0], s + c[1], s + c[2]]  # noqa: E731
    # each branch was taken: the oracle's result holds the repaired values there
    assert at(0, (8, 8, 8)) == 1e-4 and at(0, (8, 31, 8)) == 1.5 and at(18, (31, 8, 8)) == 1e-6
    assert at(0, (20, 8, 30)) == 1e-4 and at(0, (8, 20, 30)) == 1e-4 and at(0, (30, 20, 20)) == 1e-4
    assert at(18, (20, 30, 8)) == 1e-6 and at(18, (14, 30, 30)) == 1e-6
    assert [at(v, (31, 31, 31)) for v in range(1, 7)] == [1, 0, 0, 1, 0, 1]
    assert [at(v, (30, 30, 14)) for v in range(1, 7)] == [1, 0, 0, 1, 0, 1]
    lvl.close()


def _fp64_tiles(ga, chi, n, threshold=0.2):
    """Interior mask of the tiles the mixed form runs in FP64: a tile (the mixed kernel's
    launch tile) whose smallest interior chi is below the threshold or not finite."""
    import ctypes

    t = (ctypes.c_int * 3)()
    ga.library().gr_stage_tile(ga.PREC_MIXED, t)
    tx, ty, tz = t
    mask = np.zeros((n[2], n[1], n[0]), dtype=bool)
    for k0 in range(0, n[2], tz):
        for j0 in range(0, n[1], ty):
            for i0 in range(0, n[0], tx):
                c = chi[k0:k0 + tz, j0:j0 + ty, i0:i0 + tx]
                if not (c >= threshold).all():
                    mask[k0:k0 + tz, j0:j0 + ty, i0:i0 + tx] = True
    return mask


@pytest.mark.parametrize("variant", [0, 2])
def test_mixed_form_routes_nonfinite_tiles_to_fp64(engine, oracle, variant):
    """The mixed form classifies tiles whose smallest chi is below the threshold or not
    finite (the NaN and infinite chi cells, the 5e-5 cube) as FP64 tiles and runs the rest in
    FP32. (a) The RHS of the seeded state: the same non-finite set in both forms; on the tiles
    routed to FP64 the mixed form's RHS meets the FP64 parity bar against the FP64 form's,
    elsewhere it shows FP32 rounding. (b) One RK3
    step: every non-finite value of the FP64 form is non-finite in the mixed form, which
    overflows in addition only where the FP64 state has grown beyond the FP32 algebra's range
    (|value| > 1.8e19, whose square overflows: K = 5.9e32 next to the chi = 5e-5 cube after the
    step), and the repaired values are the same constants in both forms. Both forms run the
    FP64 stage kernel with the given halo staging (0: through registers; 2: straight to LDS),
    which in the mixed form covers the FP64 tiles from a small grid on a second stream."""
    import granite_amd as ga

    g, u = seeded_state(oracle)
    I = g.interior()
    p = ga.CCZ4Params()
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.set_tuning(ga.TUNE_STAGE, 0, variant)
    lvl.upload()
    say(f"seeded state, RHS in both precisions, FP64 staging variant {variant}")
    r64 = lvl.rhs(0, p, ga.PREC_FP64)[I]
    r32 = lvl.rhs(0, p, ga.PREC_MIXED)[I]
    lvl.close()
    assert np.array_equal(~np.isfinite(r32), ~np.isfinite(r64))
    tiles = _fp64_tiles(ga, u[0][I[1:]], g.n)
    assert tiles.any() and not tiles.all()
    # FP64 tiles: the FP64 kernel's arithmetic (launched over the flagged tiles only, so its
    # code is scheduled and contracted on its own: equal to the FP64 parity bar, not bitwise)
    fin = np.isfinite(r64)
    on64 = fin & tiles[None]
    e64 = normwise_rel(np.where(on64, r32, 0.0), np.where(on64, r64, 0.0))
    assert e64.max() < TOL, e64
    # the other tiles ran in FP32: far from the FP64 result at the FP64 bar
    on32 = fin & ~tiles[None]
    e32 = normwise_rel(np.where(on32, r32, 0.0), np.where(on32, r64, 0.0))
    assert e32.max() > 1e-9, e32

    out = {}
    for prec in (ga.PREC_FP64, ga.PREC_MIXED):
        blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
        blk.array[...] = u
        lvl = ga.Level([blk], engine)
        lvl.set_tuning(ga.TUNE_STAGE, 0, variant)
        lvl.upload()
        say(f"seeded state, one RK3 step, precision {prec}, FP64 staging variant {variant}")
        lvl.rk3_step(p, 0.0625, precision=prec, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
        lvl.download()
        out[prec] = blk.array.copy()
        lvl.close()
    a, b = out[ga.PREC_MIXED], out[ga.PREC_FP64]
    assert not (np.isfinite(a) & ~np.isfinite(b)).any()
    only32 = ~np.isfinite(a) & np.isfinite(b)
    assert (np.abs(b[only32]) > 1.8e19).all(), b[only32]
    s = 4
    for v, c, val in ((0, (8, 8, 8), 1e-4), (0, (8, 31, 8), 1.5), (18, (31, 8, 8), 1e-6), (0, (20, 8, 30), 1e-4),
                      (0, (8, 20, 30), 1e-4), (18, (20, 30, 8), 1e-6)):
        assert a[v, s + c[0], s + c[1], s + c[2]] == val


def test_first_nonfinite_and_cfl(engine, oracle):
    """The NaN scan's first hit in the reference's order (field, then k, j, i over the
    interior; cell indices counted with ghosts) and the advective CFL guard's maximum."""
    import granite_amd as ga

    g = Geom((18, 11, 9), 4, (-2.0, -1.0, -1.0), (2.5, 1.75, 1.25))
    u = oracle.init_bl(g, [1.0], [[0.1, 0.0, 0.0]])
    rng = np.random.default_rng(3)
    for v in (19, 20, 21):
        u[v] = rng.standard_normal(u[v].shape) * 0.3
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    say("first non-finite on a finite state, CFL maximum")
    assert lvl.first_nonfinite() is None
    dt = 0.07
    I = g.interior()
    local = sum(np.abs(u[19 + d][I[1:]]) * dt / g.dx[d] for d in range(3))
    assert abs(lvl.cfl(dt) - local.max()) <= 1e-15 * local.max()
    # poison: ghosts are not scanned, later fields and later cells lose to earlier ones
    u[3, 0, 0, 0] = np.nan           # ghost
    u[7, 6, 5, 9] = np.inf           # field 7
    u[5, 10, 4, 12] = -np.inf        # field 5, k = 10
    u[5, 6, 9, 4] = np.nan           # field 5, k = 6: the first in the scan order
    u[19, 5, 5, 5] = np.nan          # NaN shift: ignored by the CFL maximum
    blk.array[...] = u
    lvl.upload()
    say("first non-finite on a poisoned state")
    b, v, i, j, k, val = lvl.first_nonfinite()
    assert (b, v, i, j, k) == (0, 5, 4, 9, 6) and np.isnan(val)
    local = sum(np.abs(u[19 + d][I[1:]]) * dt / g.dx[d] for d in range(3))
    fin = np.where(np.isnan(local), 0.0, local)
    assert np.isnan(local).sum() == 1
    assert abs(lvl.cfl(dt) - fin.max()) <= 1e-15 * fin.max()
    lvl.close()


def test_gauge_wave_initial_data(engine, oracle):
    """setGaugeWaveData on the device (sin, pow, sqrt of the device's math library) against
    the oracle's (the host's libm): equal to a few ulp."""
    import granite_amd as ga

    g = Geom((24, 6, 5), 4, (0.0, 0.0, 0.0), (1.0, 0.25, 0.2))
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    lvl = ga.Level([blk], engine)
    say("gauge-wave initial data")
    lvl.init_gauge_wave(0.1, 1.0)
    lvl.download()
    ref = oracle.init_gauge_wave(g, 0.1, 1.0)
    assert np.array_equal(blk.array == 0.0, ref == 0.0)
    assert np.max(np.abs(blk.array - ref)) <= 4e-16 * np.max(np.abs(ref))
    lvl.close()
