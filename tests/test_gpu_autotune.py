"""Every kernel variant the autotuner chooses from, against the oracle, and the tuner itself.

The stage kernel's two halo stagings (prefetch through registers into a padded LDS box;
direct global-to-LDS loads into an unpadded box) are two schedules of the same arithmetic:
each meets the FP64 parity bar. The ghost fill and the AMR transfers compute every value the
same way whatever their field grouping: each grouping is bit-identical to the oracle. The
tuner's choice is a launch form only: a step after tuning equals a step with the same choices
forced, bit for bit."""
import os
import sys

import numpy as np
import pytest

from oracle.hierarchy import OracleHierarchy, build_levels
from oracle.oracle import Geom, group_rel, normwise_rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def bl_case(oracle):
    # C2 spacing around the puncture; even row length (20 + 8) so both stagings apply
    g = Geom((20, 12, 16), 4, (-2.5, -1.5, -2.0), (2.5, 1.5, 2.0))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.05, -0.02, 0.01]]), 17, 1e-6, range(22))
    return g, u


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_stage_variants_against_oracle(engine, oracle, variant):
    import granite_amd as ga

    g, u = bl_case(oracle)
    I = g.interior()
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    lvl.set_tuning(ga.TUNE_STAGE, 0, variant)
    assert lvl.tuning(ga.TUNE_STAGE, 0) == variant
    p = ga.CCZ4Params()
    say(f"staging variant {variant}: RHS")
    e = normwise_rel(lvl.rhs(0, p)[I], oracle.rhs(g, u)[I])
    assert e.max() < TOL, e
    say(f"staging variant {variant}: two RK3 steps")
    ref = u.copy()
    for step in range(2):
        lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=step == 1)
        oracle.fill_ghosts(g, ref, 2)
        oracle.rk3_step(g, ref, 0.0625, 0)
    lvl.download()
    e = group_rel(blk.array, ref)
    assert e.max() < TOL, e
    lvl.close()


def test_direct_staging_refused_for_odd_rows(engine, oracle):
    import granite_amd as ga

    g = Geom((13, 8, 8), 4, (-1.6, -1.0, -1.0), (1.65, 1.0, 1.0))
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    lvl = ga.Level([blk], engine)
    for v in (1, 2):
        with pytest.raises(ga.EngineError):
            lvl.set_tuning(ga.TUNE_STAGE, 0, v)
    assert lvl.tuning(ga.TUNE_STAGE, 0) == 0
    lvl.close()


@pytest.mark.parametrize("groups", [1, 2, 4, 11])
def test_ghost_groups_bitwise(engine, oracle, groups):
    import granite_amd as ga

    g = Geom((13, 9, 11), 4, (-1.3, -0.4, 0.2), (1.1, 0.6, 1.6))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.3, 0.1, 0.7]]), 11, 0.1, range(22))
    for mode in (0, 1, 2):
        blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
        blk.array[...] = u
        lvl = ga.Level([blk], engine)
        lvl.upload()
        lvl.set_tuning(ga.TUNE_GHOST, 0, groups)
        say(f"ghost fill mode {mode}, {groups} field groups")
        lvl.fill_ghosts(mode)
        lvl.download()
        assert np.array_equal(blk.array, oracle.fill_ghosts(g, u.copy(), mode))
        lvl.close()


@pytest.mark.parametrize("groups", [1, 2, 4, 11])
def test_transfer_groups_bitwise(engine, oracle, groups):
    import granite_amd as ga

    geo = [
        [([16, 16, 16], [-8.0] * 3, [8.0] * 3)],
        [([12, 16, 16], [-5.0, -4.0, -4.0], [1.0, 4.0, 4.0]), ([8, 12, 12], [1.0, -3.0, -3.0], [5.0, 3.0, 3.0])],
        [([32, 12, 12], [-2.0, -4.5, -1.5], [6.0, -1.5, 1.5])],
    ]
    ol = build_levels(oracle, geo, masses=[0.5, 0.5], positions=[[-1.0, 0.0, 0.0], [1.2, 0.0, 0.0]])
    for l, lv in enumerate(ol):
        for gg, u in lv:
            oracle.add_noise(gg, u, 31 + l, 1e-3, range(22))
    levels, blocks = [], []
    for lv in ol:
        bl = [ga.GridBlock(i, 0, gg.n, gg.lo, gg.hi, gg.ng) for i, (gg, _) in enumerate(lv)]
        for b, (_, u) in zip(bl, lv):
            b.array[...] = u
        L = ga.Level(bl, engine)
        L.upload()
        levels.append(L)
        blocks.append(bl)
    H = ga.Hierarchy(levels, engine)
    OH = OracleHierarchy(oracle, ol)
    for l in (1, 2):
        H.set_tuning(l, groups)
        assert H.tuning(l) == groups
    say(f"transfers with {groups} field groups")
    for l in (2, 1):
        H.restrict(l)
        OH.restrict(l)
    for l in (1, 2):
        H.prolongate(l)
        OH.prolongate(l)
    for L, bl, lv in zip(levels, blocks, ol):
        L.download()
        for b, (_, u) in zip(bl, lv):
            assert np.array_equal(b.array, u)
    H.close()
    for L in levels:
        L.close()


def test_autotune_choice_is_a_launch_form(engine, oracle, tmp_path):
    """After gr_level_autotune the level runs the chosen variants; a step equals a step of a
    level with the same choices forced; the table written by gr_tune_save gives a new level of
    the same size class the same choices after gr_tune_clear + gr_tune_load."""
    import granite_amd as ga

    g, u = bl_case(oracle)
    p = ga.CCZ4Params()
    ga.tune_clear()
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    say("autotune a level")
    times = lvl.autotune(p, 0.0625)
    assert len(times["stage"]) == 1 and len(times["stage"][0]) == 3
    assert all(t is not None and t > 0 for t in times["stage"][0])
    assert sorted(times["ghost"]) == [1, 2, 4, 11] and all(t > 0 for t in times["ghost"].values())
    sv, gg = lvl.tuning(ga.TUNE_STAGE, 0), lvl.tuning(ga.TUNE_GHOST, 0)
    st = times["stage"][0]
    assert st[sv] == min(st)
    assert times["ghost"][gg] == min(times["ghost"].values())
    lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.download()
    tuned = blk.array.copy()
    lvl.close()
    table = os.path.join(str(tmp_path), "tune.txt")
    ga.tune_save(table)
    ga.tune_clear()
    fresh = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    fresh.array[...] = u
    lvl = ga.Level([fresh], engine)
    assert lvl.tuning(ga.TUNE_STAGE, 0) == 0 and lvl.tuning(ga.TUNE_GHOST, 0) == 4  # defaults
    lvl.set_tuning(ga.TUNE_STAGE, 0, sv)
    lvl.set_tuning(ga.TUNE_GHOST, 0, gg)
    lvl.upload()
    lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.download()
    assert np.array_equal(fresh.array, tuned)
    lvl.close()
    ga.tune_load(table)
    lvl = ga.Level([ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)], engine)
    assert (lvl.tuning(ga.TUNE_STAGE, 0), lvl.tuning(ga.TUNE_GHOST, 0)) == (sv, gg)
    lvl.close()
    ga.tune_clear()
