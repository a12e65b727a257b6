"""Batched levels: all blocks of a level advance with one stage launch and one ghost-fill
launch per RK3 stage. Each block must come out exactly as when it is evolved alone
(bitwise: the same kernels compute the same cells), and small uneven blocks agree with
the oracle to the FP64 tolerance."""
import os
import sys

import numpy as np
import pytest

from oracle.oracle import Geom, group_rel, normwise_rel, ulp_spread

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def uneven_blocks(ga, h=0.125, seed=5, count=7):
    rng = np.random.default_rng(seed)
    edges = [8, 12, 16, 20, 24, 36]
    blocks = []
    for i in range(count):
        n = [int(rng.choice(edges)) for _ in range(3)]
        lo = [float(rng.integers(-12, 12)) * h for _ in range(3)]
        hi = [lo[d] + n[d] * h for d in range(3)]
        blocks.append(ga.GridBlock(i, 0, n, lo, hi, 4))
    return blocks


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_level_equals_blocks_alone(engine, precision):
    import granite_amd as ga

    p = ga.CCZ4Params()
    masses, pos = [0.5, 0.4], [[-0.6, 0.1, 0.0], [0.7, -0.05, 0.1]]
    blocks = uneven_blocks(ga)
    lvl = ga.Level(blocks, engine)
    lvl.init_brill_lindquist(masses, pos)
    say(f"batched level of {len(blocks)} blocks, precision {precision}")
    for _ in range(3):
        lvl.rk3_step(p, 0.03125, precision=precision, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.sync()
    lvl.download()
    for i, b in enumerate(blocks):
        alone = ga.GridBlock(0, 0, b.numCells(), b.lowerCorner(), b.upperCorner(), 4)
        one = ga.Level([alone], engine)
        one.init_brill_lindquist(masses, pos)
        say(f"block {i} alone n={b.numCells()}")
        for _ in range(3):
            one.rk3_step(p, 0.03125, precision=precision, pre_fill=ga.BC_OUTFLOW, use_graph=False)
        one.download()
        assert np.array_equal(alone.array, b.array), i
        one.close()
    lvl.close()


def test_batched_level_against_oracle(engine, oracle):
    import granite_amd as ga

    p = ga.CCZ4Params()
    blocks = uneven_blocks(ga, seed=9, count=3)
    lvl = ga.Level(blocks, engine)
    masses, pos = [1.0], [[0.05, -0.02, 0.01]]
    lvl.init_brill_lindquist(masses, pos)
    refs = []
    for b in blocks:
        g = Geom(b.numCells(), 4, b.lowerCorner(), b.upperCorner())
        refs.append((g, oracle.init_bl(g, masses, pos)))
    say("batched level vs oracle, 2 steps")
    for _ in range(2):
        lvl.rk3_step(p, 0.03125, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
        for g, u in refs:
            oracle.fill_ghosts(g, u, 2)
            oracle.rk3_step(g, u, 0.03125, 0)
    lvl.download()
    for (g, u), b in zip(refs, blocks):
        e = group_rel(b.array, u)
        assert e.max() < 1e-12, e
    lvl.close()


def test_level_with_unequal_spacing(engine):
    """Blocks of one level whose dx differ (by one ulp, as separately built boxes of one level
    can; or entirely) run in one launch per spacing, each with its own stencil constants, and
    come out bit-identical to the same block evolved alone."""
    import granite_amd as ga

    p = ga.CCZ4Params()
    h = 0.4

    def box(i, n, lo):
        # GridBlock's dx = (hi - lo) / n of hi = lo + n * h: h or an ulp either side of it
        return ga.GridBlock(i, 0, n, lo, [lo[d] + n[d] * h for d in range(3)], 4)

    a = box(0, (8, 10, 12), (-18.31, -18.31, -18.31))   # dx 0.4 - 1ulp, 0.4, 0.4 + 1ulp
    b = box(1, (8, 14, 8), (-5.79, -5.79, 12.04))       # 0.4, 0.4, 0.4 - 1ulp
    c = ga.GridBlock(2, 0, (8, 8, 8), (-3.2, -3.2, -3.2), (3.2, 3.2, 3.2), 4)  # dx 0.8
    dxs = {tuple(x.cellSpacing()) for x in (a, b, c)}
    say(f"level with spacings {sorted(dxs)}")
    assert len(dxs) == 3
    masses, pos = [0.5], [[0.3, 0.2, 0.1]]
    lvl = ga.Level([a, b, c], engine)
    lvl.init_brill_lindquist(masses, pos)
    for _ in range(2):
        lvl.rk3_step(p, 0.05, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.download()
    for blk in (a, b, c):
        alone = ga.GridBlock(0, 0, blk.numCells(), blk.lowerCorner(), blk.upperCorner(), 4)
        one = ga.Level([alone], engine)
        one.init_brill_lindquist(masses, pos)
        for _ in range(2):
            one.rk3_step(p, 0.05, pre_fill=ga.BC_OUTFLOW)
        one.download()
        assert np.array_equal(alone.array, blk.array)
        one.close()
    lvl.close()


def c4_levels(ga, engine):
    from granite_amd import workloads

    layout = workloads.uneven_amr_levels()
    out = []
    for dx, blocks in layout:
        grids = [ga.GridBlock(i, 0, n, lo, hi, 4) for i, (n, lo, hi) in enumerate(blocks)]
        out.append((dx, grids))
    return out


@pytest.mark.parametrize("precision,variant", [(0, 0), (1, 0), (0, 2), (1, 2)])
def test_c4_layout_batched_equals_blocks_alone(engine, precision, variant, tmp_path):
    """The batched-AMR workload itself (granite_amd.workloads.uneven_amr_levels: 250 blocks,
    edges from {16, 24, 32, 48, 64}, 4 levels): one RK3 step of every level in one stage launch
    per level; every block (FP64) or every fourth block (mixed) bit-identical to the block
    evolved alone. With the FP64 halo staging through registers (variant 0, the default) and
    straight to LDS in 8x4x8 parts (variant 2, taken from a tuning table for every block-size
    class, so the batched levels and the single blocks alike run it)."""
    import granite_amd as ga

    from granite_amd import workloads

    ga.tune_clear()
    if variant:
        table = os.path.join(str(tmp_path), "tune.txt")
        with open(table, "w") as f:
            f.write("".join(f"0 {c} 0 {variant}\n" for c in range(4)))
        ga.tune_load(table)
    try:
        _c4_batched_equals_alone(ga, workloads, engine, precision)
    finally:
        ga.tune_clear()


def _c4_batched_equals_alone(ga, workloads, engine, precision):
    p = ga.CCZ4Params()
    masses, pos = workloads.PUNCTURES
    levels = c4_levels(ga, engine)
    assert sum(len(g) for _, g in levels) == 250
    for li, (dx, grids) in enumerate(levels):
        lvl = ga.Level(grids, engine)
        lvl.init_brill_lindquist(masses, pos)
        say(f"C4 level {li}: {len(grids)} blocks batched, precision {precision}")
        lvl.rk3_step(p, 0.25 * dx, precision=precision, pre_fill=ga.BC_OUTFLOW, use_graph=True)
        lvl.download()
        lvl.close()
        step = 1 if precision == 0 else 4
        for i in range(0, len(grids), step):
            g = grids[i]
            alone = ga.GridBlock(0, 0, g.numCells(), g.lowerCorner(), g.upperCorner(), 4)
            one = ga.Level([alone], engine)
            one.init_brill_lindquist(masses, pos)
            one.rk3_step(p, 0.25 * dx, precision=precision, pre_fill=ga.BC_OUTFLOW)
            one.download()
            assert np.array_equal(alone.array, g.array), (li, i)
            one.close()
            g.array = None  # release the host copy


def test_c4_layout_against_oracle(engine, oracle, oracle_release):
    """The smallest block of each level of the batched-AMR workload, in its level's batched
    layout, against the oracle.

    (a) The RHS of the initial state, per field within max(1e-12, 2 x the departure of the
    reference's Release build from the strict oracle on the same input). Far from the
    punctures the RHS of K, Theta and A~ is a small remainder of O(1) stencil sums (K's is
    ~2 |grad psi|^2 / psi^8), so any change of summation order moves it: the Release build
    itself departs by up to 4.7e-9 (K) there, 3e-14 near the punctures.
    (b) One batched RK3 step of the level, each field within max(1e-12, 2 x spread), where
    spread is the oracle's own change under 1-ulp perturbations of its input
    (oracle.ulp_spread): A~ after one step is small beside the metric's last-bit noise that
    the second-derivative stencils feed into it (spread up to 1.7e-10 of its size)."""
    import granite_amd as ga

    from granite_amd import workloads

    p = ga.CCZ4Params()
    masses, pos = workloads.PUNCTURES
    for li, (dx, grids) in enumerate(c4_levels(ga, engine)):
        lvl = ga.Level(grids, engine)
        lvl.init_brill_lindquist(masses, pos)
        i = min(range(len(grids)), key=lambda q: int(np.prod(grids[q].numCells())))
        b = grids[i]
        gg = Geom(b.numCells(), 4, b.lowerCorner(), b.upperCorner())
        u = oracle.init_bl(gg, masses, pos)
        I = gg.interior()
        say(f"C4 level {li}: RHS of block {i} vs oracle")
        ref = oracle.rhs(gg, u)[I]
        release = normwise_rel(oracle_release.rhs(gg, u)[I], ref)
        e = normwise_rel(lvl.rhs(i, p)[I], ref)
        assert (e <= np.maximum(1e-12, 2.0 * release)).all(), (li, "rhs", e, release)
        say(f"C4 level {li}: RK3 step vs oracle")
        lvl.rk3_step(p, 0.25 * dx, pre_fill=ga.BC_OUTFLOW)
        lvl.download(i)

        def step(x):
            oracle.fill_ghosts(gg, x, 2)
            return oracle.rk3_step(gg, x, 0.25 * dx, 0)

        spread = ulp_spread(step, u)
        ref = step(u.copy())
        e = group_rel(b.array, ref)
        tol = np.maximum(1e-12, 2.0 * spread)
        assert (e <= tol).all(), (li, e, spread)
        lvl.close()


@pytest.mark.parametrize("precision,use_graph", [(0, False), (0, True), (1, True)])
def test_level_set_step_equals_levels_one_by_one(engine, precision, use_graph):
    """gr_levels_rk3_step (each level's ghost fills on a second stream, overlapping the next
    levels' stage kernels) gives every level the state that stepping the levels one by one
    gives, bit for bit: three batched levels of uneven blocks with their own dt, two steps."""
    import granite_amd as ga

    p = ga.CCZ4Params()
    masses, pos = [0.5, 0.5], [[-0.6, 0.0, 0.0], [0.6, 0.05, 0.0]]
    specs = [
        (0.25, [((16, 8, 8), (-2.0, -1.0, -1.0)), ((12, 16, 8), (2.0, -2.0, -1.0)), ((8, 8, 12), (-1.0, 1.0, 1.0))]),
        (0.125, [((24, 16, 8), (-1.5, -1.0, -0.5)), ((16, 16, 16), (1.5, -1.0, -1.0))]),
        (0.0625, [((32, 16, 8), (-1.0, -0.5, -0.25))]),
    ]

    def build():
        out = []
        for dx, blocks in specs:
            grids = [ga.GridBlock(i, 0, n, lo, tuple(l + dx * k for l, k in zip(lo, n)), 4)
                     for i, (n, lo) in enumerate(blocks)]
            lvl = ga.Level(grids, engine)
            lvl.init_brill_lindquist(masses, pos)
            out.append((lvl, grids, 0.25 * dx))
        return out

    together = build()
    say(f"level set, precision {precision}, graph {use_graph}")
    for _ in range(2):
        ga.levels_rk3_step([lv for lv, _, _ in together], [dt for _, _, dt in together], p, precision=precision,
                           bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW, use_graph=use_graph)
    apart = build()
    for _ in range(2):
        for lv, _, dt in apart:
            lv.rk3_step(p, dt, precision=precision, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
    for (a, ga_, _), (b, gb, _) in zip(together, apart):
        a.download()
        b.download()
        for x, y in zip(ga_, gb):
            assert np.array_equal(x.array, y.array)
        a.close()
        b.close()
