"""Batched levels: all blocks of a level advance with one stage launch and one ghost-fill
launch per RK3 stage. Each block must come out exactly as when it is evolved alone
(bitwise: the same kernels compute the same cells), and small uneven blocks agree with
the oracle to the FP64 tolerance."""
import sys

import numpy as np
import pytest

from oracle.oracle import Geom, group_rel

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


def test_level_rejects_mixed_spacing(engine):
    import granite_amd as ga

    a = ga.GridBlock(0, 0, (8, 8, 8), (0, 0, 0), (1, 1, 1), 4)
    b = ga.GridBlock(1, 0, (8, 8, 8), (0, 0, 0), (2, 2, 2), 4)
    with pytest.raises(ga.EngineError) as ei:
        ga.Level([a, b], engine)
    assert ei.value.code == -2
