"""Matter coupling of the device-resident RK3 step.

sspRK3Step hands computeRHS the matter sources rho, S_i, S_ij, S of each stage's state
(src/main.cpp:295-301). On the device a level holds borrowed device pointers to those arrays
per block (gr_level_set_sources); gr_level_rk3_stage runs one stage so a caller can recompute
the sources between stages. Checked against the oracle's RK3 step with the same sources in
every stage, and stage by stage against the whole step."""
import sys

import numpy as np
import pytest

from golden_cases import sources
from oracle.oracle import Geom, group_rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def device_sources(g, seed):
    from hipmem import DeviceArray

    host = sources(g, seed)
    dev = [DeviceArray(a) for a in host]
    return host, dev


def bl_block(oracle, n=(20, 16, 12), lo=(-2.5, -2.0, -1.5), hi=(2.5, 2.0, 1.5)):
    g = Geom(n, 4, lo, hi)
    return g, oracle.init_bl(g, [1.0], [[0.05, -0.02, 0.01]])


@pytest.mark.parametrize("variant", [0, 2])
def test_rk3_step_with_sources(engine, oracle, variant):
    """Both halo stagings of the FP64 stage kernel (registers; direct loads into LDS, whose
    field barriers count the source loads among the operations in flight)."""
    import granite_amd as ga

    g, u = bl_block(oracle)
    host, dev = device_sources(g, 7)
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    blk.array[...] = u
    lvl = ga.Level([blk], engine)
    lvl.upload()
    lvl.set_sources(0, *[t.data_ptr() for t in dev])
    lvl.set_tuning(ga.TUNE_STAGE, 0, variant)
    p = ga.CCZ4Params()
    ref = u.copy()
    for step in range(2):
        say(f"rk3 with sources, staging variant {variant}, step {step}")
        lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=step == 1)
        oracle.fill_ghosts(g, ref, 2)
        oracle.rk3_step(g, ref, 0.0625, 0, sources=host)
    lvl.download()
    e = group_rel(blk.array, ref)
    assert e.max() < TOL, e
    # the sources changed the result (vacuum differs)
    vac = u.copy()
    for _ in range(2):
        oracle.fill_ghosts(g, vac, 2)
        oracle.rk3_step(g, vac, 0.0625, 0)
    assert group_rel(blk.array, vac).max() > 1e-6
    lvl.close()


def test_stage_by_stage_equals_step_and_clearing_sources(engine, oracle):
    import granite_amd as ga

    g, u = bl_block(oracle, n=(16, 12, 10), lo=(-2.0, -1.5, -1.25), hi=(2.0, 1.5, 1.25))
    _, dev = device_sources(g, 8)
    p = ga.CCZ4Params()
    outs = []
    for by_stage in (False, True):
        blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
        blk.array[...] = u
        lvl = ga.Level([blk], engine)
        lvl.upload()
        lvl.set_sources(0, *[t.data_ptr() for t in dev])
        say(f"sources, by stage {by_stage}")
        if by_stage:
            lvl.rk3_stage(p, 0.0625, 1, pre_fill=ga.BC_OUTFLOW)
            lvl.rk3_stage(p, 0.0625, 2)
            lvl.rk3_stage(p, 0.0625, 3)
        else:
            lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW)
        lvl.download()
        outs.append(blk.array.copy())
        # cleared sources give the vacuum step
        blk.array[...] = u
        lvl.upload()
        lvl.set_sources(0)
        lvl.rk3_step(p, 0.0625, pre_fill=ga.BC_OUTFLOW, use_graph=True)
        lvl.download()
        outs.append(blk.array.copy())
        lvl.close()
    assert np.array_equal(outs[0], outs[2])
    ref = u.copy()
    oracle.fill_ghosts(g, ref, 2)
    oracle.rk3_step(g, ref, 0.0625, 0)
    assert group_rel(outs[1], ref).max() < TOL
    assert np.array_equal(outs[1], outs[3])


def test_batched_level_with_sources_on_one_block(engine, oracle):
    """A batch whose second block carries sources and whose first is vacuum: each block
    bit-identical to the same block stepped alone."""
    import granite_amd as ga

    ga_blocks = [ga.GridBlock(0, 0, (12, 10, 8), (-1.5, -1.25, -1.0), (1.5, 1.25, 1.0), 4),
                 ga.GridBlock(1, 0, (10, 12, 9), (0.25, -1.5, -1.125), (2.75, 1.5, 1.125), 4)]
    g1 = Geom(ga_blocks[1].numCells(), 4, ga_blocks[1].lowerCorner(), ga_blocks[1].upperCorner())
    _, dev = device_sources(g1, 9)
    p = ga.CCZ4Params()
    masses, pos = [1.0], [[0.1, 0.0, 0.0]]
    lvl = ga.Level(ga_blocks, engine)
    lvl.init_brill_lindquist(masses, pos)
    lvl.set_sources(1, *[t.data_ptr() for t in dev])
    say("batched level, sources on block 1")
    lvl.rk3_step(p, 0.03125, pre_fill=ga.BC_OUTFLOW, use_graph=True)
    lvl.download()
    for i, b in enumerate(ga_blocks):
        alone = ga.GridBlock(0, 0, b.numCells(), b.lowerCorner(), b.upperCorner(), 4)
        one = ga.Level([alone], engine)
        one.init_brill_lindquist(masses, pos)
        if i == 1:
            one.set_sources(0, *[t.data_ptr() for t in dev])
        one.rk3_step(p, 0.03125, pre_fill=ga.BC_OUTFLOW)
        one.download()
        assert np.array_equal(alone.array, b.array), i
        one.close()
    lvl.close()
