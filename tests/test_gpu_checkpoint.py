"""Checkpoint staging (gr_checkpoint) on the GPU: a snapshot taken between hierarchy steps
holds exactly the state of every block at that point (the data HDF5Writer::writeCheckpoint
stores, src/io/hdf5_io.cpp:166-232), although the evolution goes on while it is copied out."""
import sys

import numpy as np
import pytest

from granite_amd.hierarchy import Sphere, tracking_levels

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def small_hierarchy(ga, engine):
    sph = [Sphere((-1.5, 0.0, 0.0), 1.0, 2), Sphere((1.5, 0.0, 0.0), 1.0, 2)]
    geo = tracking_levels([16] * 3, [-8.0] * 3, [8.0] * 3, sph, 3)
    levels = [ga.Level([ga.GridBlock(i, l, n, lo, hi, 4) for i, (n, lo, hi) in enumerate(lv)], engine)
              for l, lv in enumerate(geo)]
    H = ga.Hierarchy(levels, engine)
    levels[0].init_brill_lindquist([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]])
    for l in range(1, len(levels)):
        H.prolongate(l)
    H.fill_ghosts(0)
    return levels, H


def states(levels):
    out = []
    for L in levels:
        L.download()
        out.extend(b.array.copy() for b in L.blocks)
    return out


def test_snapshot_is_the_state_at_begin(engine):
    import granite_amd as ga

    say("checkpoint: 3-level hierarchy, snapshots after steps 1 and 3")
    levels, H = small_hierarchy(ga, engine)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    ck = ga.Checkpoint(levels, engine)
    assert len(ck) == sum(len(L.blocks) for L in levels)
    assert ck.done()  # nothing pending yet
    for step, take in ((1, True), (2, False), (3, True)):
        H.step(p, 0.5, ga.PREC_FP64, use_graph=True)
        if not take:
            continue
        ck.begin()
        want = states(levels)  # the download is ordered after the snapshot on the engine stream
        for _ in range(2):  # the evolution goes on while the copy runs
            H.step(p, 0.5, ga.PREC_FP64, use_graph=True)
        ck.wait()
        assert ck.done()
        for i, w in enumerate(want):
            got = ck.block(i)
            assert got.shape == w.shape
            assert np.array_equal(got, w), f"step {step} block {i}"
        moved = states(levels)
        assert any(not np.array_equal(m, w) for m, w in zip(moved, want))  # later steps changed the state
        say(f"step {step}: {len(want)} blocks bit-identical to the state at begin")
    with pytest.raises(ga.EngineError):
        ck.block(len(want))
    ck.close()
    H.close()
    for L in levels:
        L.close()
