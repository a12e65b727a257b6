"""Berger-Oliger hierarchy on the device (gr_hierarchy) against the oracle's restatement of
AMRHierarchy (itself pinned bitwise to the reference's subcycle on the CPU).

Transfers (prolongation, restriction, fillGhostZones with uncovered faces) are exact
copies and averages computed in the same order: bit-identical. Whole coarse steps go through
the FP64 stage kernels: within 1e-12 of the oracle per field (tensor-group scale)."""
import json
import os
import sys

import numpy as np
import pytest

from granite_amd.hierarchy import PUNCTURES, Sphere, binary_puncture_levels, tracking_levels
from oracle.hierarchy import OracleHierarchy, build_levels
from oracle.oracle import DEFAULT_PARAMS, Geom, group_rel

pytestmark = pytest.mark.gpu

TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def sphere_geometry(case):
    if case == "merged":
        sph = [Sphere((-1.5, 0.0, 0.0), 1.0, 2), Sphere((1.5, 0.0, 0.0), 1.0, 2)]
        return tracking_levels([16] * 3, [-8.0] * 3, [8.0] * 3, sph, 3), sph
    if case == "split":
        sph = [Sphere((-5.0, 0.0, 0.0), 1.0, 2), Sphere((5.0, 0.0, 0.0), 1.0, 2)]
        return tracking_levels([16] * 3, [-8.0] * 3, [8.0] * 3, sph, 3), sph
    # hand-made: two parent blocks, a child sticking out of both on x-high and y-low
    geo = [
        [([16, 16, 16], [-8.0] * 3, [8.0] * 3)],
        [([12, 16, 16], [-5.0, -4.0, -4.0], [1.0, 4.0, 4.0]), ([8, 12, 12], [1.0, -3.0, -3.0], [5.0, 3.0, 3.0])],
        [([32, 12, 12], [-2.0, -4.5, -1.5], [6.0, -1.5, 1.5])],
    ]
    sph = [Sphere((-1.0, 0.0, 0.0), 1.0, 2), Sphere((1.2, 0.0, 0.0), 1.0, 2)]
    return geo, sph


def device_hierarchy(ga, engine, oracle_levels):
    levels = []
    blocks = []
    for lv in oracle_levels:
        bl = [ga.GridBlock(i, 0, g.n, g.lo, g.hi, g.ng) for i, (g, _) in enumerate(lv)]
        for b, (_, u) in zip(bl, lv):
            b.array[...] = u
        L = ga.Level(bl, engine)
        L.upload()
        levels.append(L)
        blocks.append(bl)
    return levels, blocks, ga.Hierarchy(levels, engine)


def compare(levels, blocks, oracle_levels, exact, what):
    for l, (L, bl, lv) in enumerate(zip(levels, blocks, oracle_levels)):
        L.download()
        for i, (b, (_, u)) in enumerate(zip(bl, lv)):
            if exact:
                assert np.array_equal(b.array, u), f"{what}: level {l} block {i}"
            else:
                e = group_rel(b.array, u)
                assert np.max(e) < TOL, f"{what}: level {l} block {i} worst field {int(np.argmax(e))} err {np.max(e):.3e}"


@pytest.mark.parametrize("case", ["merged", "split", "uncovered"])
def test_transfers_bitwise(engine, oracle, case):
    import granite_amd as ga

    geo, sph = sphere_geometry(case)
    masses = [0.5] * len(sph)
    pos = [list(s.center) for s in sph]
    ol = build_levels(oracle, geo, masses=masses, positions=pos)
    # give every level distinct data so that each transfer shows
    for l, lv in enumerate(ol):
        for g, u in lv:
            oracle.add_noise(g, u, 11 + l, 1e-3, range(22))
    levels, blocks, H = device_hierarchy(ga, engine, ol)
    OH = OracleHierarchy(oracle, ol)
    nl = len(ol)
    say(f"transfers {case}: restrict")
    for l in range(nl - 1, 0, -1):
        H.restrict(l)
        OH.restrict(l)
    compare(levels, blocks, ol, True, "restrict")
    say(f"transfers {case}: fill ghosts")
    for l in range(nl):
        H.fill_ghosts(l)
        OH.fill_ghosts(l)
    compare(levels, blocks, ol, True, "fillGhostZones")
    say(f"transfers {case}: prolongate")
    for l in range(1, nl):
        H.prolongate(l)
        OH.prolongate(l)
    compare(levels, blocks, ol, True, "prolongate")
    H.close()
    for L in levels:
        L.close()


@pytest.mark.parametrize("case,use_graph,variant", [("merged", False, 0), ("merged", True, 0), ("split", True, 0),
                                                    ("uncovered", True, 0), ("merged", True, 2)])
def test_coarse_steps_against_oracle(engine, oracle, case, use_graph, variant, tmp_path):
    """Two Berger-Oliger coarse steps of the device hierarchy against the oracle's, every level
    within the FP64 bar; variant 2 takes the FP64 stage kernel's direct-to-LDS staging for
    every level from a tuning table."""
    import granite_amd as ga

    ga.tune_clear()
    if variant:
        table = os.path.join(str(tmp_path), "tune.txt")
        with open(table, "w") as f:
            f.write("".join(f"0 {c} 0 {variant}\n" for c in range(4)))
        ga.tune_load(table)
    try:
        _coarse_steps_against_oracle(ga, engine, oracle, case, use_graph)
    finally:
        ga.tune_clear()


def _coarse_steps_against_oracle(ga, engine, oracle, case, use_graph):
    geo, sph = sphere_geometry(case)
    masses = [0.5] * len(sph)
    pos = [list(s.center) for s in sph]
    ol = build_levels(oracle, geo, masses=masses, positions=pos)
    levels, blocks, H = device_hierarchy(ga, engine, ol)
    OH = OracleHierarchy(oracle, ol)
    prm = dict(DEFAULT_PARAMS)
    p = ga.CCZ4Params(**prm)
    dt0 = 0.25 * ol[0][0][0].dx[0]
    H.fill_ghosts(0)
    OH.fill_ghosts(0)
    for s in range(2):
        say(f"hierarchy {case} graph={use_graph} coarse step {s}")
        H.step(p, dt0, ga.PREC_FP64, use_graph=use_graph)
        OH.step(dt0, prm)
    H.sync()
    compare(levels, blocks, ol, False, f"{case} after 2 coarse steps")
    sh, sm, n = H.constraints_l2()
    oh, om, on = OH.constraints_l2()
    assert n == on
    assert abs(sh - oh) <= 1e-9 * oh and abs(sm - om) <= 1e-9 * om
    H.close()
    for L in levels:
        L.close()


def test_mixed_coarse_steps_close_to_fp64(engine, oracle):
    import granite_amd as ga

    geo, sph = sphere_geometry("merged")
    masses = [0.5] * len(sph)
    pos = [list(s.center) for s in sph]
    ol = build_levels(oracle, geo, masses=masses, positions=pos)
    runs = []
    for prec in (ga.PREC_FP64, ga.PREC_MIXED):
        levels, blocks, H = device_hierarchy(ga, engine, ol)
        p = ga.CCZ4Params()
        H.fill_ghosts(0)
        say(f"mixed vs fp64 hierarchy, precision {prec}")
        for _ in range(4):
            H.step(p, 0.25, prec, use_graph=True)
        H.sync()
        for L in levels:
            L.download()
        runs.append(([[b.array.copy() for b in bl] for bl in blocks], H.constraints_l2()))
        H.close()
        for L in levels:
            L.close()
    (a, ca), (b, cb) = runs
    for la, lb in zip(a, b):
        for ua, ub in zip(la, lb):
            assert np.max(group_rel(ub, ua)) < 1e-4
    assert abs(np.sqrt(cb[0] / cb[2]) - np.sqrt(ca[0] / ca[2])) <= 0.01 * np.sqrt(ca[0] / ca[2])


def test_rejects_wrong_ratio(engine):
    import granite_amd as ga

    a = ga.Level([ga.GridBlock(0, 0, [8] * 3, [0.0] * 3, [8.0] * 3, 4)], engine)
    b = ga.Level([ga.GridBlock(0, 1, [8] * 3, [0.0] * 3, [8.0 / 3] * 3, 4)], engine)
    with pytest.raises(ga.EngineError):
        ga.Hierarchy([a, b], engine)
    a.close()
    b.close()


def test_binary_hierarchy_t10(engine):
    """The 4-level binary-puncture hierarchy to t = 10 M: FP64 against the oracle's run
    (tests/golden/binary_hierarchy_t10.json) and mixed precision against FP64."""
    import granite_amd as ga

    path = os.path.join(HERE, "golden", "binary_hierarchy_t10.json")
    if not os.path.exists(path):
        pytest.skip("golden run not generated")
    gold = json.load(open(path))
    geo = binary_puncture_levels()
    assert json.loads(json.dumps(geo)) == gold["geometry"]
    masses, pos = PUNCTURES
    p = ga.CCZ4Params(**gold["params"])
    out = {}
    for prec in (ga.PREC_FP64, ga.PREC_MIXED):
        levels = [ga.Level([ga.GridBlock(0, l, n, lo, hi, 4) for n, lo, hi in lv], engine) for l, lv in enumerate(geo)]
        H = ga.Hierarchy(levels, engine)
        levels[0].init_brill_lindquist(masses, pos)
        for l in range(1, len(levels)):
            H.prolongate(l)
        H.fill_ghosts(0)
        tel = {}
        for step in range(1, gold["steps"] + 1):
            H.step(p, gold["dt0"], prec, use_graph=True)
            if step in (1, 10, 20, 30):
                tel[step] = H.constraints_l2()
        say(f"binary hierarchy precision {prec}: H_L2(t=10) = {np.sqrt(tel[30][0] / tel[30][2]):.6e}")
        out[prec] = (tel, levels, H)
    for rec in gold["telemetry"]:
        sh, sm, n = out[ga.PREC_FP64][0][rec["step"]]
        assert n == rec["count"]
        assert abs(sh - rec["ham_sumsq"]) <= 1e-8 * rec["ham_sumsq"], rec
        assert abs(sm - rec["mom_sumsq"]) <= 1e-8 * rec["mom_sumsq"], rec
    tel64, telmx = out[ga.PREC_FP64][0][30], out[ga.PREC_MIXED][0][30]
    h64, hmx = np.sqrt(tel64[0] / tel64[2]), np.sqrt(telmx[0] / telmx[2])
    assert abs(hmx - h64) <= 0.01 * h64
    for prec in out:
        tel, levels, H = out[prec]
        H.close()
        for L in levels:
            L.close()
