"""The oracle's Berger-Oliger orchestration and level geometry, pinned to the reference.

Against the compiled reference (oracle/_ref, skipped where the reference sources were not
present at build time):
  * tracking_levels() reproduces AMRHierarchy::initialize/regrid block for block, bitwise;
  * OracleHierarchy.step reproduces AMRHierarchy::subcycle run by the reference itself (with
    the oracle's SSP-RK3 step as its per-level callback), bitwise on every level.
Without it:
  * the masked outflow fill against the plain outflow fill and its own slab structure;
  * skipping the overwritten work gives bit-identical states.
"""
import numpy as np
import pytest

from granite_amd.hierarchy import Sphere, binary_puncture_levels, cell_steps_per_coarse_step, puncture_spheres, \
    tracking_levels
from oracle.hierarchy import OracleHierarchy, ReferenceHierarchy, build_levels, uncovered_faces
from oracle.oracle import DEFAULT_PARAMS, Geom, Oracle, Reference

GEOMETRY_CASES = [
    # ncells, extent, spheres (center, radius, min_level), max_levels
    (16, 8.0, [((-1.5, 0, 0), 1.0, 2), ((1.5, 0, 0), 1.0, 2)], 3),           # merged boxes
    (16, 8.0, [((-5.0, 0, 0), 1.0, 2), ((5.0, 0, 0), 1.0, 2)], 3),           # two blocks per level
    (20, 10.0, [((9.0, -9.5, 0.3), 0.5, 3)], 4),                              # clamped at the domain
    (12, 6.0, [((-2, 0, 0), 1.0, 2), ((0, 0, 0), 1.0, 1), ((2.2, 0.5, 0), 1.0, 2)], 3),  # chained merges
    (24, 64.0, [((-3.0, 0, 0), 1.0, 3), ((3.0, 0, 0), 1.0, 3)], 4),          # the binary at low resolution
]


def _ref_or_skip():
    if not Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    return Reference()


@pytest.mark.parametrize("case", range(len(GEOMETRY_CASES)))
def test_geometry_matches_reference_regrid(case):
    ref = _ref_or_skip()
    nc, ext, sph, maxl = GEOMETRY_CASES[case]
    spheres = [Sphere(tuple(float(c) for c in ce), r, m) for ce, r, m in sph]
    R = ReferenceHierarchy(ref, [nc] * 3, [-ext] * 3, [ext] * 3, spheres, maxl)
    try:
        got = R.blocks()
    finally:
        R.close()
    ours = tracking_levels([nc] * 3, [-ext] * 3, [ext] * 3, spheres, maxl)
    flat = [(l, n, lo, hi) for l, lv in enumerate(ours) for n, lo, hi in lv]
    assert len(flat) == len(got)
    for (l, n, lo, hi), (rl, rn, rlo, rhi, rdx) in zip(flat, got):
        assert l == rl and list(n) == rn
        assert lo == rlo and hi == rhi  # bitwise
        assert list(Geom(n, 4, lo, hi).dx) == rdx


def test_binary_puncture_geometry():
    """The 4-level binary hierarchy: one merged box per level, ratio 2, nested."""
    lv = binary_puncture_levels()
    assert [len(x) for x in lv] == [1, 1, 1, 1]
    dx = [Geom(x[0][0], 4, x[0][1], x[0][2]).dx[0] for x in lv]
    for l in range(1, 4):
        assert abs(dx[l] * 2 - dx[l - 1]) <= 1e-15 * dx[l - 1]
        n, lo, hi = lv[l][0]
        pn, plo, phi = lv[l - 1][0]
        assert all(plo[d] <= lo[d] and hi[d] <= phi[d] for d in range(3))
        # y and z: a single sphere-centred box of 96 fine cells; x: the union of two
        assert n[1] == n[2] == 96 and n[0] > 96
    assert [x[0][0] for x in lv] == [[96] * 3, [106, 96, 96], [114, 96, 96], [132, 96, 96]]
    assert cell_steps_per_coarse_step(lv) == 96 ** 3 + 2 * 106 * 96 * 96 + 4 * 114 * 96 * 96 + 8 * 132 * 96 * 96


def test_binary_puncture_geometry_matches_reference():
    ref = _ref_or_skip()
    spheres = puncture_spheres([0.5, 0.5], [[-3.0, 0, 0], [3.0, 0, 0]], 4)
    R = ReferenceHierarchy(ref, [96] * 3, [-64.0] * 3, [64.0] * 3, spheres, 4)
    try:
        got = R.blocks()
    finally:
        R.close()
    ours = [(l, n, lo, hi) for l, lv in enumerate(binary_puncture_levels()) for n, lo, hi in lv]
    assert [(l, n, lo, hi) for l, n, lo, hi, _ in got] == ours


def _tiny(oracle, case):
    nc, ext, sph, maxl = GEOMETRY_CASES[case]
    spheres = [Sphere(tuple(float(c) for c in ce), r, m) for ce, r, m in sph]
    geo = tracking_levels([nc] * 3, [-ext] * 3, [ext] * 3, spheres, maxl)
    masses = [0.5] * len(sph)
    pos = [list(s.center) for s in spheres]
    return spheres, geo, masses, pos


@pytest.mark.parametrize("case", [0, 1])
def test_subcycle_matches_reference(oracle, case):
    """Two coarse steps of the reference's own subcycle vs the oracle's orchestration."""
    ref = _ref_or_skip()
    nc, ext, sph, maxl = GEOMETRY_CASES[case]
    spheres, geo, masses, pos = _tiny(oracle, case)
    levels = build_levels(oracle, geo, masses=masses, positions=pos)
    flat = [u for lv in levels for _, u in lv]
    R = ReferenceHierarchy(ref, [nc] * 3, [-ext] * 3, [ext] * 3, spheres, maxl)
    try:
        assert len(R.blocks()) == len(flat)
        for i, u in enumerate(flat):
            R.view(i)[...] = u
        prm = dict(DEFAULT_PARAMS)

        def evolve(level, n, ng, lo, dx, u, dt):
            g = Geom(n, ng, lo, [lo[d] + n[d] * dx[d] for d in range(3)])
            assert list(g.dx) == list(dx)
            oracle.rk3_step(g, u, dt, 0, **prm)

        H = OracleHierarchy(oracle, levels)
        R.fill_ghosts(0)
        H.fill_ghosts(0)
        dt0 = 0.25 * levels[0][0][0].dx[0]
        for _ in range(2):
            R.step(dt0, evolve)
            H.step(dt0, prm)
            for i, u in enumerate(flat):
                assert np.array_equal(R.view(i), u), f"block {i}"
    finally:
        R.close()


@pytest.mark.parametrize("case", [0, 1, 3])
def test_skipping_overwritten_work_is_exact(oracle, case):
    spheres, geo, masses, pos = _tiny(oracle, case)
    a = build_levels(oracle, geo, masses=masses, positions=pos)
    b = build_levels(oracle, geo, masses=masses, positions=pos)
    Ha, Hb = OracleHierarchy(oracle, a), OracleHierarchy(oracle, b)
    prm = dict(DEFAULT_PARAMS)
    dt0 = 0.25 * a[0][0][0].dx[0]
    for _ in range(2):
        Ha.step(dt0, prm)
        Hb.step(dt0, prm, skip_overwritten=True)
    for la, lb in zip(a, b):
        for (_, ua), (_, ub) in zip(la, lb):
            assert np.array_equal(ua, ub)


def test_outflow_faces_all_equals_plain_outflow(oracle):
    g = Geom((6, 5, 7), 4, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(3)
    u = rng.standard_normal(g.shape)
    a = oracle.fill_ghosts(g, u.copy(), 2)
    b = oracle.fill_outflow_faces(g, u.copy(), 0x3F)
    assert np.array_equal(a, b)
    assert np.array_equal(oracle.fill_outflow_faces(g, u.copy(), 0), u)


@pytest.mark.parametrize("face", range(6))
def test_outflow_single_face_slab(oracle, face):
    g = Geom((6, 5, 7), 4, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(4 + face)
    u = rng.standard_normal(g.shape)
    out = oracle.fill_outflow_faces(g, u.copy(), 1 << face)
    d, side = face // 2, face % 2
    ax = 3 - d  # array axis of spatial axis d
    ng, n = g.ng, g.n[d]
    sl_ghost = [slice(None)] * 4
    sl_ghost[ax] = slice(0, ng) if side == 0 else slice(ng + n, n + 2 * ng)
    src = [slice(None)] * 4
    src[ax] = slice(ng, ng + 1) if side == 0 else slice(ng + n - 1, ng + n)
    expect = u.copy()
    expect[tuple(sl_ghost)] = np.broadcast_to(u[tuple(src)], expect[tuple(sl_ghost)].shape)
    assert np.array_equal(out, expect)


def test_uncovered_faces():
    parent = Geom((8, 8, 8), 4, (0, 0, 0), (8, 8, 8))
    inside = Geom((8, 8, 8), 4, (2, 2, 2), (6, 6, 6))
    assert uncovered_faces(inside, [parent]) == 0
    flush = Geom((8, 8, 8), 4, (0, 0, 0), (4, 4, 4))
    assert uncovered_faces(flush, [parent]) == 0
    sticking = Geom((12, 8, 8), 4, (4, 2, 2), (10, 6, 6))  # x-high beyond the parent
    assert uncovered_faces(sticking, [parent]) == 1 << 1
    assert uncovered_faces(sticking, [parent, Geom((8, 8, 8), 4, (8, 0, 0), (16, 8, 8))]) == 0
