"""Mixed-precision form (FP32 stencils and algebra; FP64 state, combine, metric inverse,
near-puncture tiles and constraint monitor) against the FP64 oracle and the FP64 form.

The mixed form is not a parity form: its RHS is reported and bounded against the FP64
oracle (relative to each field's tensor-group magnitude), and its evolution is held to the
Hamiltonian-constraint criterion of the mixed form (L2 within 1% of the FP64 run)."""
import sys

import numpy as np
import pytest

from oracle.oracle import Geom, group_rel

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def level_from(engine, g, u):
    import granite_amd as ga

    b = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    b.array[...] = u
    lvl = ga.Level([b], engine)
    lvl.upload()
    return lvl, b


@pytest.mark.parametrize("noise", [0.0, 1e-8])
def test_mixed_rhs_error_against_oracle(engine, oracle, noise):
    import granite_amd as ga

    # C2/C3 spacing around the puncture: FP64 tiles near it, FP32 tiles elsewhere
    g = Geom((48, 40, 32), 4, (-6.0, -5.0, -4.0), (6.0, 5.0, 4.0))
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    oracle.rk3_step(g, u, 0.0625, 0)
    if noise:
        oracle.add_noise(g, u, 31, noise, [1, 2, 3, 4, 5, 6, 13])
    lvl, blk = level_from(engine, g, u)
    p = ga.CCZ4Params()
    ref = oracle.rhs(g, u)
    I = g.interior()
    say(f"mixed rhs noise={noise}")
    mixed = lvl.rhs(0, p, ga.PREC_MIXED)
    fp64 = lvl.rhs(0, p, ga.PREC_FP64)
    e64 = group_rel(fp64[I], ref[I])
    e32 = group_rel(mixed[I], ref[I])
    say("fp64 " + np.array2string(e64, precision=1))
    say("mixed " + np.array2string(e32, precision=1))
    assert e64.max() < 1e-12
    # measured: <= 4e-5 (K), <= 5e-6 (Theta), ~1e-6 (A~), <= 6e-7 elsewhere
    assert e32.max() < 1e-4
    lvl.close()


def test_mixed_policy_extremes(engine, oracle):
    import granite_amd as ga

    g = Geom((24, 16, 16), 4, (-3.0, -2.0, -2.0), (3.0, 2.0, 2.0))
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    oracle.rk3_step(g, u, 0.0625, 0)  # every tensor group non-trivial
    lvl, blk = level_from(engine, g, u)
    p = ga.CCZ4Params()
    say("mixed policy: every tile FP64")
    lvl.set_mixed_policy(1e9)
    a = lvl.rhs(0, p, ga.PREC_MIXED)
    b = lvl.rhs(0, p, ga.PREC_FP64)
    assert np.array_equal(a, b)
    say("mixed policy: default (FP64 below chi = 0.2) and every tile FP32")
    I = g.interior()
    lvl.set_mixed_policy(0.2)
    d = lvl.rhs(0, p, ga.PREC_MIXED)
    lvl.set_mixed_policy(-1.0)
    c = lvl.rhs(0, p, ga.PREC_MIXED)
    e_def = group_rel(d[I], b[I])
    e_all = group_rel(c[I], b[I])
    say("default " + np.array2string(e_def, precision=1))
    say("all fp32 " + np.array2string(e_all, precision=1))
    # the FP64 tiles near the puncture can only reduce the error
    assert e_def.max() < 1e-4
    assert e_all.max() >= 0.99 * e_def.max()
    lvl.close()


def test_mixed_evolution_matches_fp64_constraint(engine):
    """BL puncture, 64^3 on [-8, 8]^3 (dx = 0.25 as C2), 40 RK3 steps (t = 2.5 M): the
    Hamiltonian-constraint L2 of the mixed run is within 1% of the FP64 run, and the states
    stay close."""
    import granite_amd as ga

    n = 64
    out = {}
    for prec in (ga.PREC_FP64, ga.PREC_MIXED):
        blk = ga.GridBlock(0, 0, (n, n, n), (-8.0,) * 3, (8.0,) * 3, 4)
        lvl = ga.Level([blk], engine)
        lvl.init_brill_lindquist([1.0], [[0.0, 0.0, 0.0]])
        say(f"evolve precision={prec}")
        for _ in range(40):
            lvl.rk3_step(ga.CCZ4Params(), 0.0625, precision=prec, pre_fill=ga.BC_OUTFLOW, use_graph=True)
        lvl.sync()
        assert lvl.nonfinite() == 0
        s, c = lvl.constraints_l2()
        lvl.download()
        out[prec] = (np.sqrt(s / c), blk.array.copy())
        lvl.close()
    h64, u64 = out[ga.PREC_FP64]
    h32, u32 = out[ga.PREC_MIXED]
    say(f"H L2 fp64 {h64:.6e} mixed {h32:.6e} rel {abs(h32 - h64) / h64:.2e}")
    assert abs(h32 - h64) <= 0.01 * h64
    assert group_rel(u32, u64).max() < 1e-4


def test_c3_full_size_rhs_error(engine, oracle):
    """192^3 puncture block on [-16,16]^3 with N(0, 1e-8^2) on the conformal metric and K:
    the mixed RHS against the FP64 form on every cell, and both against the oracle on three
    sub-boxes (the RHS of a cell depends only on its stencil, so a sub-box with its own
    ghosts cut from the full block reproduces it)."""
    import granite_amd as ga

    g = Geom((192, 192, 192), 4, (-16.0,) * 3, (16.0,) * 3)
    blk = ga.GridBlock(0, 0, g.n, g.lo, g.hi, g.ng)
    lvl = ga.Level([blk], engine)
    lvl.init_brill_lindquist([1.0], [[0.0, 0.0, 0.0]])
    lvl.add_noise(20240611, 1e-8, [1, 2, 3, 4, 5, 6, 13])
    lvl.download()
    p = ga.CCZ4Params()
    say("C3 192^3: mixed and FP64 RHS")
    mixed = lvl.rhs(0, p, ga.PREC_MIXED)
    fp64 = lvl.rhs(0, p, ga.PREC_FP64)
    I = g.interior()
    e = group_rel(mixed[I], fp64[I])
    say("C3 mixed vs FP64 (all cells) " + np.array2string(e, precision=1))
    assert e.max() < 1e-4
    # normwise scales over the whole block: per field (parity) and per tensor group (mixed)
    fscale = np.array([max(float(np.max(np.abs(fp64[I][v]))), 1e-300) for v in range(22)])
    scale = fscale.copy()
    for grp in ([0], range(1, 7), range(7, 13), [13], range(14, 17), [17], [18], range(19, 22)):
        scale[list(grp)] = fscale[list(grp)].max()
    u = blk.array
    ng = g.ng
    for name, c0 in (("puncture", (80, 80, 80)), ("middle", (120, 96, 40)), ("far", (8, 150, 160))):
        n = 32
        sl = (slice(None),) + tuple(slice(c0[2 - a], c0[2 - a] + n + 2 * ng) for a in range(3))
        lo = [g.lo[d] + c0[d] * g.dx[d] for d in range(3)]
        sub = Geom((n, n, n), ng, lo, [lo[d] + n * g.dx[d] for d in range(3)])
        sub.dx = g.dx.copy()  # the full block's spacing, bit for bit (the RHS reads only dx)
        su = np.ascontiguousarray(u[sl])
        ref = oracle.rhs(sub, su)
        J = (slice(None),) + tuple(slice(c0[2 - a] + ng, c0[2 - a] + ng + n) for a in range(3))
        # normwise over the block, as the parity bar is defined (far from the puncture the
        # RHS is small: the metric's is the dissipation of the 1e-8 noise)
        d64 = np.max(np.abs(fp64[J] - ref[sub.interior()]).reshape(22, -1), axis=1)
        d32 = np.max(np.abs(mixed[J] - ref[sub.interior()]).reshape(22, -1), axis=1)
        e64 = d64 / fscale
        e32 = d32 / scale
        say(f"C3 sub-box {name}: fp64 {e64.max():.1e} mixed {e32.max():.1e}")
        assert e64.max() < 1e-12
        assert e32.max() < 1e-4
    lvl.close()
