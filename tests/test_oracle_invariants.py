"""Checks that pin the oracle from mathematics alone (no reference run needed).

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed operand, a wrong ghost-fill order) fails it."""
import numpy as np
import pytest

from oracle.oracle import Geom, normwise_rel

V = dict(CHI=0, GXX=1, GXY=2, GXZ=3, GYY=4, GYZ=5, GZZ=6, AXX=7, K=13, GHX=14, THETA=17, ALPHA=18, BX=19)


def flat(oracle, n=(8, 8, 8), lo=(-1, -1, -1), hi=(1, 1, 1)):
    g = Geom(n, 4, lo, hi)
    return g, oracle.init_flat(g)


def test_flat_rhs_and_constraints_vanish(oracle):
    # tests/spacetime/test_ccz4_flat.cpp:44-63,106-149 (thresholds 1e-10)
    g, u = flat(oracle, (12, 12, 12), (-10, -10, -10), (10, 10, 10))
    I = g.interior()
    assert np.max(np.abs(oracle.rhs(g, u, ko_sigma=0.0)[I])) < 1e-10
    h, m = oracle.constraints(g, u)
    assert np.max(np.abs(h)) < 1e-10 and np.max(np.abs(m)) < 1e-10


def test_vacuum_equals_zero_sources(oracle):
    # tests/spacetime/test_ccz4_flat.cpp:65-104 (EXPECT_DOUBLE_EQ)
    g = Geom((6, 6, 6), 4, (-1, -1, -1), (1, 1, 1))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.1, 0.2, 0.3]]), 5, 1e-3, range(1, 22))
    N = g.stride
    z = (np.zeros(N), np.zeros(3 * N), np.zeros(6 * N), np.zeros(N))
    assert np.array_equal(oracle.rhs(g, u), oracle.rhs(g, u, sources=z))


def test_ko_null_space_and_alternating_mode(oracle):
    # KO annihilates polynomials of degree <= 5 per axis and maps (-1)^i to -64 sigma/(64h) (-1)^i
    g = Geom((8, 8, 8), 4, (0, 0, 0), (1, 1, 1))
    x, y, z = np.meshgrid(g.coords(0), g.coords(1), g.coords(2), indexing="ij")
    u = g.zeros()
    u[0] = (1 + x - 2 * x**2 + 0.5 * x**5 + y**3 - z**4 + x * y * z**2).transpose(2, 1, 0)
    r = oracle.ko(g, u, g.zeros(), 0.3)
    assert np.max(np.abs(r[g.interior()][0])) < 1e-9
    i = np.arange(g.nt[0])
    u2 = g.zeros()
    u2[5] = np.broadcast_to(((-1.0) ** i)[None, None, :], u2[5].shape)
    r2 = oracle.ko(g, u2, g.zeros(), 0.3)
    expect = -0.3 / (64 * g.dx[0]) * 64 * u2[5]
    assert np.allclose(r2[g.interior()][5], expect[g.interior()[1:]], rtol=1e-13, atol=0)


def _lapse_rhs(oracle, g, alpha, beta, chi_center):
    u = oracle.init_flat(g)
    u[V["ALPHA"]] = alpha
    for d in range(3):
        u[V["BX"] + d] = beta[d]
    r = oracle.rhs(g, u, ko_sigma=0.0, chi_blend_center=chi_center)
    return r[g.interior()][V["ALPHA"]]


def test_upwind_and_centred_advection_exact_on_quartics(oracle):
    # with K = 0 the lapse RHS is beta . grad(alpha); both the one-sided 5-point and the centred
    # 4th-order first derivatives are exact on quartics. chi = 1 with blend centre 0.05 gives
    # weight 1 (upwind only); a blend centre of 10 gives weight 0 (centred only).
    g = Geom((8, 8, 8), 4, (-1, -1, -1), (1, 1, 1))
    x, y, z = np.meshgrid(g.coords(0), g.coords(1), g.coords(2), indexing="ij")
    a = (2 + 0.3 * x - 0.2 * x**2 + 0.05 * x**4 + 0.1 * y**3 - 0.07 * z**4).transpose(2, 1, 0)
    ax = (0.3 - 0.4 * x + 0.2 * x**3).transpose(2, 1, 0)
    ay = (0.3 * y**2).transpose(2, 1, 0)
    az = (-0.28 * z**3).transpose(2, 1, 0)
    I = g.interior()[1:]
    for beta in [(0.5, -0.25, 0.125), (-0.5, 0.25, -0.75)]:
        expect = beta[0] * ax[I] + beta[1] * ay[I] + beta[2] * az[I]
        for cc in (0.05, 10.0):
            got = _lapse_rhs(oracle, g, a, [np.full_like(a, b) for b in beta], cc)
            assert np.allclose(got, expect - 2 * 0.0, rtol=0, atol=1e-12)


def test_second_derivatives_through_gamma_driver(oracle):
    # flat metric, beta^x = q(x,y,z) quadratic with cross terms: the Gamma-hat RHS at beta-only
    # data is T5 + T6 - Gh.d beta + advection = lap(beta) + 1/3 grad(div beta)
    g = Geom((8, 8, 8), 4, (-1, -1, -1), (1, 1, 1))
    x, y, z = np.meshgrid(g.coords(0), g.coords(1), g.coords(2), indexing="ij")
    bx = (0.01 * (x * y + 2 * x * z + 3 * x**2 - y * z)).transpose(2, 1, 0)
    u = oracle.init_flat(g)
    u[V["BX"]] = bx
    r = oracle.rhs(g, u, ko_sigma=0.0, chi_blend_center=10.0)
    I = g.interior()
    # lap(bx) = 0.06; grad(div beta) = grad(d bx/dx) = (0.06, 0.01, 0.02)
    # advection term beta.grad(Gh) = 0 (Gh = 0); -Gh.d beta = 0
    assert np.allclose(r[I][V["GHX"]], 0.06 + 0.06 / 3, atol=1e-12)
    assert np.allclose(r[I][V["GHX"] + 1], 0.01 / 3, atol=1e-12)
    assert np.allclose(r[I][V["GHX"] + 2], 0.02 / 3, atol=1e-12)


def test_rk3_linear_shift_decay_exact(oracle):
    # flat space, uniform beta, periodic ghosts: d beta/dt = -eta beta exactly, and one SSP-RK3
    # step multiplies beta by 1 - z + z^2/2 - z^3/6 with z = eta dt
    g = Geom((8, 8, 8), 4, (0, 0, 0), (1, 1, 1))
    u = oracle.init_flat(g)
    u[V["BX"]] = 0.3
    u[V["BX"] + 1] = -0.2
    dt, eta = 0.1, 2.0
    oracle.rk3_step(g, u, dt, bc_mode=1, eta=eta)
    zz = eta * dt
    amp = 1 - zz + zz**2 / 2 - zz**3 / 6
    assert np.allclose(u[V["BX"]], 0.3 * amp, rtol=1e-14, atol=0)
    assert np.allclose(u[V["BX"] + 1], -0.2 * amp, rtol=1e-14, atol=0)
    assert np.array_equal(u[V["CHI"]], np.ones_like(u[0]))


def _stage_sequence(oracle, g, u0, dt, floors_each_stage=True, project_each_stage=False):
    """One SSP-RK3 step composed from separately pinned pieces in the order of
    TimeIntegrator::sspRK3Step (src/main.cpp:673-689): per stage the RHS (orc_rhs, bitwise
    equal to the compiled reference), the combine d = a1 s1 + a2 s2 + a3 dt r on interior
    cells (src/main.cpp:306-335, written out here in numpy), the floors on every cell
    (src/main.cpp:588-608), after the last stage only the algebraic constraints
    (src/main.cpp:612-671), then the 1/r ghost fill (src/main.cpp:351-521). The keyword
    switches build the misplaced orders the check must reject."""
    I = g.interior()
    u = u0.copy()
    stage = g.zeros()
    coef = [(1.0, "u", 0.0, "u", 1.0), (0.75, "u", 0.25, "s", 0.25), (1.0 / 3.0, "u", 2.0 / 3.0, "s", 2.0 / 3.0)]
    for k, (a1, s1, a2, s2, a3) in enumerate(coef):
        src = u if k == 0 else stage
        r = oracle.rhs(g, src)
        dst = u if k == 2 else stage
        x1 = (u if s1 == "u" else stage)[I].copy()
        x2 = (u if s2 == "u" else stage)[I].copy()
        dst[I] = a1 * x1 + a2 * x2 + a3 * dt * r[I]
        if floors_each_stage or k == 2:
            oracle.floors(g, dst)
        if project_each_stage or k == 2:
            oracle.algebraic(g, dst)
        oracle.fill_ghosts(g, dst, 0)
    return u


def test_rk3_step_is_the_reference_stage_sequence(oracle):
    """orc_rk3_step equals, bit for bit, the reference's stage sequence composed from pinned
    pieces; on a state with chi above its cap and below its floor, a conformal metric off unit
    determinant and a traced A~, floors applied only at the end or the algebraic projection
    applied at every stage give a different step."""
    g = Geom((12, 10, 8), 4, (-1.5, -1.25, -1.0), (1.5, 1.25, 1.0))
    u0 = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.05, -0.02, 0.01]]), 5, 1e-3, range(22))
    u0[V["GXX"]:V["GXX"] + 6, 5:9, 5:10, 5:11] *= 1.2
    for a in (0, 3, 5):  # xx, yy, zz
        u0[V["AXX"] + a, 4:10, 4:11, 4:12] += 0.01
    u0[V["CHI"], 6, 6, 6:9] = 2.0  # above the 1.5 cap through every stage
    u0[V["CHI"], 8, 9, 6] = 5e-5
    dt = 0.05
    step = u0.copy()
    oracle.rk3_step(g, step, dt, bc_mode=0)
    assert np.isfinite(step).all()
    assert np.array_equal(step.view(np.uint64), _stage_sequence(oracle, g, u0, dt).view(np.uint64))
    for kw in (dict(floors_each_stage=False), dict(project_each_stage=True)):
        other = _stage_sequence(oracle, g, u0, dt, **kw)
        assert normwise_rel(other[g.interior()], step[g.interior()]).max() > 1e-10, kw


def test_rk3_with_uniform_sources_follows_the_reduced_ode(oracle):
    """Spatially uniform flat data with uniform matter sources rho, S (S_i = S_ij = 0), zero
    shift, periodic ghosts: every derivative and the Ricci tensor vanish, the metric and A~
    stay put, and the CCZ4 equations reduce to an ODE system for (chi, alpha, K, Theta):
        chi'   = 2/3 chi alpha K
        alpha' = -2 alpha K
        K'     = alpha K^2 / 3 + 4 pi alpha (rho + S) + kappa1 (1 - kappa2) alpha Theta
        Theta' = alpha (K^2 / 3 - 8 pi rho) - alpha kappa1 (2 + kappa2) Theta
    One SSP-RK3 step of the oracle (sources held in all three stages) must equal the same
    scheme applied to this ODE; dropping the sources from any stage, or a wrong stage weight,
    fails it."""
    g = Geom((6, 6, 6), 4, (0, 0, 0), (1, 1, 1))
    u = oracle.init_flat(g)
    u[V["K"]] = 0.05
    u[V["THETA"]] = -0.02
    u[V["ALPHA"]] = 0.9
    u[V["CHI"]] = 0.8
    N = g.stride
    rho0, S0 = 0.013, 0.004
    src = (np.full(N, rho0), np.zeros(3 * N), np.zeros(6 * N), np.full(N, S0))
    k1, k2, dt = 0.02, 0.1, 0.05

    def f(y):
        chi, al, K, Th = y
        return np.array([2 / 3 * chi * al * K, -2 * al * K,
                         al * K * K / 3 + 4 * np.pi * al * (rho0 + S0) + k1 * (1 - k2) * al * Th,
                         al * (K * K / 3 - 8 * np.pi * rho0) - al * k1 * (2 + k2) * Th])

    y0 = np.array([0.8, 0.9, 0.05, -0.02])
    y1 = y0 + dt * f(y0)
    y2 = 0.75 * y0 + 0.25 * y1 + 0.25 * dt * f(y1)
    y3 = y0 / 3 + 2 / 3 * y2 + 2 / 3 * dt * f(y2)
    oracle.rk3_step(g, u, dt, bc_mode=1, sources=src, kappa1=k1, kappa2=k2)
    for v, want in zip((V["CHI"], V["ALPHA"], V["K"], V["THETA"]), y3):
        assert np.allclose(u[v], want, rtol=1e-13, atol=1e-16), (v, u[v].ravel()[0], want)
    for v in (1, 4, 6):
        assert np.array_equal(u[v], np.ones_like(u[v]))
    # (round-off only: the 1e-30 regularisations in A~, the dissipation sums of uniform fields)
    assert np.max(np.abs(u[7:13])) < 1e-30 and np.max(np.abs(u[19:22])) < 1e-18
    # null sources are the vacuum step, bit for bit
    a = oracle.init_flat(g)
    a[V["K"]] = 0.05
    b = a.copy()
    oracle.rk3_step(g, a, dt, bc_mode=1)
    z = (np.zeros(N), np.zeros(3 * N), np.zeros(6 * N), np.zeros(N))
    oracle.rk3_step(g, b, dt, bc_mode=1, sources=z)
    assert np.array_equal(a, b)


def test_floors(oracle):
    g = Geom((1, 1, 1), 4, (0, 0, 0), (1, 1, 1))
    u = oracle.init_flat(g)
    vals_chi = [np.nan, np.inf, -np.inf, -1.0, 5e-5, 2.0, 0.7]
    exp_chi = [1e-4, 1.5, 1e-4, 1e-4, 1e-4, 1.5, 0.7]
    vals_a = [np.nan, np.inf, -1.0, 1e-7, 0.3, 2.0, -np.inf]
    exp_a = [1e-6, np.inf, 1e-6, 1e-6, 0.3, 2.0, 1e-6]
    # main.cpp:599-605: +inf chi -> 1e-4 then capped? no: !finite -> 1e-4 first
    exp_chi[1] = 1e-4
    exp_a[1] = 1e-6
    flat_c = u[V["CHI"]].reshape(-1)
    flat_a = u[V["ALPHA"]].reshape(-1)
    flat_c[: len(vals_chi)] = vals_chi
    flat_a[: len(vals_a)] = vals_a
    oracle.floors(g, u)
    assert np.array_equal(u[V["CHI"]].reshape(-1)[: len(vals_chi)], exp_chi)
    assert np.array_equal(u[V["ALPHA"]].reshape(-1)[: len(vals_a)], exp_a)


def test_algebraic_constraints(oracle):
    g = Geom((4, 4, 4), 4, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(0)
    u = oracle.init_flat(g)
    s = u.shape[1:]
    gm = np.zeros((3, 3) + s)
    for a in range(3):
        gm[a, a] = 1.0
    pert = 0.2 * rng.standard_normal((3, 3) + s)
    gm = gm + 0.5 * (pert + pert.transpose(1, 0, 2, 3, 4)) + 0.3 * np.eye(3)[:, :, None, None, None]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for p, (a, b) in enumerate(pairs):
        u[V["GXX"] + p] = gm[a, b]
        u[V["AXX"] + p] = rng.standard_normal(s)
    u.reshape(22, -1)[V["GXX"]:V["GXX"] + 6, 0] = [1e-6, 0, 0, 1e-6, 0, 1e-6]  # degenerate -> identity
    oracle.algebraic(g, u)
    G = np.empty((3, 3) + s)
    A = np.empty((3, 3) + s)
    for p, (a, b) in enumerate(pairs):
        G[a, b] = G[b, a] = u[V["GXX"] + p]
        A[a, b] = A[b, a] = u[V["AXX"] + p]
    Gm = np.moveaxis(G, (0, 1), (-2, -1))
    assert np.allclose(np.linalg.det(Gm), 1.0, rtol=0, atol=1e-13)
    tr = np.einsum("...ij,...ij->...", np.linalg.inv(Gm), np.moveaxis(A, (0, 1), (-2, -1)))
    assert np.max(np.abs(tr)) < 1e-12
    assert np.array_equal(u.reshape(22, -1)[V["GXX"]:V["GXX"] + 6, 0], [1, 0, 0, 1, 0, 1])


def _three_pass_sommerfeld(g, u):
    """Literal restatement of the X, Y, Z passes (main.cpp:395-483) in numpy."""
    finf = np.zeros(22)
    finf[[0, 1, 4, 6, 18]] = 1.0
    ng = g.ng
    X = [g.coords(d) for d in range(3)]
    u = u.copy()
    for d in range(3):
        ib, ie = ng, g.nt[d] - ng - 1
        for side, (b, gs) in enumerate([(ib, [ib - 1 - q for q in range(ng)]), (ie, [ie + 1 + q for q in range(ng)])]):
            for gi in gs:
                c_int = [X[0], X[1], X[2]]
                c_gh = [X[0], X[1], X[2]]
                c_int = [np.full_like(X[e], X[e][b]) if e == d else X[e] for e in range(3)]
                c_gh = [np.full_like(X[e], X[e][gi]) if e == d else X[e] for e in range(3)]
                xi, yi, zi = np.meshgrid(*c_int, indexing="ij")
                xg, yg, zg = np.meshgrid(*c_gh, indexing="ij")
                rint = (np.sqrt(xi * xi + yi * yi + zi * zi) + 1e-30).transpose(2, 1, 0)
                rgh = (np.sqrt(xg * xg + yg * yg + zg * zg) + 1e-30).transpose(2, 1, 0)
                sl_b = [slice(None)] * 3
                sl_g = [slice(None)] * 3
                sl_b[2 - d] = b
                sl_g[2 - d] = gi
                rint_s = rint[tuple(sl_b)]
                rgh_s = rgh[tuple(sl_g)]
                for v in range(22):
                    fint = u[v][tuple(sl_b)]
                    u[v][tuple(sl_g)] = finf[v] + (fint - finf[v]) * rint_s / rgh_s
    return u


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ghost_fill_against_literal_passes(oracle, mode):
    g = Geom((6, 5, 7), 4, (-1.3, -0.4, 0.2), (1.1, 0.6, 1.6))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0.3, 0.1, 0.7]]), 11, 0.1, range(22))
    got = oracle.fill_ghosts(g, u.copy(), mode)
    I = g.interior()
    assert np.array_equal(got[I], u[I])
    if mode == 0:
        assert np.array_equal(got, _three_pass_sommerfeld(g, u))
    elif mode == 1:
        n, ng = g.n, g.ng
        idx = [ng + (np.arange(g.nt[d]) - ng) % n[d] for d in range(3)]
        assert np.array_equal(got, u[:, idx[2]][:, :, idx[1]][:, :, :, idx[0]])
    else:
        idx = [np.clip(np.arange(g.nt[d]), g.ng, g.ng + g.n[d] - 1) for d in range(3)]
        assert np.array_equal(got, u[:, idx[2]][:, :, idx[1]][:, :, :, idx[0]])


def test_sommerfeld_fixed_point_on_inverse_radius(oracle):
    # f = f_inf + c / r is reproduced in every ghost cell (the chain of ratios telescopes)
    g = Geom((8, 8, 8), 4, (-2, -2, -2), (2, 2, 2))
    x, y, z = np.meshgrid(g.coords(0), g.coords(1), g.coords(2), indexing="ij")
    r = (np.sqrt(x * x + y * y + z * z) + 1e-30).transpose(2, 1, 0)
    u = g.zeros()
    for v in range(22):
        u[v] = (1.0 if v in (0, 1, 4, 6, 18) else 0.0) + 0.7 / r
    got = oracle.fill_ghosts(g, u.copy(), 0)
    assert np.allclose(got, u, rtol=1e-14, atol=1e-15)


def test_brill_lindquist_values(oracle):
    # psi(1) = 1.5 and psi(0.5) = 2 for m = 1 (tests/initial_data/test_brill_lindquist.cpp)
    g = Geom((8, 1, 1), 4, (-0.125, -0.125, -0.125), (1.875, 0.125, 0.125))  # centres at 0, .25, ..., 1.75
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    xs = g.coords(0)
    j = k = g.ng
    for xv, psi in [(1.0, 1.5), (0.5, 2.0)]:
        i = int(np.argmin(np.abs(xs - xv)))
        assert abs(u[0, k, j, i] - psi**-4) < 1e-14
        assert abs(u[18, k, j, i] - psi**-2) < 1e-14
        # Gamma-hat^x = (8/3) dpsi/dx / psi with dpsi/dx = -m x / (2 r^3)
        assert abs(u[14, k, j, i] - (8 / 3) * (-1 / (2 * xv**2)) / psi) < 1e-13


def test_gauge_wave_values(oracle):
    # alpha(L/4) = sqrt(1 + A) (tests/spacetime/test_gauge_wave.cpp:26-50)
    g = Geom((64, 8, 8), 4, (-0.5 / 64, 0, 0), (1 - 0.5 / 64, 0.125, 0.125))  # a centre at x = 0.25
    u = oracle.init_gauge_wave(g, 0.01, 1.0)
    i = int(np.argmin(np.abs(g.coords(0) - 0.25)))
    assert abs(u[18, 6, 6, i] - np.sqrt(1.01)) < 1e-14
    assert abs(u[0, 6, 6, i] - 1.01 ** (-1 / 3)) < 1e-14
