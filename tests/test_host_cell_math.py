"""The stage kernel's per-cell arithmetic (granite_amd/csrc/ccz4_cell.hpp), compiled for the
host with g++, against the oracle. This checks the kernel's algebra without a GPU; the GPU
parity tests check the kernel itself."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Geom, _p, normwise_rel, param_array

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hcell(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hcell") / "libhcell.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fno-fast-math", "-ffp-contract=off",
                    "-Wno-unknown-pragmas", "-o", out, os.path.join(HERE, "host_cell_harness.cpp")], check=True)
    L = C.CDLL(out)

    def rhs(g, u, sources=None, ko_exact=1, **prm):
        r = g.zeros()
        rho = Si = Sij = S = None
        if sources is not None:
            rho, Si, Sij, S = sources
        L.hcell_rhs((C.c_int * 3)(*g.n), C.c_int(g.ng), (C.c_double * 3)(*g.dx), param_array(**prm), _p(u), _p(r),
                    _p(rho), _p(Si), _p(Sij), _p(S), C.c_int(ko_exact))
        return r

    return rhs


def cases(oracle):
    g = Geom((8, 8, 8), 4, (-1, -1, -1), (1, 1, 1))
    yield "bl", g, oracle.init_bl(g, [1.0], [[0, 0, 0]])
    yield "bl_noise", g, oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0, 0, 0]]), 1234, 1e-3, range(1, 22))
    g2 = Geom((16, 12, 10), 4, (-2, -1.5, -1.25), (2, 1.5, 1.25))
    u2 = oracle.init_bl(g2, [0.5, 0.5], [[-0.8, 0.1, 0], [0.8, -0.1, 0.05]])
    yield "binary_noise8", g2, oracle.add_noise(g2, u2, 7, 1e-8, [1, 2, 3, 4, 5, 6, 13])
    g3 = Geom((16, 4, 4), 4, (0, 0, 0), (1, 0.25, 0.25))
    yield "gauge", g3, oracle.init_gauge_wave(g3, 0.1, 1.0)


def test_cell_math_matches_oracle(oracle, hcell):
    for name, g, u in cases(oracle):
        I = g.interior()
        for sigma in (0.0, 0.1):
            e = normwise_rel(hcell(g, u, ko_sigma=sigma)[I], oracle.rhs(g, u, ko_sigma=sigma)[I])
            assert e.max() < 1e-13, (name, sigma, e)


def test_cell_math_with_sources(oracle, hcell):
    g = Geom((8, 8, 8), 4, (-1, -1, -1), (1, 1, 1))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0, 0, 0]]), 5, 1e-3, range(1, 22))
    rng = np.random.default_rng(3)
    N = g.stride
    src = (rng.standard_normal(N), rng.standard_normal(3 * N), rng.standard_normal(6 * N), rng.standard_normal(N))
    I = g.interior()
    e = normwise_rel(hcell(g, u, sources=src)[I], oracle.rhs(g, u, sources=src)[I])
    assert e.max() < 1e-13, e


def test_exact_order_dissipation_is_bitwise_on_ko_dominated_fields(oracle, hcell):
    # at t = 0 with zero shift the metric RHS is the dissipation term alone
    g = Geom((8, 8, 8), 4, (-1, -1, -1), (1, 1, 1))
    u = oracle.add_noise(g, oracle.init_bl(g, [1.0], [[0, 0, 0]]), 9, 1e-8, [1, 2, 3, 4, 5, 6])
    I = g.interior()
    a = hcell(g, u)[I]
    b = oracle.rhs(g, u)[I]
    assert np.array_equal(a[1:7], b[1:7])
