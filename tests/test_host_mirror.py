"""The C++17 host mirror (granite_amd/host/granite_hip/ccz4_hip.hpp) as a GRANITE translation
unit would use it: compiled with g++ against include/granite_ccz4.h and linked to the
library; on a GPU its results are checked against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle.oracle import Geom, group_rel, normwise_rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mirror") / "host_mirror_driver")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "granite_amd", "host"), os.path.join(ROOT, "tests", "cpp", "host_mirror_driver.cpp"),
           "-L", os.path.join(ROOT, "granite_amd"), "-lgranite_ccz4", "-Wl,-rpath," + os.path.join(ROOT, "granite_amd"),
           "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def write_input(path, g: Geom, u):
    with open(path, "wb") as f:
        f.write(struct.pack("4i", *g.n, g.ng))
        f.write(struct.pack("6d", *g.lo, *g.hi))
        f.write(np.ascontiguousarray(u).tobytes())


def test_mirror_compiles_and_fails_cleanly_without_gpu(driver, tmp_path):
    import granite_amd

    try:
        granite_amd.Engine(0)
        pytest.skip("a GPU is present; the GPU test covers this driver")
    except granite_amd.EngineError:
        pass
    g = Geom((4, 4, 4), 4, (0, 0, 0), (1, 1, 1))
    write_input(tmp_path / "in.bin", g, g.zeros())
    out = subprocess.run([driver, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert out.returncode == 1 and "no usable gfx950 device" in out.stdout


@pytest.mark.gpu
def test_mirror_against_oracle(driver, oracle, tmp_path):
    g = Geom((20, 16, 12), 4, (-2.5, -2.0, -1.5), (2.5, 2.0, 1.5))
    u = oracle.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    oracle.rk3_step(g, u, 0.0625, 0)
    write_input(tmp_path / "in.bin", g, u)
    print("[gpu] C++ host mirror driver", flush=True)
    out = subprocess.run([driver, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "expected error" in out.stdout
    assert "hierarchy ok" in out.stdout  # Hierarchy + Checkpoint through the mirror
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    n = u.size
    N = g.stride
    r1 = raw[:n].reshape(u.shape)
    r2 = raw[n:2 * n].reshape(u.shape)
    ham = raw[2 * n:2 * n + N].reshape(u.shape[1:])
    mom = raw[2 * n + N:2 * n + 4 * N].reshape(u.shape[1:] + (3,))
    s = raw[2 * n + 4 * N:].reshape(u.shape)
    I = g.interior()
    assert normwise_rel(r1[I], oracle.rhs(g, u)[I]).max() < 1e-12
    assert np.array_equal(r1, r2)
    h, m = oracle.constraints(g, u)
    assert np.max(np.abs(ham - h)) <= 1e-12 * np.max(np.abs(h))
    assert np.max(np.abs(mom - m)) <= 1e-12 * np.max(np.abs(m))
    ref = u.copy()
    oracle.rk3_step(g, ref, 0.0625, 0)
    oracle.rk3_step(g, ref, 0.0625, 0)
    assert group_rel(s, ref).max() < 1e-12


REFERENCE = "/root/reference"


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "include", "granite", "core", "grid.hpp")),
                    reason="reference sources not present on this machine")
def test_mirror_with_the_reference_types(tmp_path):
    """The mirror instantiated with GRANITE's own granite::GridBlock and
    granite::spacetime::CCZ4Params (the reference's headers and src/core/grid.cpp, compiled
    where they lie): every method compiles against the real types, the geometry handed to the
    engine is the block's own, and without a device the engine's constructor throws the
    runtime_error the reference's callers would see."""
    import granite_amd

    exe = str(tmp_path / "reference_types_mirror")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(REFERENCE, "include"), "-I",
           os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "granite_amd", "host"),
           os.path.join(ROOT, "tests", "cpp", "reference_types_mirror.cpp"),
           os.path.join(REFERENCE, "src", "core", "grid.cpp"), "-L", os.path.join(ROOT, "granite_amd"),
           "-lgranite_ccz4", "-Wl,-rpath," + os.path.join(ROOT, "granite_amd"), "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    try:
        granite_amd.Engine(0)
        have_gpu = True
    except granite_amd.EngineError:
        have_gpu = False
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if have_gpu:
        assert out.returncode == 0 and "ran on the device" in out.stdout, out.stdout + out.stderr
    else:
        assert out.returncode == 1 and "no usable gfx950 device" in out.stdout, out.stdout + out.stderr
