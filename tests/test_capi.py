"""The C ABI library loads and exports every entry point include/granite_ccz4.h declares.
No kernel is launched here."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "granite_ccz4.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gr_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    import granite_amd

    path = granite_amd.library_path()
    if not os.path.exists(path):
        pytest.fail("libgranite_ccz4.so not built: run __graft_entry__.build()")
    return ctypes.CDLL(path)


def test_header_declares_the_boundary():
    names = declared()
    for must in ["gr_engine_create", "gr_ccz4_rhs", "gr_ccz4_dissipation", "gr_ccz4_constraints",
                 "gr_level_create", "gr_level_rk3_step", "gr_level_constraints_l2"]:
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "granite_amd", "libgranite_ccz4.so")],
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gr_\w+)", out))
    assert set(declared()) <= exported


def test_host_only_entry_points(lib):
    lib.gr_version.restype = ctypes.c_char_p
    assert lib.gr_version().startswith(b"granite-ccz4")
    t = (ctypes.c_int * 3)()
    assert lib.gr_stage_tile(0, t) == 0 and all(v > 0 for v in t)
    assert lib.gr_stage_tile(7, t) != 0
    assert lib.gr_engine_create(0, None) == -1  # null output pointer is rejected before any device call


def test_library_contains_gfx950_code(tmp_path):
    # the embedded offload bundle of our own build names its target
    sec = tmp_path / "fatbin.bin"
    out = subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin",
                          os.path.join(ROOT, "granite_amd", "libgranite_ccz4.so"), str(sec)], capture_output=True)
    if out.returncode != 0:
        pytest.skip("objcopy not available")
    assert b"gfx950" in sec.read_bytes()


def test_bench_refuses_more_than_one_gpu():
    # one GPU is measured; a request for more exits with a message before touching a device
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present: the engine can be created")
def test_no_device_fails_loudly(lib):
    # without a gfx950 device the engine is not created and nothing falls back to the CPU
    eng = ctypes.c_void_p()
    assert lib.gr_engine_create(0, ctypes.byref(eng)) == -4  # GR_ERR_NODEV
    assert eng.value is None
    import granite_amd as ga

    with pytest.raises(ga.EngineError):
        ga.Engine(0)
