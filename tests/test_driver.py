"""granite_main (granite_amd/host/granite_main.cpp) on the CPU: parameter parsing and the
level geometry it prints (--dry-run touches no GPU) against granite_amd.hierarchy, which is
pinned to the reference's regrid. The GPU run is in test_gpu_driver.py."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from granite_amd import build as B
from granite_amd.hierarchy import puncture_spheres, tracking_levels

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "driver")
OFFSET = 1.3621415e-6  # the reference driver's grid phase shift (src/main.cpp:835-841)
LEVEL = re.compile(r"AMR: Level (\d+)\s+block_id=(\d+)\s+ncells=\((\d+)x(\d+)x(\d+)\)\s+lo=\(([^)]*)\)\s+hi=\(([^)]*)\)")


def run(*args, check=False):
    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    res = subprocess.run([B.DRIVER, *args], capture_output=True, text=True, timeout=60)
    if check:
        assert res.returncode == 0, res.stderr
    return res


def test_dry_run_geometry_matches_tracking_levels():
    out = run(os.path.join(FIX, "bl_binary_small.yaml"), "--dry-run", check=True).stdout
    assert "Grid: 24 x 24 x 24" in out and "dx = 1, dt = 0.25" in out and "t_final = 1.5" in out
    assert "Brill-Lindquist: 2 BH(s), ADM mass = 1 M_sun" in out
    got = [(int(m[1]), [int(m[3]), int(m[4]), int(m[5])], [float(v) for v in m[6].split(",")],
            [float(v) for v in m[7].split(",")]) for m in LEVEL.finditer(out)]
    spheres = puncture_spheres([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]], 15)
    lo = [-12.0 + OFFSET] * 3
    hi = [12.0 + OFFSET] * 3
    ref = tracking_levels([24] * 3, lo, hi, spheres, 15)
    assert len(ref) == 4 and len(got) == 3
    for (l, n, glo, ghi), lv in zip(got, ref[1:]):
        (rn, rlo, rhi), = lv
        assert n == rn
        # printed with 6 significant digits
        assert all(abs(a - b) <= 1e-5 * max(1.0, abs(b)) for a, b in zip(glo + ghi, rlo + rhi))


def test_dry_run_single_level_gauge_wave():
    out = run(os.path.join(FIX, "gauge_wave_small.yaml"), "--dry-run", check=True).stdout
    assert "Gauge wave: A=0.1, L=1" in out
    assert "Grid: 16 x 8 x 8" in out and "dx = 0.0625, dt = 0.015625" in out
    assert not LEVEL.search(out)


def test_rejects_unsupported_initial_data(tmp_path):
    p = tmp_path / "tp.yaml"
    p.write_text("initial_data:\n  type: two_punctures\n")
    res = run(str(p), "--dry-run")
    assert res.returncode == 2 and "elliptic" in res.stderr


def test_reports_yaml_errors(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("grid:\n  ncells: [8, 8\n")
    res = run(str(p), "--dry-run")
    assert res.returncode == 1 and "YAML parsing error" in res.stderr
    p.write_text("grid:\n  ncells: [8, 8, 8]\n  ghost_cells: 3\n")
    res = run(str(p), "--dry-run")
    assert res.returncode == 2 and "ghost_cells" in res.stderr


REFERENCE_YAML = sorted(glob.glob("/root/reference/benchmarks/**/*.yaml", recursive=True))


@pytest.mark.skipif(not REFERENCE_YAML, reason="reference benchmark files not present on this machine")
@pytest.mark.parametrize("path", REFERENCE_YAML, ids=os.path.basename)
def test_reference_parameter_files_parse(path):
    """Every parameter file of the reference's benchmarks reads without a YAML error (the
    driver may still refuse a setup it does not cover, e.g. two_punctures, with exit 2)."""
    res = run(path, "--dry-run")
    assert "YAML parsing error" not in res.stderr, res.stderr
    assert res.returncode in (0, 2), res.stderr


@pytest.fixture(scope="module")
def yaml_harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("yaml") / "yaml_harness"
    src = os.path.join(HERE, "cpp", "yaml_harness.cpp")
    inc = os.path.join(os.path.dirname(HERE), "granite_amd", "host")
    res = subprocess.run(["g++", "-std=c++17", "-O1", "-I", inc, src, "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return str(exe)


def test_yaml_block_scalars(yaml_harness, tmp_path):
    """Folded (>) and literal (|) block scalars with clip, strip (-) and keep (+) chomping,
    a '#' inside the text, a blank line inside a folded scalar, and the mapping that follows."""
    p = tmp_path / "b.yaml"
    p.write_text(
        "scenario:\n"
        "  folded: >\n"
        "    five black holes\n"
        "    # of equal mass\n"
        "\n"
        "    at a pentagon\n"
        "  literal: |\n"
        "    line one\n"
        "      indented\n"
        "  strip: >-\n"
        "    a\n"
        "    b\n"
        "  keep: |+\n"
        "    k\n"
        "\n"
        "  name: B5  # comment\n"
        "after: 3\n")

    def get(*keys):
        res = subprocess.run([yaml_harness, str(p), *keys], capture_output=True, text=True)
        assert res.returncode == 0, res.stderr
        return res.stdout.strip()

    assert get("scenario", "folded") == "[five black holes # of equal mass\\nat a pentagon\\n]"
    assert get("scenario", "literal") == "[line one\\n  indented\\n]"
    assert get("scenario", "strip") == "[a b]"
    assert get("scenario", "keep") == "[k\\n\\n]"
    assert get("scenario", "name") == "[B5]"
    assert get("after") == "[3]"


def test_driver_installed_where_the_reference_launcher_looks():
    """scripts/run_granite.py:292-297 runs build/bin/granite_main under its default build
    directory; the build places the driver there too."""
    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    assert os.path.exists(B.DRIVER_LAUNCHER)
    res = subprocess.run([B.DRIVER_LAUNCHER, os.path.join(FIX, "gauge_wave_small.yaml"), "--dry-run"],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "Grid: 16 x 8 x 8" in res.stdout


def test_checkpoint_reader_layout(tmp_path):
    """The reader takes the flat layout granite_main writes: header, then per block its
    metadata and nvars * stride values (x fastest); a short file is an error."""
    import struct

    from granite_amd.checkpoint import read_checkpoint

    rng = np.random.default_rng(3)
    blocks = [(0, 4, (3, 2, 1), rng.standard_normal((22, 9, 10, 11))),
              (1, 4, (2, 2, 2), rng.standard_normal((22, 10, 10, 10)))]
    head = b"GRCKPT01" + struct.pack("<iid", 7, len(blocks), 1.75)
    body = b""
    for level, ng, n, data in blocks:
        body += struct.pack("<6i", level, ng, *n, 22) + struct.pack("<6d", -1.0, -2.0, -3.0, 0.5, 0.25, 0.125)
        body += data.astype("<f8").tobytes()
    path = tmp_path / "c.gck"
    path.write_bytes(head + body)
    ck = read_checkpoint(path)
    assert ck.step == 7 and ck.time == 1.75 and len(ck.blocks) == 2
    for got, (level, ng, n, data) in zip(ck.blocks, blocks):
        assert got.level == level and got.ng == ng and got.n == n
        assert np.array_equal(got.lo, [-1.0, -2.0, -3.0]) and np.array_equal(got.dx, [0.5, 0.25, 0.125])
        assert np.array_equal(got.data, data)
    path.write_bytes(head + body[:-8])
    with pytest.raises(ValueError):
        read_checkpoint(path)
