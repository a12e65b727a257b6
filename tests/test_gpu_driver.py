"""granite_main on the GPU: a small binary Brill-Lindquist run prints the reference's
telemetry lines, and its numbers are those of the same hierarchy driven through the Python
binding of the C ABI."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from granite_amd import build as B
from granite_amd.hierarchy import puncture_spheres, tracking_levels

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
OFFSET = 1.3621415e-6
STEP = re.compile(r"step=(\d+)\s+t=(\S+)\s+α_center=(\S+)\s+\|\|H\|\|₂=(\S+)\s+\[Blocks:\s*(\d+)\]")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_driver_run_matches_binding(engine, tmp_path):
    import granite_amd as ga
    from granite_amd.checkpoint import read_checkpoint

    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    say("granite_main bl_binary_small.yaml")
    res = subprocess.run([B.DRIVER, os.path.join(HERE, "driver", "bl_binary_small.yaml")], capture_output=True,
                         text=True, timeout=120, cwd=tmp_path)
    assert res.returncode == 0, res.stderr[-2000:]
    out = res.stdout
    assert "Evolution complete." in out
    for s in range(1, 7):
        assert f"[NaN@step={s}] all finite" in out
    rows = [(int(m[1]), float(m[2]), float(m[3]), float(m[4]), int(m[5])) for m in STEP.finditer(out)]
    assert [r[0] for r in rows] == [2, 4, 6] and all(r[4] == 4 for r in rows)

    # the same run through the binding
    spheres = puncture_spheres([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]], 15)
    geo = tracking_levels([24] * 3, [-12.0 + OFFSET] * 3, [12.0 + OFFSET] * 3, spheres, 15)
    levels = [ga.Level([ga.GridBlock(0, l, n, lo, hi, 4) for n, lo, hi in lv], engine) for l, lv in enumerate(geo)]
    H = ga.Hierarchy(levels, engine)
    levels[0].init_brill_lindquist([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]])
    for l in range(1, 4):
        H.prolongate(l)
    H.fill_ghosts(0)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    mine = []
    for step in range(1, 7):
        H.step(p, 0.25, ga.PREC_FP64, use_graph=True)
        if step % 2 == 0:
            sh, _, n = H.constraints_l2()
            mine.append((step, levels[0].read_cell(0, 18, 16, 16, 16), np.sqrt(sh / n)))
    for (s, t, a, h, _), (s2, a2, h2) in zip(rows, mine):
        say(f"step {s}: driver alpha_c={a} H={h}  binding alpha_c={a2:.6g} H={h2:.6g}")
        assert s == s2 and abs(t - 0.25 * s) < 1e-12
        assert abs(a - a2) <= 1e-5 * abs(a2) and abs(h - h2) <= 1e-5 * h2  # printed to 6 digits

    # the checkpoint at t = t_final (step 6) holds the binding's final state, bit for bit
    path = "output/bl_binary_small/checkpoint_000006.gck"
    assert f"Writing checkpoint: {path} (t=1.5)" in out
    ck = read_checkpoint(tmp_path / path)
    assert ck.step == 6 and ck.time == 1.5
    mine_blocks = [(l, b) for l, L in enumerate(levels) for b in L.blocks]
    assert len(ck.blocks) == len(mine_blocks)
    for L in levels:
        L.download()
    for cb, (l, b) in zip(ck.blocks, mine_blocks):
        assert cb.level == l and cb.ng == 4 and cb.n == tuple(b.numCells())
        assert np.array_equal(cb.dx, [b.dx(d) for d in range(3)])
        assert np.array_equal(cb.data, b.array), f"level {l}"
    say(f"checkpoint: {len(ck.blocks)} blocks equal to the binding's state")
    H.close()
    for L in levels:
        L.close()


def _driver(tmp_path, cfg, *args):
    res = subprocess.run([B.DRIVER, os.path.join(HERE, "driver", cfg), *args], capture_output=True, text=True,
                         timeout=120, cwd=tmp_path)
    assert res.returncode == 0, res.stderr[-2000:]
    assert "Evolution complete." in res.stdout
    return res.stdout


def test_driver_gauge_wave_matches_binding(engine, tmp_path):
    """One-level gauge-wave run (flat metric, lapse 1 + A sin(2 pi x / L) on every cell,
    src/main.cpp:1042-1064): the final checkpoint equals the same hierarchy stepped through
    the binding, bit for bit."""
    import math

    import granite_amd as ga
    from granite_amd.checkpoint import read_checkpoint

    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    say("granite_main gauge_wave_small.yaml")
    out = _driver(tmp_path, "gauge_wave_small.yaml")
    assert "Gauge wave: A=0.1, L=1" in out
    ck = read_checkpoint(tmp_path / "output" / "checkpoint_000008.gck")
    assert ck.step == 8 and abs(ck.time - 0.125) < 1e-15 and len(ck.blocks) == 1

    lo = [v + OFFSET for v in (-0.5, -0.25, -0.25)]
    hi = [v + OFFSET for v in (0.5, 0.25, 0.25)]
    blk = ga.GridBlock(0, 0, (16, 8, 8), lo, hi, 4)
    u = blk.array
    u[0] = u[1] = u[4] = u[6] = 1.0
    for i in range(u.shape[3]):
        u[18, :, :, i] = 1.0 + 0.1 * math.sin(2.0 * math.pi * blk.x(0, i) / 1.0)
    L = ga.Level([blk], engine)
    L.upload()
    H = ga.Hierarchy([L], engine)
    H.fill_ghosts(0)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    t = 0.0
    for _ in range(8):
        dt = 0.25 / 16
        if t + dt > 0.125:
            dt = 0.125 - t
        H.step(p, dt, ga.PREC_FP64, use_graph=True)
        t += dt
    L.download()
    assert np.array_equal(ck.blocks[0].data, blk.array)
    say("gauge wave: checkpoint equal to the binding's state")
    H.close()
    L.close()


def test_driver_mixed_precision_matches_binding(engine, tmp_path):
    """--precision mixed: the binary run's final checkpoint equals the binding's mixed-form
    hierarchy step for step (the same kernels, the same tile classification)."""
    import granite_amd as ga
    from granite_amd.checkpoint import read_checkpoint

    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    say("granite_main bl_binary_small.yaml --precision mixed")
    _driver(tmp_path, "bl_binary_small.yaml", "--precision", "mixed")
    ck = read_checkpoint(tmp_path / "output" / "bl_binary_small" / "checkpoint_000006.gck")
    spheres = puncture_spheres([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]], 15)
    geo = tracking_levels([24] * 3, [-12.0 + OFFSET] * 3, [12.0 + OFFSET] * 3, spheres, 15)
    levels = [ga.Level([ga.GridBlock(0, l, n, lo, hi, 4) for n, lo, hi in lv], engine) for l, lv in enumerate(geo)]
    H = ga.Hierarchy(levels, engine)
    levels[0].init_brill_lindquist([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]])
    for l in range(1, len(levels)):
        H.prolongate(l)
    H.fill_ghosts(0)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    for _ in range(6):
        H.step(p, 0.25, ga.PREC_MIXED, use_graph=True)
    mine = [b for L in levels for b in L.blocks]
    for L in levels:
        L.download()
    assert len(ck.blocks) == len(mine)
    for cb, b in zip(ck.blocks, mine):
        assert np.array_equal(cb.data, b.array)
    say(f"mixed: {len(mine)} blocks equal to the binding's mixed-form state")
    H.close()
    for L in levels:
        L.close()


# the patterns python/granite_analysis/core/parsers.py:33-38 reads these lines with
RE_NAN_ST = re.compile(r"\[NaN@step=(\d+)\]\s+ST\s+var=(\d+)\s+\((\d+),(\d+),(\d+)\)\s*=\s*(\S+)")
RE_CFL_W = re.compile(r"\[CFL-(?:WARN|GUARD)\].*CFL=([\d.eE+\-]+)")


def _yaml(tmp_path, name, **over):
    text = open(os.path.join(HERE, "driver", "bl_binary_small.yaml")).read()
    for key, val in over.items():
        text = re.sub(rf"(\n\s*{key}:)\s*[^\n#]*", rf"\1 {val} ", text, count=1)
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_driver_nan_line_parses_like_the_reference(engine, tmp_path):
    """A NaN dissipation strength (applyDissipation only skips sigma <= 0) poisons every
    interior RHS value in the first step: chi and the lapse are floored, the conformal metric
    is reset to the identity by the algebraic constraints, and the first non-finite value in
    the reference's scan order is A~_xx at the first interior cell of the coarsest block. The
    driver prints it in the reference's format, which the telemetry parser's pattern reads."""
    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    cfg = _yaml(tmp_path, "nan.yaml", ko_sigma="nan", t_final="0.5", output_interval="1")
    say("granite_main with ko_sigma = nan")
    res = subprocess.run([B.DRIVER, cfg, "--no-checkpoints"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path)
    assert res.returncode == 0, res.stderr[-2000:]
    hits = [m.groups() for m in RE_NAN_ST.finditer(res.stdout)]
    assert [h[:5] for h in hits] == [("1", "7", "4", "4", "4"), ("2", "7", "4", "4", "4")], res.stdout[-3000:]
    assert all(h[5].lstrip("-") == "nan" for h in hits)
    assert "all finite" not in res.stdout


def test_driver_cfl_guard_lines(engine, tmp_path):
    """At CFL 2 the shift driven by the punctures' Gamma-hat exceeds the advective limit in the
    first coarse steps: the driver prints the reference's [CFL-GUARD] line for every level
    step above 0.95, in subcycle order, with the values the binding's CFL monitor gives."""
    import granite_amd as ga

    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    cfg = _yaml(tmp_path, "cfl.yaml", cfl="2.0", t_final="4.0")
    say("granite_main at CFL 2")
    res = subprocess.run([B.DRIVER, cfg, "--no-checkpoints"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path)
    assert res.returncode == 0, res.stderr[-2000:]
    printed = [float(m[1]) for m in RE_CFL_W.finditer(res.stdout)]
    spheres = puncture_spheres([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]], 15)
    geo = tracking_levels([24] * 3, [-12.0 + OFFSET] * 3, [12.0 + OFFSET] * 3, spheres, 15)
    levels = [ga.Level([ga.GridBlock(0, l, n, lo, hi, 4) for n, lo, hi in lv], engine) for l, lv in enumerate(geo)]
    H = ga.Hierarchy(levels, engine)
    H.set_cfl_monitor(True)
    levels[0].init_brill_lindquist([0.5, 0.5], [[1.5, 0, 0], [-1.5, 0, 0]])
    for l in range(1, 4):
        H.prolongate(l)
    H.fill_ghosts(0)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    want = []
    for _ in range(2):
        H.step(p, 2.0, ga.PREC_FP64, use_graph=True)
        vals = H.cfl()
        assert [lv for lv, _ in vals] == [0, 1, 2, 3, 3, 2, 3, 3, 1, 2, 3, 3, 2, 3, 3]
        want += [v for _, v in vals if v > 0.95]
    say(f"CFL lines: driver {printed[:4]}..., binding {want[:4]}...")
    assert len(want) >= 1 and len(printed) == len(want)
    assert all(abs(a - b) <= 1e-5 * b for a, b in zip(printed, want))
    H.close()
    for L in levels:
        L.close()
