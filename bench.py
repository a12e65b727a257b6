"""Headline benchmark: cell-updates/s per SSP-RK3 step of the CCZ4 engine on one MI355X.

Workload (one "step" = one full RK3 step of the hot path over the block): a single
Brill-Lindquist puncture, M = 1 at the origin, one uniform 128^3 block over [-16, 16]^3
with 4 ghost cells (dx = 0.25, dt = CFL 0.25 * dx = 0.0625), kappa1 = 0.02, eta = 2,
KO sigma = 0.1, FP64, level-0 ghost semantics of the reference (outflow fill before the
step, 1/r fall-off after every stage). Initial data is generated on the device; the state
is resident in HBM before timing starts. Synthetic data (no checkpoint is loaded).

Prints one JSON line. Also reports the stage kernel's roofline (algorithmic bytes per
launch / its average launch time from HIP events on the engine stream) and a CPU baseline
timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector
N = 128
HALF = 16.0
N_NOISY = 192
DT = 0.25 * (2 * HALF / N)
BYTES_PER_CELL_STEP = 1408  # 3 stages x (22-field read + write) + 2 x 22-field read of u^n, FP64


def host_cpu():
    """The host's CPU as lscpu and the scheduler report it."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": None,
            "physical_cores": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline():
    """The compiled reference (oracle/_ref, Release flags) on this host's cores with the
    reference launcher's OpenMP placement (scripts/run_granite.py:138-149): every core of the
    process's affinity, but no more than the CPU share a scheduler grants the job through
    OMP_NUM_THREADS when it sets it; the median of 3 runs of 5 steps' RHS work, plus a
    single-thread leg."""
    host = host_cpu()
    threads = host["affinity"]
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and 0 < int(share) < threads:
        threads = int(share)
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--n", str(N), "--half-width", str(HALF)],
                         capture_output=True, text=True, env=env, timeout=900, cwd=ROOT)
    if out.returncode != 0:
        return {"value": None, "unit": "cell-updates/s", "cores": threads, "kind": None,
                "sample": "failed: " + out.stderr[-300:], "host": host}
    res = json.loads(out.stdout.strip().splitlines()[-1])
    res["host"] = host
    res["omp"] = {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores",
                  "share_from_env": share}
    # the same code on one thread, on a 48^3 block at the same spacing (median of 3 runs of one
    # step's RHS work)
    env1 = dict(env, OMP_NUM_THREADS="1")
    one = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--n", "48", "--half-width",
                          str(HALF * 48 / N), "--steps", "1"], capture_output=True, text=True, env=env1, timeout=600,
                         cwd=ROOT)
    if one.returncode == 0:
        r1 = json.loads(one.stdout.strip().splitlines()[-1])
        res["single_thread"] = {"value": r1["value"], "cores": r1["cores"], "sample": r1["sample"]}
    # the 192^3 block with 1e-8 noise on the metric and K (the noisy_192_mixed workload), same
    # cores: the median of 3 runs of one step's RHS work
    big = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--n", str(N_NOISY), "--half-width",
                          str(HALF), "--steps", "1", "--noise", "1e-8"], capture_output=True, text=True, env=env,
                         timeout=900, cwd=ROOT)
    if big.returncode == 0:
        r3 = json.loads(big.stdout.strip().splitlines()[-1])
        res["noisy_192"] = {"value": r3["value"], "cores": r3["cores"], "sample": r3["sample"],
                            "seconds_per_step": r3["seconds_per_step"]}
    return res


def copy_calibration(nbytes):
    """Five plain device-to-device copies of nbytes (the HIP runtime's copy kernel): the
    FETCH_SIZE / WRITE_SIZE counters of one such copy read ~nbytes each when the counters
    report HBM bytes exactly; their ratio to nbytes is the calibration factor."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "copies": 5}), flush=True)


STAGING = {0: "registers -> padded LDS, 8x4x4 parts", 1: "direct global -> LDS, 8x4x4 parts",
           2: "direct global -> LDS, 8x4x8 parts"}


def level_tuning(ga, lvl, times=None):
    """The autotuner's choices in force on a level (and the per-variant times it measured)."""
    ngroups = len(times["stage"]) if times else 1
    out = {"stage_staging": [STAGING[lvl.tuning(ga.TUNE_STAGE, g)] for g in range(ngroups)],
           "ghost_field_groups": lvl.tuning(ga.TUNE_GHOST, 0)}
    if times:
        out["stage_ms_per_variant"] = times["stage"]
        out["ghost_ms_per_groups"] = {str(k): v for k, v in times["ghost"].items()}
    return out


def run_batched(ga, eng, p, args):
    """Uneven AMR blocks (granite_amd.workloads.uneven_amr_levels: ~250 blocks, edges in
    {16..64}, 4 levels): one RK3 step of every level (each with dt = 0.25 dx_l), either
    with one stage launch per level (batched; FP64 and mixed) or with every block in its
    own level (per-block launches, FP64). Value = interior cells / time of that step."""
    from granite_amd import workloads

    layout = workloads.uneven_amr_levels()
    total = workloads.cells(layout)
    masses, pos = workloads.PUNCTURES
    steps = max(3, args.steps // 5)
    out = {"blocks": sum(len(b) for _, b in layout), "levels": len(layout), "cells": total}

    def timed(levels, precision, tune, together=False):
        tuning = []
        for lv, dx in levels:
            lv.init_brill_lindquist(masses, pos)
            if tune:
                tuning.append(level_tuning(ga, lv, lv.autotune(p, 0.25 * dx) if args.autotune else None))
        def step():
            if together:
                # the level set as one schedule: each level's ghost fills overlap the
                # following levels' stage kernels (gr_levels_rk3_step)
                ga.levels_rk3_step([lv for lv, _ in levels], [0.25 * dx for _, dx in levels], p, precision=precision,
                                   pre_fill=ga.BC_OUTFLOW, use_graph=True)
            else:
                for lv, dx in levels:
                    lv.rk3_step(p, 0.25 * dx, precision=precision, pre_fill=ga.BC_OUTFLOW, use_graph=True)

        for _ in range(2):
            step()
        for lv, _ in levels:
            lv.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        for lv, _ in levels:
            lv.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        bad = sum(lv.nonfinite() for lv, _ in levels)
        r = {"ms_per_step": ms, "value": total / (ms * 1e-3), "nonfinite": bad}
        if tuning:
            r["tuning_per_level"] = tuning
        return r

    def make(batched):
        levels = []
        for dx, blocks in layout:
            grids = [ga.GridBlock(i, 0, n, lo, hi, 4) for i, (n, lo, hi) in enumerate(blocks)]
            if batched:
                levels.append((ga.Level(grids, eng), dx))
            else:
                levels += [(ga.Level([g], eng), dx) for g in grids]
        return levels

    for name, batched, prec, together in (("fp64_batched", True, ga.PREC_FP64, True),
                                          ("mixed_batched", True, ga.PREC_MIXED, True),
                                          ("fp64_batched_levels_one_by_one", True, ga.PREC_FP64, False),
                                          ("fp64_per_block", False, ga.PREC_FP64, False)):
        levels = make(batched)
        out[name] = timed(levels, prec, batched, together)
        for lv, _ in levels:
            lv.close()
    out["batched_speedup_vs_per_block"] = out["fp64_per_block"]["ms_per_step"] / out["fp64_batched"]["ms_per_step"]
    out["level_set_speedup_vs_one_by_one"] = (out["fp64_batched_levels_one_by_one"]["ms_per_step"] /
                                              out["fp64_batched"]["ms_per_step"])
    out["mixed_batched_vs_fp64_per_block"] = out["fp64_per_block"]["ms_per_step"] / out["mixed_batched"]["ms_per_step"]
    out["steps"] = steps
    return out


def run_noisy_192(ga, eng, p, steps, tune=True, autotune=True):
    """The mixed-precision single-puncture case: BL M = 1, one 192^3 block on [-16,16]^3
    (dx = 1/6, dt = 1/24) with N(0, 1e-8^2) added to the conformal metric and K (seed
    20240611, the engine's portable generator). Per-field max RHS error of the mixed form
    relative to the FP64 form (tensor-group scale, all interior cells; the FP64 form matches
    the oracle to 1e-12), then `steps` RK3 steps of each form with their ||H||_2."""
    import numpy as np

    n, half = N_NOISY, HALF
    dt = 0.25 * (2 * half / n)
    noise_fields = [1, 2, 3, 4, 5, 6, 13]
    out = {"block": [n] * 3, "dt": dt, "noise": 1e-8, "steps": steps}
    groups = [[0], list(range(1, 7)), list(range(7, 13)), [13], [14, 15, 16], [17], [18], [19, 20, 21]]
    res = {}
    for name, prec in (("fp64", ga.PREC_FP64), ("mixed", ga.PREC_MIXED)):
        blk = ga.GridBlock(0, 0, (n, n, n), (-half,) * 3, (half,) * 3, 4)
        lvl = ga.Level([blk], eng)
        lvl.init_brill_lindquist([1.0], [[0.0, 0.0, 0.0]])
        lvl.add_noise(20240611, 1e-8, noise_fields)
        if tune:
            out.setdefault("tuning", level_tuning(ga, lvl, lvl.autotune(p, dt) if autotune else None))
        r = lvl.rhs(0, p, prec)[:, 4:-4, 4:-4, 4:-4]
        res[name] = r
        st = dict(precision=prec, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
        lvl.rk3_step(p, dt, use_graph=True, **st)
        lvl.sync()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            lvl.rk3_step(p, dt, use_graph=True, **st)
        lvl.sync()
        ms = (time.perf_counter() - t0) * 1e3 / max(steps - 1, 1)
        sumsq, count = lvl.constraints_l2()
        out[name] = {"ms_per_step": ms, "value": n ** 3 / (ms * 1e-3), "ham_l2": (sumsq / count) ** 0.5,
                     "nonfinite": lvl.nonfinite(), "t": steps * dt}
        lvl.close()
    a, b = res["mixed"], res["fp64"]
    scale = np.empty(22)
    for grp in groups:
        scale[grp] = max(float(np.max(np.abs(b[grp]))), 1e-300)
    err = [float(np.max(np.abs(a[v] - b[v])) / scale[v]) for v in range(22)]
    out["mixed_rhs_rel_err_per_field"] = err
    out["mixed_rhs_rel_err_max"] = max(err)
    out["mixed_speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    out["mixed_ham_l2_rel_vs_fp64"] = abs(out["mixed"]["ham_l2"] / out["fp64"]["ham_l2"] - 1.0)
    return out


def checkpoint_staging(ga, eng, levels, H, p, dt0, prec, ms_step):
    """gr_checkpoint on the hierarchy's whole state (every block, ghosts included, as
    writeCheckpoint stores it): the device snapshot's time on the engine stream, the
    device -> pinned host transfer, and coarse steps run while that transfer is in flight."""
    ck = ga.Checkpoint(levels, eng)
    nbytes = sum(22 * b.totalSize() * 8 for lv in levels for b in lv.blocks)
    ck.begin()
    ck.wait()  # first touch of the pinned pages
    H.sync()
    t0 = time.perf_counter()
    ck.begin()
    H.sync()
    t1 = time.perf_counter()
    ck.wait()
    t2 = time.perf_counter()
    ck.begin()
    for _ in range(3):
        H.step(p, dt0, prec, use_graph=True)
    H.sync()
    t3 = time.perf_counter()
    ck.wait()
    ck.close()
    return {"bytes": nbytes, "snapshot_ms": (t1 - t0) * 1e3, "d2h_ms": (t2 - t1) * 1e3,
            "d2h_GBps": nbytes / (t2 - t1) / 1e9,
            "coarse_step_ms_during_transfer": (t3 - t2) * 1e3 / 3, "coarse_step_ms_alone": ms_step}


def run_hierarchy(ga, eng, p, args):
    """The 4-level binary-puncture hierarchy (granite_amd.hierarchy.binary_puncture_levels:
    96^3 on [-64,64]^3 plus tracking-sphere levels of 106x96^2, 114x96^2, 132x96^2 cells at
    dx 2/3, 1/3, 1/6; punctures m = 0.5 at (+-3,0,0)) evolved to t = 10 M: 30 Berger-Oliger
    coarse steps of dt0 = 1/3, each one replay of a captured HIP graph (level l steps 2^l
    times). Value = interior cell updates of all levels per second, over coarse steps 2-30
    (step 1 captures the graph). Constraint telemetry at t = 10 for FP64 and mixed."""
    from granite_amd import hierarchy as hy

    geo = hy.binary_puncture_levels()
    per_step = hy.cell_steps_per_coarse_step(geo)
    masses, pos = hy.PUNCTURES
    dt0 = 0.25 * (128.0 / 96)
    out = {"levels": [[n, lo, hi] for lv in geo for n, lo, hi in lv], "cell_updates_per_coarse_step": per_step,
           "coarse_steps": 30, "dt0": dt0}
    for name, prec in (("fp64", ga.PREC_FP64), ("mixed", ga.PREC_MIXED)):
        levels = [ga.Level([ga.GridBlock(0, l, n, lo, hi, 4) for n, lo, hi in lv], eng) for l, lv in enumerate(geo)]
        H = ga.Hierarchy(levels, eng)
        levels[0].init_brill_lindquist(masses, pos)
        for l in range(1, len(levels)):
            H.prolongate(l)
        tuning = None
        if args.autotune or args.tune_table:
            if args.autotune:
                H.autotune(p, dt0)
            tuning = {"per_level": [level_tuning(ga, lv) for lv in levels],
                      "transfer_field_groups": [H.tuning(l) for l in range(1, len(levels))]}
        H.fill_ghosts(0)
        H.step(p, dt0, prec, use_graph=True)
        H.sync()
        t0 = time.perf_counter()
        for _ in range(29):
            H.step(p, dt0, prec, use_graph=True)
        H.sync()
        ms = (time.perf_counter() - t0) * 1e3 / 29
        sh, sm, n = H.constraints_l2()
        bad = sum(lv.nonfinite() for lv in levels)
        out[name] = {"ms_per_coarse_step": ms, "value": per_step / (ms * 1e-3), "ham_l2_t10": (sh / n) ** 0.5,
                     "mom_l2_t10": (sm / n) ** 0.5, "nonfinite": bad}
        if tuning:
            out[name]["tuning"] = tuning
        if name == "fp64":
            out["checkpoint_staging"] = checkpoint_staging(ga, eng, levels, H, p, dt0, prec, ms)
        H.close()
        for lv in levels:
            lv.close()
    out["mixed_speedup_vs_fp64"] = out["fp64"]["ms_per_coarse_step"] / out["mixed"]["ms_per_coarse_step"]
    out["mixed_ham_l2_rel_vs_fp64"] = abs(out["mixed"]["ham_l2_t10"] / out["fp64"]["ham_l2_t10"] - 1.0)
    gpath = os.path.join(ROOT, "tests", "golden", "binary_hierarchy_t10.json")
    if os.path.exists(gpath):
        g = json.load(open(gpath))["telemetry"][-1]
        ref = (g["ham_sumsq"] / g["count"]) ** 0.5
        out["oracle_ham_l2_t10"] = ref
        out["fp64_ham_l2_rel_vs_oracle"] = abs(out["fp64"]["ham_l2_t10"] / ref - 1.0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-AMR-blocks section")
    ap.add_argument("--no-hierarchy", action="store_true", help="skip the binary-puncture hierarchy section")
    ap.add_argument("--no-noisy", action="store_true", help="skip the 192^3 noisy mixed-precision section")
    ap.add_argument("--no-tune", action="store_true", help="keep the default kernel variants (no autotuning)")
    ap.add_argument("--tune-table", metavar="PATH",
                    help="load this tuning table (written by --tune-save) instead of autotuning")
    ap.add_argument("--tune-save", metavar="PATH", help="write the tuning table in force at the end of the run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path (the level would shard "
              "per block across GPUs; see the sharding design). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import granite_amd as ga

    # variants: measured here (autotune), taken from a saved table, or the defaults
    args.autotune = not args.no_tune and not args.tune_table
    if args.tune_table:
        ga.tune_load(args.tune_table)
    eng = ga.Engine(0)
    p = ga.CCZ4Params(kappa1=0.02, kappa2=0.0, eta=2.0, ko_sigma=0.1)
    cells = N ** 3

    def run_form(precision):
        blk = ga.GridBlock(0, 0, (N, N, N), (-HALF,) * 3, (HALF,) * 3, 4)
        lvl = ga.Level([blk], eng)
        lvl.init_brill_lindquist([1.0], [[0.0, 0.0, 0.0]])
        tuning = None
        if args.autotune or args.tune_table:
            tuning = level_tuning(ga, lvl, lvl.autotune(p, DT) if args.autotune else None)
        step = dict(precision=precision, bc_mode=ga.BC_SOMMERFELD, pre_fill=ga.BC_OUTFLOW)
        for _ in range(args.warmup):
            lvl.rk3_step(p, DT, use_graph=True, **step)
        lvl.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            lvl.rk3_step(p, DT, use_graph=True, **step)
        lvl.sync()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        # per-kernel device times (HIP events on the engine stream, eager launches)
        lvl.profile(True)
        for k in range(3):
            lvl.profile_read(k)
        for _ in range(args.profile_steps):
            lvl.rk3_step(p, DT, use_graph=False, **step)
        prof = [lvl.profile_read(k) for k in range(3)]
        lvl.profile(False)
        r = {
            "ms_per_step": ms,
            "value": cells / (ms * 1e-3),
            "stage_avg_ms": prof[0][0] / max(prof[0][1], 1),
            "stage_launches": prof[0][1],
            "final_avg_ms": prof[1][0] / max(prof[1][1], 1),
            "final_launches": prof[1][1],
            "ghost_avg_ms": prof[2][0] / max(prof[2][1], 1),
            "nonfinite": lvl.nonfinite(),
            "tuning": tuning,
            "t": (args.warmup + args.steps + args.profile_steps) * DT,
        }
        sumsq, count = lvl.constraints_l2()
        r["ham_l2"] = (sumsq / count) ** 0.5 if count else None
        lvl.close()
        return r

    f64 = run_form(ga.PREC_FP64)
    mix = run_form(ga.PREC_MIXED)
    batched = None if args.no_batched else run_batched(ga, eng, p, args)
    hier = None if args.no_hierarchy else run_hierarchy(ga, eng, p, args)
    noisy = None if args.no_noisy else run_noisy_192(ga, eng, p, 100, tune=args.autotune or bool(args.tune_table),
                                                          autotune=args.autotune)
    value = f64["value"]
    ms_per_step = f64["ms_per_step"]

    # stage kernels of RK3 stages 1 and 2 alternate: stage 1 reads u (stencil) and writes
    # u1 = 352 B/cell; stage 2 reads u1 (stencil) and u^n and writes u2 = 528 B/cell
    stage_avg_ms = f64["stage_avg_ms"]
    n_stage = f64["stage_launches"]
    stage_bytes = (352 + 528) / 2 * cells
    final_avg_ms = f64["final_avg_ms"]
    n_final = f64["final_launches"]
    final_bytes = 528 * cells
    achieved = stage_bytes / (stage_avg_ms * 1e-3) / 1e9
    traffic = None
    compute = None
    tpath = os.path.join(ROOT, "profiles", "stage_kernel_traffic_bl128.json")
    if os.path.exists(tpath):
        prof = json.load(open(tpath))
        traffic = prof.get("hbm_bytes_per_launch")
        if prof.get("fp64_flop_per_launch") and prof.get("cells_per_launch") == cells:
            # the kernel's own FP64 arithmetic (rocprofv3 instruction counters of this
            # workload) at the live launch time, against the FP64 vector peak; the issue
            # floors are the launch times with every SIMD issuing a VALU instruction
            # (4 cycles per wave64 instruction) every cycle
            flop = prof["fp64_flop_per_launch"]
            tf = flop / (stage_avg_ms * 1e-3) / 1e12
            compute = {
                "bound": "valu",
                "achieved": tf,
                "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s",
                "frac": tf / PEAK_FP64_TFLOPS,
                "flop_per_cell": flop / cells,
                "valu_issue_floor_ms": prof["valu_issue_floor_ms"],
                "fp64_issue_floor_ms": prof["fp64_issue_floor_ms"],
                "valu_issue_frac": prof["valu_issue_floor_ms"] / stage_avg_ms,
                "hbm_frac_at_valu_floor": stage_bytes / (prof["valu_issue_floor_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
            }

    cpu = {"value": None, "unit": "cell-updates/s", "cores": 0, "kind": None, "sample": "skipped (--no-cpu)"}
    if not args.no_cpu:
        cpu = cpu_baseline()

    line = {
        "metric": "cell-updates/s per RK3 step on 1 GPU, FP64 vs mixed/batched, at matched H-constraint",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "single Brill-Lindquist puncture M=1, one 128^3 block on [-16,16]^3, ng=4, FP64 RK3 step",
            "block": [N, N, N],
            "ghosts": 4,
            "dx": 2 * HALF / N,
            "dt": DT,
            "precision": "fp64",
            "stage_tile": list(_tile(ga)),
            "tuning": f64.get("tuning"),
            "graph": True,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "stage_kernel<MODE_STAGE> (RK3 stages 1-2: RHS + KO + combine + floors)",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            # HBM-side bytes moved per second at the live launch time (from `traffic`)
            "traffic_GBs": traffic / (stage_avg_ms * 1e-3) / 1e9 if traffic else None,
            "algorithmic_bytes_per_launch": stage_bytes,
            "avg_launch_ms": stage_avg_ms,
            "launches": n_stage,
        },
        "compute": compute,
        "cpu_baseline": cpu,
        "step_algorithmic_GBs": BYTES_PER_CELL_STEP * value / 1e9,
        "kernels_ms": {
            "stage_1_2_avg": stage_avg_ms,
            "stage_3_avg": final_avg_ms,
            "stage_3_GBs": final_bytes / (final_avg_ms * 1e-3) / 1e9 if n_final else None,
            "ghost_fill_avg": f64["ghost_avg_ms"],
        },
        "checks": {"nonfinite": f64["nonfinite"], "ham_l2": f64["ham_l2"], "t": f64["t"]},
        "batched_amr_blocks": batched,
        "binary_hierarchy_t10": hier,
        "noisy_192_mixed": noisy,
        "forms": {
            "fp64": f64,
            "mixed": dict(mix, speedup_vs_fp64=mix["value"] / value,
                          ham_l2_rel_vs_fp64=abs(mix["ham_l2"] - f64["ham_l2"]) / f64["ham_l2"],
                          note="FP32 stencils/algebra on tiles with min chi >= 0.2, FP64 elsewhere; FP64 state, "
                               "combine, metric inverse and monitor; same HBM bytes as FP64 (1408 B/cell-step)"),
        },
    }
    if cpu.get("value"):
        line["gpu_over_cpu"] = value / cpu["value"]
    if noisy and (cpu.get("noisy_192") or {}).get("value"):
        noisy["gpu_over_cpu"] = noisy["fp64"]["value"] / cpu["noisy_192"]["value"]
    if args.tune_save:
        ga.tune_save(args.tune_save)
    print(json.dumps(line), flush=True)
    eng.close()


def _tile(ga):
    import ctypes

    t = (ctypes.c_int * 3)()
    ga.library().gr_stage_tile(0, t)
    return tuple(t)


if __name__ == "__main__":
    main()
