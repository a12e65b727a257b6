"""Times the CPU path on the host cores (bench.py's cpu_baseline leg; test infrastructure).

Run as a subprocess so that the OpenMP settings of the reference launcher
(scripts/run_granite.py:138-149: OMP_NUM_THREADS = cores, OMP_PROC_BIND=close,
OMP_PLACES=cores) are in place before the OpenMP runtime starts. Never touches a GPU.

kind "reference": the compiled GRANITE sources (oracle/_ref/libgranite_ref_fast.so, built
with the reference Release flags) evaluate CCZ4Evolution::computeRHS (dissipation
included) three times on the block, i.e. the RHS work of one RK3 step; the cheaper
combine / floor / ghost passes of the step are not included, which can only make the CPU
rate look better. kind "port": the oracle's own full RK3 step (same flags), used when the
compiled reference is absent.

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from oracle.oracle import Geom, Oracle, Reference  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--half-width", type=float, default=16.0)
    ap.add_argument("--steps", type=int, default=5, help="RK3 steps' worth of RHS work per timed run")
    ap.add_argument("--runs", type=int, default=3, help="timed runs; the median is reported")
    ap.add_argument("--noise", type=float, default=0.0,
                    help="amplitude of the noise added to the conformal metric and K (0: none)")
    ap.add_argument("--seed", type=int, default=20240611)
    args = ap.parse_args()
    n = args.n
    w = args.half_width
    g = Geom((n, n, n), 4, (-w,) * 3, (w,) * 3)
    port = Oracle(fast=True)
    u = port.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    block = f"{n}^3 Brill-Lindquist block"
    if args.noise > 0:
        port.add_noise(g, u, args.seed, args.noise, [1, 2, 3, 4, 5, 6, 13])
        block += f" with {args.noise:g} noise on the conformal metric and K"
    threads = port.num_threads()
    times = []
    if Reference.available(fast=True):
        ref = Reference(fast=True)
        rhs = g.zeros()
        ref.rhs(g, u, rhs)  # warm-up (threads, pages)
        for _ in range(args.runs):
            t0 = time.perf_counter()
            for _ in range(3 * args.steps):
                ref.rhs(g, u, rhs)
            times.append((time.perf_counter() - t0) / args.steps)
        kind = "reference"
        sample = (f"median of {args.runs} runs, each {args.steps} RK3 steps' RHS work (3 x computeRHS + "
                  f"applyDissipation per step) of the compiled reference on the {block}, after "
                  f"one warm-up evaluation; the combine / floor / ghost passes of a step are excluded")
    else:
        stage = g.zeros()
        rhs = g.zeros()
        port.rk3_step(g, u.copy(), 0.0625, 0, stage=stage, rhs=rhs)  # warm-up
        for _ in range(args.runs):
            t0 = time.perf_counter()
            for _ in range(args.steps):
                port.rk3_step(g, u, 0.0625, 0, stage=stage, rhs=rhs)
            times.append((time.perf_counter() - t0) / args.steps)
        kind = "port"
        sample = f"median of {args.runs} runs of {args.steps} full SSP-RK3 steps of the oracle on the {block}"
    dt = sorted(times)[len(times) // 2]
    cells = n ** 3
    print(json.dumps({"value": cells / dt, "unit": "cell-updates/s", "cores": threads, "kind": kind,
                      "sample": sample, "seconds_per_step": dt,
                      "runs_cell_updates_per_s": [cells / t for t in times]}))


if __name__ == "__main__":
    main()
