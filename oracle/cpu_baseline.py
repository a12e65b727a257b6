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
    ap.add_argument("--reps", type=int, default=3, help="RHS evaluations timed (3 = one RK3 step's RHS work)")
    args = ap.parse_args()
    n = args.n
    w = args.half_width
    g = Geom((n, n, n), 4, (-w,) * 3, (w,) * 3)
    port = Oracle(fast=True)
    u = port.init_bl(g, [1.0], [[0.0, 0.0, 0.0]])
    threads = port.num_threads()
    if Reference.available(fast=True):
        ref = Reference(fast=True)
        gw = Geom((32, 32, 32), 4, (-4.0,) * 3, (4.0,) * 3)
        ref.rhs(gw, port.init_bl(gw, [1.0], [[0.0, 0.0, 0.0]]))  # warm-up (threads, pages)
        rhs = g.zeros()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            ref.rhs(g, u, rhs)
        dt = (time.perf_counter() - t0) * 3 / args.reps
        kind = "reference"
        scaled = "" if args.reps == 3 else ", time scaled to 3"
        sample = (f"{args.reps} x computeRHS+applyDissipation of the compiled reference on the {n}^3 Brill-Lindquist "
                  f"block{scaled} (the RHS work of one RK3 step; combine/floor/ghost passes excluded)")
    else:
        stage = g.zeros()
        rhs = g.zeros()
        port.rk3_step(Geom((32, 32, 32), 4, (-4.0,) * 3, (4.0,) * 3),
                      port.init_bl(Geom((32, 32, 32), 4, (-4.0,) * 3, (4.0,) * 3), [1.0], [[0, 0, 0]]), 0.0625)
        t0 = time.perf_counter()
        port.rk3_step(g, u, 0.0625, 0, stage=stage, rhs=rhs)
        dt = time.perf_counter() - t0
        kind = "port"
        sample = f"one full SSP-RK3 step of the oracle on the {n}^3 Brill-Lindquist block"
    cells = n ** 3
    print(json.dumps({"value": cells / dt, "unit": "cell-updates/s", "cores": threads, "kind": kind,
                      "sample": sample, "seconds": dt}))


if __name__ == "__main__":
    main()
