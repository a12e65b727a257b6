"""Writes tests/golden/binary_hierarchy_t10.json: the strict-IEEE oracle's run of the 4-level
binary-puncture hierarchy to t = 10 M (test fixture generator; run on the CPU).

Geometry: 96^3 on [-64, 64]^3 plus the three tracking-sphere levels the reference's regrid
builds for Brill-Lindquist punctures of mass 0.5 at (+-3, 0, 0) (granite_amd/hierarchy.py).
Initial data as src/main.cpp:962-1134 sets it up: Brill-Lindquist on level 0, finer levels
prolongated from their parents, then fillGhostZones(0). Then 30 coarse steps of dt0 =
0.25 * dx0 = 1/3 through the oracle's restatement of AMRHierarchy::subcycle (with the work
the reference overwrites left out, which is bit-identical). Recorded: the constraint
telemetry of src/main.cpp:1216-1250 at steps 1, 10, 20 and 30 (sum of H^2, sum of |M|^2,
cell count over all blocks' interiors), and per level and field the sum of squares and the
largest magnitude of the interior state at t = 10.

    python tests/golden/make_hierarchy_golden.py      (about 6 minutes, one core)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from granite_amd.hierarchy import PUNCTURES, binary_puncture_levels  # noqa: E402
from oracle.hierarchy import OracleHierarchy, build_levels  # noqa: E402
from oracle.oracle import DEFAULT_PARAMS, Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "binary_hierarchy_t10.json")
RECORD = (1, 10, 20, 30)


def main():
    o = Oracle()
    geo = binary_puncture_levels()
    masses, pos = PUNCTURES
    levels = build_levels(o, geo, masses=masses, positions=pos)
    H = OracleHierarchy(o, levels)
    H.fill_ghosts(0)
    prm = dict(DEFAULT_PARAMS)
    dt0 = 0.25 * float(levels[0][0][0].dx[0])
    t = 0.0
    telemetry = []
    t0 = time.time()
    for step in range(1, 31):
        H.step(dt0, prm, skip_overwritten=True)
        t += dt0
        if step in RECORD:
            sh, sm, n = H.constraints_l2()
            telemetry.append(dict(step=step, t=t, ham_sumsq=sh, mom_sumsq=sm, count=n))
        print(f"step {step} t={t:.4f} {time.time() - t0:.0f}s", flush=True)
    fields = []
    for lv in levels:
        g, u = lv[0]
        s = g.ng
        inner = u[:, s:s + g.n[2], s:s + g.n[1], s:s + g.n[0]]
        fields.append(dict(sumsq=[float(np.sum(inner[v] ** 2)) for v in range(22)],
                           maxabs=[float(np.max(np.abs(inner[v]))) for v in range(22)]))
    out = dict(geometry=geo, dt0=dt0, params=prm, masses=masses, positions=pos, steps=30, t_final=t,
               telemetry=telemetry, levels=fields)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
