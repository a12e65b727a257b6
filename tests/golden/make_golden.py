"""Writes the golden fixtures of tests/golden/*.npz from the compiled GRANITE reference.

Runs only where /root/reference exists (it builds oracle/_ref from the reference sources
with oracle/Makefile). Each fixture holds: the block geometry, the recipe of its input
(built with the oracle's initial-data and seeded-noise routines, whose own agreement with
the reference is tested separately), the SHA-256 of the input bytes, and the reference's
outputs on interior cells. Inputs are regenerated at test time and checked against the
stored hash before the outputs are compared.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Geom, Oracle, Reference, build  # noqa: E402

NOISE_FIELDS = [1, 2, 3, 4, 5, 6, 13, 19, 20, 21]


def make_input(O: Oracle, rec: dict):
    g = Geom(rec["n"], rec["ng"], rec["lo"], rec["hi"])
    if rec["kind"] == "bl":
        u = O.init_bl(g, rec["masses"], rec["positions"])
    elif rec["kind"] == "gauge":
        u = O.init_gauge_wave(g, rec["A"], rec["L"])
    else:
        raise ValueError(rec["kind"])
    if rec.get("noise", 0.0) > 0.0:
        O.add_noise(g, u, rec["seed"], rec["noise"], rec.get("noise_fields", NOISE_FIELDS))
    return g, u


def sources(g: Geom, seed: int):
    N = g.stride
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(N) * 1e-3, rng.standard_normal(3 * N) * 1e-3, rng.standard_normal(6 * N) * 1e-3,
            rng.standard_normal(N) * 1e-3)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


CASES = {
    # (i) BL puncture, 8^3 block that straddles the puncture (chi from ~0.03 to ~0.4)
    "bl_puncture": dict(kind="bl", n=[8, 8, 8], ng=4, lo=[-1.0, -1.0, -1.0], hi=[1.0, 1.0, 1.0], masses=[1.0],
                        positions=[[0.0, 0.0, 0.0]], ops=["rhs", "constraints"]),
    # (ii) the same block with seeded noise on the metric, K and the shift
    "bl_puncture_noise": dict(kind="bl", n=[8, 8, 8], ng=4, lo=[-1.0, -1.0, -1.0], hi=[1.0, 1.0, 1.0], masses=[1.0],
                              positions=[[0.0, 0.0, 0.0]], noise=1e-3, seed=20240611,
                              ops=["rhs", "rhs_nodiss", "ko", "constraints", "rhs_sources"]),
    # two unequal punctures, anisotropic block, tiny noise (1e-8) as in the mixed-precision case
    "bl_binary_noise8": dict(kind="bl", n=[10, 8, 6], ng=4, lo=[-2.5, -1.0, -0.75], hi=[2.5, 1.0, 0.75],
                             masses=[0.5, 0.3], positions=[[-1.1, 0.05, 0.0], [1.3, -0.1, 0.07]], noise=1e-8,
                             seed=7, noise_fields=[1, 2, 3, 4, 5, 6, 13], ops=["rhs", "constraints"]),
    # (iii) reference gauge-wave data, 16 x 4 x 4
    "gauge_wave": dict(kind="gauge", n=[16, 4, 4], ng=4, lo=[0.0, 0.0, 0.0], hi=[1.0, 0.25, 0.25], A=0.1, L=1.0,
                       ops=["rhs", "constraints"]),
}


def main():
    build()
    O = Oracle()
    R = Reference()
    for name, rec in CASES.items():
        g, u = make_input(O, rec)
        out = {"input_sha256": sha(u)}
        I = g.interior()
        for op in rec["ops"]:
            if op == "rhs":
                out["rhs"] = R.rhs(g, u)[I]
            elif op == "rhs_nodiss":
                out["rhs_nodiss"] = R.rhs(g, u, ko_sigma=0.0)[I]
            elif op == "ko":
                out["ko"] = R.ko(g, u, g.zeros(), 0.1)[I]
            elif op == "rhs_sources":
                out["rhs_sources"] = R.rhs(g, u, sources=sources(g, 99))[I]
            elif op == "constraints":
                h, m = R.constraints(g, u)
                out["ham"] = h[I[1:]]
                out["mom"] = m[I[1:]]
        # ID of the reference itself (BrillLindquist::apply / setGaugeWaveData) for this block
        if rec["kind"] == "bl":
            out["init"] = R.init_bl(g, rec["masses"], rec["positions"])
        else:
            out["init"] = R.init_gauge_wave(g, rec["A"], rec["L"])
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, recipe=json.dumps(rec), **out)
        print(name, os.path.getsize(path), "bytes")
    # AMR transfer on a two-level toy (amr.cpp:484-626)
    gc = Geom((8, 8, 8), 4, (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0))
    gf = Geom((8, 8, 8), 4, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    uc = O.add_noise(gc, O.init_bl(gc, [1.0], [[0.1, -0.05, 0.02]]), 3, 1e-2, list(range(22)))
    uf = O.add_noise(gf, O.init_bl(gf, [1.0], [[0.1, -0.05, 0.02]]), 4, 1e-2, list(range(22)))
    prol = R.prolongate(gc, uc, gf, np.zeros_like(uf))
    rest = R.restrict(gf, uf, gc, uc.copy())
    np.savez_compressed(os.path.join(HERE, "amr_transfer.npz"), uc_sha256=sha(uc), uf_sha256=sha(uf),
                        prolongated=prol[:4], restricted=rest[:4])
    print("amr_transfer", os.path.getsize(os.path.join(HERE, "amr_transfer.npz")), "bytes")


if __name__ == "__main__":
    main()
