"""Regenerates the inputs of the golden fixtures (tests/golden/*.npz) on any machine."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from oracle.oracle import Geom

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NOISE_FIELDS = [1, 2, 3, 4, 5, 6, 13, 19, 20, 21]
NAMES = ["bl_puncture", "bl_puncture_noise", "bl_binary_noise8", "gauge_wave"]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    rec = json.loads(str(z["recipe"]))
    return rec, z


def make_input(O, rec):
    g = Geom(rec["n"], rec["ng"], rec["lo"], rec["hi"])
    if rec["kind"] == "bl":
        u = O.init_bl(g, rec["masses"], rec["positions"])
    else:
        u = O.init_gauge_wave(g, rec["A"], rec["L"])
    if rec.get("noise", 0.0) > 0.0:
        O.add_noise(g, u, rec["seed"], rec["noise"], rec.get("noise_fields", NOISE_FIELDS))
    return g, u


def sources(g: Geom, seed: int):
    N = g.stride
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(N) * 1e-3, rng.standard_normal(3 * N) * 1e-3, rng.standard_normal(6 * N) * 1e-3,
            rng.standard_normal(N) * 1e-3)
