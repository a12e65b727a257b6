"""Build recipe of the HIP engine: gfx950 code objects + C ABI in one shared library.

    python -m granite_amd.build        (or granite_amd.build.build())

The library is written in-tree as granite_amd/libgranite_ccz4.so.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgranite_ccz4.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kernels.hip", "capi.hip")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("kernels.hpp", "ccz4_cell.hpp")] + [
    os.path.join(ROOT, "include", "granite_ccz4.h")
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions stay disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile every source into `out` (default: the in-tree library); always recompiles.
    `defines` are extra -D macros (tile shape and prefetch depth of the stage kernel)."""
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *["-D" + d for d in defines], "-I", os.path.join(ROOT, "include"), *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
