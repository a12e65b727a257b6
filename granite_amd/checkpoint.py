"""Reader of the checkpoint files granite_main writes.

A checkpoint holds what HDF5Writer::writeCheckpoint stores (src/io/hdf5_io.cpp:166-232):
the time, the step, and every block of the hierarchy (coarsest level first) with each
variable's whole storage, ghosts included (GridBlock::rawData). Without HDF5 the driver
writes it as one flat little-endian file:

    8 bytes  "GRCKPT01"
    int32    step
    int32    num_blocks
    float64  time
    per block:
      int32    level, ng, n[0], n[1], n[2], nvars
      float64  lo[3], dx[3]
      float64  nvars * stride values, var v at [v * stride, (v + 1) * stride),
               stride = prod(n[d] + 2 ng), x fastest
"""
from __future__ import annotations

import dataclasses

import numpy as np

MAGIC = b"GRCKPT01"


@dataclasses.dataclass
class CheckpointBlock:
    level: int
    ng: int
    n: tuple
    lo: np.ndarray
    dx: np.ndarray
    data: np.ndarray  # (nvars, nz + 2 ng, ny + 2 ng, nx + 2 ng)


@dataclasses.dataclass
class CheckpointData:
    step: int
    time: float
    blocks: list


def read_checkpoint(path) -> CheckpointData:
    raw = memoryview(open(path, "rb").read())
    if bytes(raw[:8]) != MAGIC:
        raise ValueError(f"{path}: not a granite checkpoint")
    pos = 8
    step, nblocks = np.frombuffer(raw, "<i4", 2, pos)
    pos += 8
    (time,) = np.frombuffer(raw, "<f8", 1, pos)
    pos += 8
    blocks = []
    for _ in range(int(nblocks)):
        level, ng, nx, ny, nz, nv = (int(v) for v in np.frombuffer(raw, "<i4", 6, pos))
        pos += 24
        lo = np.frombuffer(raw, "<f8", 3, pos).copy()
        dx = np.frombuffer(raw, "<f8", 3, pos + 24).copy()
        pos += 48
        shape = (nv, nz + 2 * ng, ny + 2 * ng, nx + 2 * ng)
        count = int(np.prod(shape))
        if pos + 8 * count > len(raw):
            raise ValueError(f"{path}: truncated block data")
        data = np.frombuffer(raw, "<f8", count, pos).reshape(shape).copy()
        pos += 8 * count
        blocks.append(CheckpointBlock(level, ng, (nx, ny, nz), lo, dx, data))
    if pos != len(raw):
        raise ValueError(f"{path}: {len(raw) - pos} trailing bytes")
    return CheckpointData(int(step), float(time), blocks)
