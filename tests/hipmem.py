"""Device buffers for the GPU tests, through the HIP runtime the engine itself links
(/opt/rocm/lib/libamdhip64.so): hipMalloc, hipMemcpy, hipFree over ctypes.

The tests hand device pointers to entry points that borrow them (matter sources); they use
this instead of a framework allocator so that the process holds one HIP runtime only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB = None
H2D, D2H = 1, 2


def _hip():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        _LIB = C.CDLL(path)
        _LIB.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _LIB.hipFree.argtypes = [C.c_void_p]
        _LIB.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _LIB.hipDeviceSynchronize.argtypes = []
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """A device copy of a float64 host array; freed by close() or when collected."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.nbytes = host.nbytes
        self.shape = host.shape
        self._ptr = C.c_void_p()
        L = _hip()
        _check(L.hipMalloc(C.byref(self._ptr), max(self.nbytes, 8)), "hipMalloc")
        _check(L.hipMemcpy(self._ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def data_ptr(self) -> int:
        return int(self._ptr.value)

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(_hip().hipMemcpy(out.ctypes.data, self._ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def close(self):
        if self._ptr is not None and self._ptr.value:
            _hip().hipFree(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
