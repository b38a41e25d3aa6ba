"""Device buffers for tests that call the C ABI's device-pointer entry points.

Uses the HIP runtime that libumapx.so itself is linked against (ctypes on
libamdhip64.so.7, which resolves to the copy already loaded with the
library), so a test process never mixes two HIP runtimes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from umapers_amd import _lib as L

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        L.lib()  # load the library first: its HIP runtime is the one to use
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


_H2D, _D2H = 1, 2


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a shape)."""

    def __init__(self, shape, dtype, host: np.ndarray | None = None):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p
        if host is not None:
            src = np.ascontiguousarray(host, dtype=self.dtype)
            assert src.shape == self.shape
            _ok(hip().hipMemcpy(self.ptr, src.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    @classmethod
    def of(cls, host: np.ndarray) -> "DeviceArray":
        return cls(host.shape, host.dtype, host)

    @classmethod
    def filled(cls, shape, dtype, value) -> "DeviceArray":
        return cls(shape, dtype, np.full(shape, value, dtype=dtype))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
