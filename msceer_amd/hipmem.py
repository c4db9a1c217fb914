"""Minimal device-memory helpers over the HIP runtime that libmsceer_hip.so links.

PyTorch ships its own HIP runtime; mixing it with /opt/rocm's in one process
gives two runtimes, so callers that hand device pointers to the C ABI from
Python allocate them here, from the same runtime instance as the library.
"""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None
_H2D, _D2H, _D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        from . import native
        native.lib()  # loads libamdhip64.so.7 as a dependency first
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [ctypes.c_int]
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class DeviceBuffer:
    """An owned hipMalloc allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def addr(self) -> int:
        return self.ptr.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _H2D), "hipMemcpy H2D")

    def upload_at(self, offset: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(ctypes.c_void_p(self.ptr.value + offset), a.ctypes.data_as(ctypes.c_void_p), a.nbytes,
                            _H2D), "hipMemcpy H2D")

    def download_at(self, offset: int, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr.value + offset), out.nbytes,
                            _D2H), "hipMemcpy D2H")
        return out

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def fill(self, byte: int):
        _ok(hip().hipMemset(self.ptr, byte, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
