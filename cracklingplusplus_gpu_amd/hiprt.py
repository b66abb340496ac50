"""Minimal ctypes plumbing over the HIP runtime (device buffers, copies, sync).

Used by the benchmark and tests to keep inputs resident in HBM and to call
issl_score_device; the scorer itself never goes through here.
"""
from __future__ import annotations

import ctypes
from ctypes import c_int, c_size_t, c_void_p, byref

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
        h.hipFree.argtypes = [c_void_p]
        h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        h.hipMemset.argtypes = [c_void_p, c_int, c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [c_int]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(c_void_p)]
        h.hipStreamSynchronize.argtypes = [c_void_p]
        h.hipStreamDestroy.argtypes = [c_void_p]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(c_size_t), ctypes.POINTER(c_size_t)]
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device buffer holding a copy of a numpy array's shape and dtype."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = c_void_p()
        _ok(hip().hipMalloc(byref(self.ptr), max(self.nbytes, 256)), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data_as(c_void_p), d.nbytes, H2D), "hipMemcpy H2D")
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(c_void_p), self.ptr, self.nbytes, D2H),
            "hipMemcpy D2H")
        return out

    def fill_bytes(self, value: int) -> None:
        _ok(hip().hipMemset(self.ptr, value, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority (hipStreamCreate)."""

    def __init__(self):
        self.handle = c_void_p()
        _ok(hip().hipStreamCreate(byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            _ok(hip().hipStreamDestroy(self.handle), "hipStreamDestroy")
        self.handle = None


def copy_device(dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
    """Device-to-device copy of nbytes (the runtime's own copy kernel)."""
    _ok(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def set_device(d: int):
    _ok(hip().hipSetDevice(d), "hipSetDevice")


def mem_info():
    free, total = c_size_t(), c_size_t()
    _ok(hip().hipMemGetInfo(byref(free), byref(total)), "hipMemGetInfo")
    return free.value, total.value
