"""Device arrays on the HIP runtime that libcoral_amd.so itself links
(/opt/rocm's libamdhip64), through ctypes: allocation, host copies and
zero fills for the closure's training buffers.

A PyTorch ROCm build carries its own copy of the HIP and HSA runtimes; a
process that initialises both copies cannot run the library's kernels on
memory of the other, so the closure takes its device memory from here.
"""
from __future__ import annotations

import ctypes
import os
from typing import Tuple

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        h = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
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
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DevArray:
    """A device buffer of `n` elements of `dtype` ("f32" or "bf16"; bf16 holds
    16-bit patterns), freed with the object."""

    def __init__(self, n: int, dtype: str = "f32", zero: bool = False):
        self.n = int(n)
        self.dtype = dtype
        self.itemsize = 4 if dtype == "f32" else 2
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 4)), "hipMalloc")
        self._p = p.value
        if zero:
            self.zero()

    @property
    def nbytes(self) -> int:
        return self.n * self.itemsize

    @property
    def ptr(self) -> int:
        return self._p

    def at(self, offset_elems: int) -> int:
        return self._p + offset_elems * self.itemsize

    def zero(self):
        _ok(hip().hipMemset(self._p, 0, self.nbytes), "hipMemset")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DevArray":
        a = np.ascontiguousarray(a, np.float32)
        d = cls(a.size, "f32")
        _ok(hip().hipMemcpy(d._p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return d

    def upload(self, a: np.ndarray, offset_elems: int = 0):
        a = np.ascontiguousarray(a, np.float32 if self.dtype == "f32" else np.uint16)
        if offset_elems + a.size > self.n:
            raise ValueError("upload beyond the buffer")
        _ok(hip().hipMemcpy(self.at(offset_elems), a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def copy_from(self, src_ptr: int, nelem: int, offset_elems: int = 0):
        """Device-to-device copy of nelem elements from src_ptr to this
        buffer at offset_elems (the null stream's order)."""
        if offset_elems + nelem > self.n:
            raise ValueError("copy beyond the buffer")
        _ok(hip().hipMemcpy(self.at(offset_elems), src_ptr, nelem * self.itemsize, 3), "hipMemcpy D2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, np.float32 if self.dtype == "f32" else np.uint16)
        _ok(hip().hipMemcpy(out.ctypes.data, self._p, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def __del__(self):
        p = getattr(self, "_p", None)
        if p and _hip is not None:
            _hip.hipFree(p)
            self._p = None
