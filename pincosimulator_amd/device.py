"""Device buffers for Python callers of the allocation-free entry point
(``twmc_advance_noalloc``), allocated through the same HIP runtime that
libtwmc_hip.so is bound to (ctypes on the runtime library the process has
loaded), so a process never drives the GPU through two runtimes.  A process
that loads PyTorch's ROCm build before libtwmc_hip.so binds the library to the
HIP runtime bundled with PyTorch instead of the system one; that is why the
tests and the benchmark use these buffers rather than torch tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_hip = None
H2D, D2H, D2D = 1, 2, 3  # hipMemcpyKind


class HipError(RuntimeError):
    pass


def runtime():
    """The HIP runtime library libtwmc_hip.so is bound to in this process."""
    global _hip
    if _hip is None:
        _lib.load()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise HipError("libtwmc_hip.so is loaded but no HIP runtime library is mapped")
        h = C.CDLL(path)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice"):
            getattr(h, name).restype = C.c_int
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        msg = runtime().hipGetErrorString(rc)
        raise HipError(f"{what}: {msg.decode() if msg else rc}")


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """nbytes of device memory on the current device (hipMalloc / hipFree)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.write(a)
        return b

    @property
    def ptr(self):
        return self._ptr

    def write(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(runtime().hipMemcpy(self._ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def read(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read larger than the buffer")
        _check(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other, nbytes=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        _check(runtime().hipMemcpy(self._ptr, other.ptr, n, D2D), "hipMemcpy D2D")

    def free(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            runtime().hipFree(self._ptr)
        self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
