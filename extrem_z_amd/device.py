"""Device buffers through the HIP runtime that libextremz.so itself uses.

Loading libamdhip64.so.7 after libextremz.so returns the runtime the engine
is already bound to, so buffers, copies and synchronisation here and the
engine's kernels share one runtime (a second HIP runtime in the process, such
as the one a framework wheel bundles, would be a separate device context).
"""
from __future__ import annotations

import ctypes

import numpy as np

H2D, D2H, D2D = 1, 2, 3
_rt = None


def runtime():
    global _rt
    if _rt is None:
        from . import lib
        lib()  # the engine first: its runtime is the one to share
        rt = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        rt.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        rt.hipFree.argtypes = [vp]
        rt.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        rt.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        rt.hipDeviceSynchronize.argtypes = []
        rt.hipSetDevice.argtypes = [ctypes.c_int]
        rt.hipGetErrorString.restype = ctypes.c_char_p
        rt.hipGetErrorString.argtypes = [ctypes.c_int]
        _rt = rt
    return _rt


def _ok(rc, what):
    if rc != 0:
        msg = runtime().hipGetErrorString(rc).decode(errors="replace")
        raise RuntimeError(f"{what}: HIP error {rc} ({msg})")


def synchronize():
    _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """nbytes of device memory on the current device; freed by free() or GC."""

    def __init__(self, nbytes, dtype=np.uint8, shape=None):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = ctypes.c_void_p()
        _ok(runtime().hipMalloc(ctypes.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, a.dtype, a.shape)
        _ok(runtime().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return b

    @classmethod
    def zeros(cls, shape, dtype):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        b = cls(n, dtype, tuple(np.atleast_1d(shape)))
        _ok(runtime().hipMemset(b.ptr, 0, max(1, n)), "hipMemset")
        return b

    def empty_like(self):
        return DeviceBuffer(self.nbytes, self.dtype, self.shape)

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _ok(runtime().hipMemcpy(self.ptr, other.ptr, self.nbytes, D2D), "hipMemcpy D2D")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
