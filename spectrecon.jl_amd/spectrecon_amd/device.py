"""Device-resident FP32 arrays allocated through the C-ABI (sr_malloc).

Shapes use the reference's index order (image [nx,ny,nz], views
[nx,nz,nview]); the memory is column-major, as numpy's order='F'.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().sr_malloc(C.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=np.float32):
        a = np.asfortranarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        check(lib().sr_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype, order="F")
        check(lib().sr_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, a):
        a = np.asfortranarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(lib().sr_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def fill_zero(self):
        check(lib().sr_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().sr_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def size(self):
        return int(np.prod(self.shape))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


class PinnedArray:
    """Page-locked host FP32 memory (sr_host_alloc) with a numpy view
    (`.array`, column-major): the source or target of asynchronous copies on a
    plan's stream (sr_plan_memcpy_async)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().sr_host_alloc(C.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value
        buf = (C.c_char * max(self.nbytes, 4)).from_address(self.ptr)
        self.array = np.frombuffer(buf, self.dtype, count=int(np.prod(self.shape))).reshape(self.shape, order="F")

    def free(self):
        if self.ptr:
            self.array = None
            lib().sr_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
