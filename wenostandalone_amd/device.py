"""Device buffers owned through the C-ABI (weno_device_alloc / free / copies)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import WenoError, check, last_error, lib


class DeviceArray:
    """A device allocation with a numpy dtype and a Fortran-order shape."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        ptr = lib.weno_device_alloc(max(self.nbytes, 1))
        if not ptr:
            raise WenoError(-2, last_error())
        self.ptr = int(ptr)

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.asarray(a)
        if dtype is not None:
            a = a.astype(dtype)
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.asfortranarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"size mismatch {a.size} != {self.size}")
        check(lib.weno_memcpy_to_device(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(lib.weno_memcpy_to_host(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes))
        return out

    def fill(self, value: float) -> None:
        self.upload(np.full(self.shape, value, dtype=self.dtype, order="F"))

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            lib.weno_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def vp(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)
