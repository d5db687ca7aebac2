"""Device context and device arrays, managed through librbffd's own HIP calls."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


class Context:
    """One librbffd context: a device, a stream, a workspace.  Not shared
    between threads."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        self._h = ctypes.c_void_p()
        check(L.rbffd_create(int(device), ctypes.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.load().rbffd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(_lib.load().rbffd_synchronize(self._h))

    def timer_start(self):
        check(_lib.load().rbffd_timer_start(self._h))

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        check(_lib.load().rbffd_timer_stop(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def info(self):
        name = ctypes.create_string_buffer(256)
        arch = ctypes.create_string_buffer(64)
        cus = ctypes.c_int32()
        check(_lib.load().rbffd_device_info(self._h, name, 256, ctypes.byref(cus), arch, 64))
        return {"name": name.value.decode(), "compute_units": cus.value, "arch": arch.value.decode()}

    # -- arrays
    def empty(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.copy_from_host(a)
        return d

    def zeros(self, shape, dtype):
        d = DeviceArray(self, shape, dtype)
        check(_lib.load().rbffd_memset_device(self._h, d.ptr, 0, d.nbytes))
        return d


class DeviceArray:
    """A contiguous device allocation with a numpy shape and dtype."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(_lib.load().rbffd_device_malloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self._p = p

    @property
    def ptr(self):
        return self._p

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def copy_from_host(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(_lib.load().rbffd_memcpy_h2d(self.ctx.handle, self._p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def copy_from_device(self, other: "DeviceArray"):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(_lib.load().rbffd_memcpy_d2d(self.ctx.handle, self._p, other.ptr, self.nbytes))

    def zero_(self):
        check(_lib.load().rbffd_memset_device(self.ctx.handle, self._p, 0, self.nbytes))
        return self

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(_lib.load().rbffd_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), self._p,
                                           self.nbytes))
        return out

    def free(self):
        if self._p:
            _lib.load().rbffd_device_free(self.ctx.handle, self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
