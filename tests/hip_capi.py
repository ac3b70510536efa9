"""Thin helpers to call libnbody_hip.so's C ABI directly from the tests."""
from __future__ import annotations

import ctypes as C

import numpy as np

from nbody_amd import native


class Ctx:
    def __init__(self, device: int = 0):
        self.lib = native.hip_lib()
        self.h = C.c_void_p()
        native.check(self.lib.nbh_ctx_create(device, C.byref(self.h)), "nbh_ctx_create")
        self.bufs = []

    def close(self):
        for b in self.bufs:
            if b.ptr:
                self.lib.nbh_free(self.h, b.ptr)
                b.ptr = None
        self.bufs = []
        if self.h:
            native.check(self.lib.nbh_ctx_destroy(self.h), "nbh_ctx_destroy")
            self.h = None

    def alloc(self, n: int) -> "DevArray":
        b = DevArray(self, n)
        self.bufs.append(b)
        return b

    def put(self, a: np.ndarray) -> "DevArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = self.alloc(a.size)
        b.write(a)
        return b

    def sync(self):
        native.check(self.lib.nbh_ctx_synchronize(self.h), "sync")

    def put_raw(self, a: np.ndarray) -> "DevArray":
        """Device copy of an array of any dtype (read back with read_raw)."""
        a = np.ascontiguousarray(a)
        b = self.alloc((a.nbytes + 7) // 8)
        native.check(self.lib.nbh_memcpy_h2d(self.h, b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), "h2d")
        return b

    def zeros_raw(self, dtype, count: int) -> "DevArray":
        return self.put_raw(np.zeros(count, dtype))


class DevArray:
    def __init__(self, ctx: Ctx, n: int):
        self.ctx = ctx
        self.n = n
        p = C.c_void_p()
        native.check(ctx.lib.nbh_malloc(ctx.h, max(8, 8 * n), C.byref(p)), "nbh_malloc")
        self.ptr = p

    def write(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.n
        native.check(self.ctx.lib.nbh_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes),
                     "h2d")

    def read(self) -> np.ndarray:
        out = np.empty(self.n)
        native.check(self.ctx.lib.nbh_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n),
                     "d2h")
        return out

    def read_raw(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= 8 * self.n
        native.check(self.ctx.lib.nbh_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes),
                     "d2h")
        return out

    def at(self, offset: int):
        """Device pointer offset by `offset` doubles."""
        return C.c_void_p(self.ptr.value + 8 * offset)


def ptr_array(ptrs):
    arr = (C.c_void_p * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = p.value if isinstance(p, C.c_void_p) else p
    return arr


def dbl_array(vals):
    a = np.ascontiguousarray(vals, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))
