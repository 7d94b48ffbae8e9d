"""MI355X (gfx950) compensated BLAS1 + Krylov library -- Python view of the C ABI.

The product is ``libhpb.so`` (HIP kernels + C ABI, ``include/hpb.h``) and the C++
templates under ``cpp/`` that mirror the reference's VectorOperations /
linear_operator / solver interfaces.  This module is a thin ctypes layer used by
the tests and the benchmark: device buffers, a handle, and typed wrappers that
return numpy values.  Every call goes to the GPU library; nothing here computes.
"""
import ctypes

import numpy as np

from ._lib import HPBError, call, lib, prototypes  # noqa: F401

__all__ = ["Handle", "DeviceArray", "HPBError", "lib", "call", "prototypes"]

_SUFFIX = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}


def suffix(dtype):
    return _SUFFIX[np.dtype(dtype)]


class Handle:
    """Owns an hpb handle (one HIP stream plus reduction workspace)."""

    def __init__(self, stream=None):
        self._h = ctypes.c_void_p()
        call("hpb_create", ctypes.byref(self._h), stream)

    @property
    def ptr(self):
        return self._h

    def close(self):
        if self._h:
            lib().hpb_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        call("hpb_synchronize", self._h)

    def num_cu(self):
        n = ctypes.c_int()
        call("hpb_device_info", self._h, ctypes.byref(n), None)
        return n.value

    def set_tuning(self, kind, dtype, block, unroll, grid_max, flags=0):
        call("hpb_set_tuning", self._h, kind, 0 if np.dtype(dtype) == np.float64 else 1, block, unroll, grid_max,
             flags)

    def get_tuning(self, kind, dtype):
        b, u, g, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("hpb_get_tuning", self._h, kind, 0 if np.dtype(dtype) == np.float64 else 1,
             ctypes.byref(b), ctypes.byref(u), ctypes.byref(g), ctypes.byref(f))
        return b.value, u.value, g.value, f.value

    def reduce_config(self, dtype, n, aligned=True, binary=True):
        """(block, unroll, grid, vector width) of a reduction; binary: two-input dot."""
        cfg = (ctypes.c_int * 4)()
        call("hpb_reduce_config", self._h, 0 if np.dtype(dtype) == np.float64 else 1, n, int(aligned), int(binary),
             cfg)
        return tuple(cfg)

    def timer_start(self):
        call("hpb_timer_start", self._h)

    def timer_stop(self):
        ms = ctypes.c_float()
        call("hpb_timer_stop", self._h, ctypes.byref(ms))
        return ms.value


class DeviceArray:
    """A device allocation of n elements of float64 / float32 / other numpy dtype."""

    def __init__(self, n, dtype=np.float64, handle=None, offset_bytes=0, _base=None):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.handle = handle
        self._base = _base
        if _base is None:
            p = ctypes.c_void_p()
            call("hpb_malloc", ctypes.byref(p), max(self.nbytes, 16) + 256)
            self._alloc = p.value
            self.addr = p.value + offset_bytes
        else:
            self._alloc = None
            self.addr = _base.addr + offset_bytes

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @property
    def ptr(self):
        return ctypes.c_void_p(self.addr)

    @classmethod
    def from_numpy(cls, arr, handle, offset_bytes=0):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype, handle, offset_bytes=offset_bytes)
        d.upload(arr)
        return d

    def view(self, start, count):
        """Sub-vector sharing this allocation (may be unaligned)."""
        return DeviceArray(count, self.dtype, self.handle, offset_bytes=start * self.dtype.itemsize, _base=self)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size == self.n
        call("hpb_memcpy_h2d", self.handle.ptr, self.ptr, arr.ctypes.data_as(ctypes.c_void_p), self.nbytes)

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        call("hpb_memcpy_d2h", self.handle.ptr, out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes)
        return out

    def free(self):
        if self._alloc:
            lib().hpb_free(ctypes.c_void_p(self._alloc))
            self._alloc = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _out4(dtype):
    return np.zeros(4, dtype=dtype)


def reduce(name, h, x, y=None):
    """Run a host-result reduction ``hpb_<name>_<f64|f32>``; returns out[4] as numpy."""
    out = _out4(x.dtype)
    args = [h.ptr, x.n, x.ptr]
    if y is not None:
        args.append(y.ptr)
    args.append(out.ctypes.data_as(ctypes.c_void_p))
    call("hpb_%s_%s" % (name, suffix(x.dtype)), *args)
    return out


def dot2(h, x, y):
    return reduce("dot2", h, x, y)


def dot(h, x, y):
    return reduce("dot", h, x, y)


def sum2(h, x):
    return reduce("sum2", h, x)


def asum2(h, x):
    return reduce("asum2", h, x)


def nrm2c(h, x):
    return reduce("nrm2c", h, x)
