"""Vectors longer than 2^31 elements (64-bit indexing), FP64, n = 2^31 + 3:
x_i = 1 below index 2^31 and 2 from there on, y_i = 1, so that a dropped,
repeated or wrapped (i mod 2^31) element changes every result.  All sums are
integers below 2^53, so the exact values are known in closed form:
x.y = 2^31 + 2*3, sum x = x.y, sum |x| = x.y, ||y||^2 = n; after y := 2x + y,
x.y = sum 2 x_i^2 + x_i.  Needs 2 x 16 GiB of device memory."""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
from conftest import say

pytestmark = pytest.mark.gpu

N = (1 << 31) + 3
SPLIT = 1 << 31


def test_reductions_and_axpy_beyond_2p31(handle):
    say("large n=%d" % N)
    x = H.DeviceArray(N, np.float64, handle)
    y = H.DeviceArray(N, np.float64, handle)
    try:
        H.call("hpb_fill_f64", handle.ptr, N, ctypes.c_double(1.0), x.ptr)
        H.call("hpb_fill_f64", handle.ptr, N - SPLIT, ctypes.c_double(2.0), ctypes.c_void_p(x.addr + 8 * SPLIT))
        H.call("hpb_fill_f64", handle.ptr, N, ctypes.c_double(1.0), y.ptr)
        dot = SPLIT + 2 * (N - SPLIT)
        for name, args, want in (("dot2", (x, y), dot), ("dot", (x, y), dot), ("sum2", (x,), dot),
                                 ("asum2", (x,), dot), ("nrm2c", (y,), N)):
            out = H.reduce(name, handle, *args)
            assert out[2] == float(want), (name, out[2], want)
        ex = np.zeros(2)
        H.call("hpb_exact_dot_f64", handle.ptr, N, x.ptr, y.ptr, ex.ctypes.data)
        assert (ex[0], ex[1]) == (float(dot), 0.0)
        H.call("hpb_axpy_f64", handle.ptr, N, ctypes.c_double(2.0), x.ptr, y.ptr)   # y = 2x + 1: 3 or 5
        out = H.dot2(handle, x, y)
        assert out[2] == float(3 * SPLIT + 2 * 5 * (N - SPLIT))
        tail = np.zeros(4)
        H.call("hpb_memcpy_d2h", handle.ptr, tail.ctypes.data_as(ctypes.c_void_p),
               ctypes.c_void_p(y.addr + 8 * (SPLIT - 1)), 32)
        assert tail.tolist() == [3.0, 5.0, 5.0, 5.0]
    finally:
        x.free()
        y.free()
