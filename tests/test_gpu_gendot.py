"""Device GenDot generator: bit-identical to the host restatement (oracle.gendot)
for the same arguments, and usable as ground truth at the benchmark size through
the exactly tracked dot (the ORO bound holds against it)."""
import ctypes
from fractions import Fraction as F

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def device_gendot(handle, n, cond, seed, n_tail):
    x = H.DeviceArray(n, np.float64, handle)
    y = H.DeviceArray(n, np.float64, handle)
    limbs = np.zeros(4, dtype=np.uint64)
    d3 = np.zeros(3)
    H.call("hpb_gendot_f64", handle.ptr, n, ctypes.c_double(float(np.log2(cond))), seed, n_tail, x.ptr, y.ptr,
           limbs.ctypes.data_as(ctypes.c_void_p), d3.ctypes.data_as(ctypes.c_void_p))
    return x, y, limbs, d3


@pytest.mark.parametrize("n,n_tail", [(2, 1), (1000, 500), (4096, 1), (70001, 35000), ((1 << 20), 1 << 19),
                                      (300000, 299000), ((1 << 21) + 3, (1 << 20) + (1 << 19) + 1)])
@pytest.mark.parametrize("cond", [1e4, 1e16, 1e32])
def test_bitwise_equal_to_host(handle, n, n_tail, cond):
    say("gendot", n, n_tail, cond)
    dx, dy, limbs, d3 = device_gendot(handle, n, cond, 42, n_tail)
    hx, hy, hl, hd = O.gendot(n, cond, 42, n_tail)
    assert dx.numpy().tobytes() == hx.tobytes()
    assert dy.numpy().tobytes() == hy.tobytes()
    assert limbs.tobytes() == hl.tobytes()
    assert d3.tobytes() == hd.tobytes()
    dx.free()
    dy.free()


def test_benchmark_size_ground_truth(handle):
    n = 1 << 27
    for cond in (1e8, 1e16, 1e24):
        say("gendot 2^27", cond)
        dx, dy, limbs, d3 = device_gendot(handle, n, cond, 0, n // 2)
        exf = O.limbs_to_fraction(limbs)
        assert d3[0] == float(exf)
        g = H.dot2(handle, dx, dy)
        g_abs = H.reduce("dot2", handle, dx, dy)  # same call twice: reproducible
        assert g.tobytes() == g_abs.tobytes()
        gam = (n - 1) * U / (1 - (n - 1) * U)
        bound = U * abs(d3[0]) + gam * gam * d3[1] * 1.0001
        assert abs(F(float(g[2])) - exf) <= F(bound)
        assert d3[2] >= cond
        dx.free()
        dy.free()


def test_rejects_out_of_range(handle):
    x = H.DeviceArray(16, np.float64, handle)
    with pytest.raises(H.HPBError):
        H.call("hpb_gendot_f64", handle.ptr, 16, ctypes.c_double(200.0), 0, 8, x.ptr, x.ptr, None, None)
    with pytest.raises(H.HPBError):
        H.call("hpb_gendot_f64", handle.ptr, 16, ctypes.c_double(10.0), 0, 16, x.ptr, x.ptr, None, None)
