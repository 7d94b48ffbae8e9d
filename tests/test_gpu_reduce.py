"""GPU parity of the compensated and plain reductions (C ABI, gfx950).

Three levels:
* bit-exact against the oracle's restatement of the same arithmetic in the
  library's blocked order (oracle.hpb_order_reduce) for every op, dtype, size
  class, alignment and several launch shapes -- catches any indexing mistake;
* Ogita-Rump-Oishi bound against the exact dot on GenDot pairs, kappa 1e4..1e32,
  and agreement with the reference's sequential Dot2 (dot_product.h:108-121)
  within twice that bound;
* size-independent properties at N = 2^27 (the benchmark size): the constant
  known-answer vector of test_threads.cpp:50-64 and exact power-of-two scaling.
"""
import ctypes
from fractions import Fraction as F

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu
U = 2.0 ** -53

OPS = ["dot2", "dot", "sum2", "sum", "asum2", "asum", "nrm2c", "nrm2"]
SIZES = [0, 1, 2, 3, 5, 63, 64, 65, 257, 1000, 4099, 65536 + 7, (1 << 20) + 3]


def _data(n, dtype, seed):
    r = np.random.default_rng(seed)
    x = (r.standard_normal(n) * np.exp2(r.integers(-8, 8, n))).astype(dtype)
    y = (r.standard_normal(n) * np.exp2(r.integers(-8, 8, n))).astype(dtype)
    return x, y


def _check_op(handle, op, x, y, offset):
    n = x.size
    binary = op in ("dot2", "dot")
    dx = H.DeviceArray.from_numpy(x, handle, offset_bytes=offset)
    dy = H.DeviceArray.from_numpy(y, handle, offset_bytes=offset) if binary else None
    gpu = H.reduce(op, handle, dx, dy)
    aligned = (dx.addr % 16 == 0) and (dy is None or dy.addr % 16 == 0)
    cfg = handle.reduce_config(x.dtype, n, aligned, binary)
    emu = O.hpb_order_reduce(op, x, y if binary else None, cfg)
    assert gpu[0].tobytes() == emu[0].tobytes(), (op, n, cfg, gpu, emu)
    assert gpu[1].tobytes() == emu[1].tobytes(), (op, n, cfg, gpu, emu)
    assert gpu[2].tobytes() == emu[2].tobytes(), (op, n, cfg, gpu, emu)
    if op in ("nrm2c", "nrm2"):
        assert gpu[3] == np.sqrt(gpu[2])
    dx.free()
    if dy is not None:
        dy.free()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("op", OPS)
def test_bitexact_vs_blocked_order(handle, dtype, op):
    for n in SIZES:
        for offset in (0, np.dtype(dtype).itemsize):
            say("reduce", op, np.dtype(dtype).name, "n=%d" % n, "offset=%d" % offset)
            x, y = _data(n, dtype, n + 17)
            _check_op(handle, op, x, y, offset)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bitexact_all_launch_shapes(handle, dtype):
    n = (1 << 18) + 5
    x, y = _data(n, dtype, 99)
    saved = [handle.get_tuning(k, dtype) for k in (0, 2)]
    try:
        for block in (256, 512):
            for unroll in (1, 2, 4, 8):
                for gmax in (1, 3, 64, 1024, 4096):
                    for flags in (0, 1):
                        handle.set_tuning(0, dtype, block, unroll, gmax, flags)
                        handle.set_tuning(2, dtype, block, unroll, gmax, flags)
                        say("shape", block, unroll, gmax, flags, np.dtype(dtype).name)
                        _check_op(handle, "dot2", x, y, 0)
                        _check_op(handle, "sum2", x, y, 0)
                        _check_op(handle, "nrm2c", x, y, 0)
    finally:
        handle.set_tuning(0, dtype, *saved[0])
        handle.set_tuning(2, dtype, *saved[1])


def _oro_bound(ex, x, y):
    n = x.size
    g = (n - 1) * U / (1 - (n - 1) * U)
    return U * abs(ex) + g * g * float(np.dot(np.abs(x), np.abs(y))) * (1 + 1e-12)


@pytest.mark.parametrize("cond", [1e4, 1e8, 1e16, 1e24, 1e32])
def test_dot2_oro_bound_gendot(handle, cond):
    n = 1 << 20
    say("gendot bound", cond)
    x, y, limbs, dd = O.gendot(n, cond, 0)
    exf = O.limbs_to_fraction(limbs)
    dx = H.DeviceArray.from_numpy(x, handle)
    dy = H.DeviceArray.from_numpy(y, handle)
    g = H.dot2(handle, dx, dy)
    ref = O.dot2(x, y)                          # reference sequential Dot2
    b = _oro_bound(float(exf), x, y)
    assert abs(F(float(g[2])) - exf) <= F(b)
    assert abs(F(ref) - exf) <= F(b)
    assert abs(float(g[2]) - ref) <= 2 * b
    # the (s, c) pair holds the compensated value before its final rounding
    assert F(float(g[0])) + F(float(g[1])) - exf <= F(b) + F(U) * abs(exf)
    naive = H.dot(handle, dx, dy)[2]
    if cond >= 1e24:
        assert abs(naive - float(exf)) > abs(float(g[2]) - float(exf))


def test_dot2_large_kat_and_scaling(handle):
    n = 1 << 27
    v = 7e-4
    say("large KAT n=2^27")
    x = np.full(n, v)
    dx = H.DeviceArray.from_numpy(x, handle)
    g = H.dot2(handle, dx, dx)
    ex = n * F(v) * F(v)
    assert abs(F(float(g[2])) - ex) <= F(U) * ex + F(((n - 1) * U) ** 2 * 1.01) * ex
    # power-of-two scaling is exact in every summation order
    H.call("hpb_scal_f64", handle.ptr, n, 4.0, dx.ptr)
    g4 = H.dot2(handle, dx, dx)
    assert g4[2] == 16.0 * g[2] and g4[0] == 16.0 * g[0]
    s = H.sum2(handle, dx)
    assert abs(F(float(s[2])) - 4 * n * F(v)) <= F(U) * 4 * n * F(v) * 2
    dx.free()


def test_async_result_on_device(handle):
    n = 100003
    x, y = _data(n, np.float64, 5)
    dx = H.DeviceArray.from_numpy(x, handle)
    dy = H.DeviceArray.from_numpy(y, handle)
    dout = H.DeviceArray(8, np.float64, handle)
    H.call("hpb_dot2_f64_async", handle.ptr, n, dx.ptr, dy.ptr, dout.ptr)
    H.call("hpb_nrm2c_f64_async", handle.ptr, n, dx.ptr, ctypes.c_void_p(dout.addr + 32))
    res = dout.numpy()
    assert res[:4].tobytes() == H.dot2(handle, dx, dy).tobytes()
    assert res[4:].tobytes() == H.nrm2c(handle, dx).tobytes()
