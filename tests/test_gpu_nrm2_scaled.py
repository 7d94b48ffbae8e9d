"""Overflow- and underflow-safe norms on the GPU (hpb_nrm2c_* / hpb_nrm2_*).

The reference's standard-precision norm is cuBLAS nrm2, which scales
(source/common/gpu_vector_operations.h:315-327).  The library's norms sum
unscaled squares while max |x_i| lies in [2^-400, 2^400] (FP64) or [2^-30, 2^30]
(FP32) and otherwise repeat the sum over x_i 2^k.  Checked here:
* bit-exact against the oracle's statement of the same rule in the library's
  order (oracle.hpb_order_nrm2), blocking and asynchronous entry points;
* closed form: a constant vector c 1 of length 4^j has norm |c| 2^j exactly;
* |x| near 1e200 / 1e-200 / 1e300 / 1e-300 / subnormal in FP64 and 1e30 / 1e-30
  in FP32 give the norm of the unscaled data times the scale (the scaled sum
  sees the same significands), where the plain sum of squares would give inf
  or lose every digit;
* inf and NaN entries: inf without NaN, NaN with one.
"""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu


def _gpu_both(handle, op, x):
    """Blocking result and asynchronous (device) result of norm op on x."""
    dx = H.DeviceArray.from_numpy(x, handle)
    blocking = H.reduce(op, handle, dx)
    suf = "f64" if x.dtype == np.float64 else "f32"
    dout = H.DeviceArray(4, x.dtype, handle)
    H.call("hpb_%s_%s_async" % (op, suf), handle.ptr, x.size, dx.ptr, dout.ptr)
    asyn = dout.numpy()
    cfg = handle.reduce_config(x.dtype, x.size, dx.addr % 16 == 0, False)
    dx.free()
    dout.free()
    return blocking, asyn, cfg


def _check(handle, op, x):
    g, a, cfg = _gpu_both(handle, op, x)
    emu = O.hpb_order_nrm2(op, x, cfg)
    assert g.tobytes() == emu.tobytes(), (op, x.dtype, x.size, g, emu)
    assert a.tobytes() == emu.tobytes(), (op, x.dtype, x.size, a, emu)
    return g


@pytest.mark.parametrize("op", ["nrm2c", "nrm2"])
@pytest.mark.parametrize("scale", [1.0, 2.0 ** 600, 2.0 ** -600, 1e200, 1e-200, 1e300, 1e-300, 1e-310])
def test_f64_scaled(handle, op, scale):
    n = 100003
    say("nrm2 f64", op, scale)
    r = np.random.default_rng(7)
    base = r.uniform(-1, 1, n)
    x = base * scale
    g = _check(handle, op, x)
    ref = np.linalg.norm(base) * scale
    assert np.isfinite(g[3]) and g[3] > 0
    tol = 1e-15 if scale > 1e-300 else 1e-9      # subnormal entries keep fewer digits
    assert abs(g[3] / ref - 1) <= tol, (g[3], ref)


@pytest.mark.parametrize("op", ["nrm2c", "nrm2"])
@pytest.mark.parametrize("scale", [1.0, 2.0 ** 40, 2.0 ** -40, 1e30, 1e-30])
def test_f32_scaled(handle, op, scale):
    n = 65539
    say("nrm2 f32", op, scale)
    r = np.random.default_rng(8)
    base = r.uniform(-1, 1, n).astype(np.float32)
    x = (base.astype(np.float64) * scale).astype(np.float32)
    g = _check(handle, op, x)
    ref = np.linalg.norm(x.astype(np.float64))
    assert np.isfinite(g[3]) and abs(float(g[3]) / ref - 1) <= 1e-6, (g[3], ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_constant_vector_closed_form(handle, dtype):
    n = 1 << 16
    for e in ((700, -700, 3, -3) if dtype == np.float64 else (50, -50, 3, -3)):
        c = dtype(0.75 * 2.0 ** e)
        say("constant", np.dtype(dtype).name, e)
        x = np.full(n, c, dtype=dtype)
        g = _check(handle, "nrm2c", x)
        assert float(g[3]) == 0.75 * 2.0 ** e * 256.0


def test_mixed_magnitudes_and_special_values(handle):
    n = 4099
    x = np.zeros(n)
    x[:10] = 1e200
    x[10:] = 1e-200                       # far below the largest entries: no effect
    g = _check(handle, "nrm2c", x)
    assert g[3] == pytest.approx(1e200 * np.sqrt(10.0), rel=1e-15)
    y = np.ones(n)
    y[17] = np.inf
    g = _check(handle, "nrm2c", y)
    assert np.isinf(g[3]) and g[3] > 0
    y[99] = np.nan
    g = _check(handle, "nrm2c", y)
    assert np.isnan(g[3])
    z = np.ones(n)
    z[5] = np.nan                          # in-range data with a NaN: NaN
    assert np.isnan(_check(handle, "nrm2c", z)[3])
    # zero vector and an empty vector need no rescaling
    assert _check(handle, "nrm2c", np.zeros(n))[3] == 0.0
    assert _check(handle, "nrm2c", np.zeros(0))[3] == 0.0


def test_rescale_then_ordinary_call(handle):
    """The rescaling state is cleared: a badly scaled norm followed by an ordinary
    one, and a norm inside a reduction sequence, give the unscaled results."""
    n = 70001
    x = np.random.default_rng(3).uniform(-1, 1, n)
    for v in (x * 1e250, x, x * 1e-250, x):
        _check(handle, "nrm2c", v)
    dx = H.DeviceArray.from_numpy(x * 1e250, handle)
    dy = H.DeviceArray.from_numpy(x, handle)
    dout = H.DeviceArray(12, np.float64, handle)
    H.call("hpb_nrm2c_f64_async", handle.ptr, n, dx.ptr, dout.ptr)
    H.call("hpb_dot2_f64_async", handle.ptr, n, dy.ptr, dy.ptr, ctypes.c_void_p(dout.addr + 32))
    H.call("hpb_nrm2c_f64_async", handle.ptr, n, dy.ptr, ctypes.c_void_p(dout.addr + 64))
    r = dout.numpy()
    assert r[3] == pytest.approx(np.linalg.norm(x) * 1e250, rel=1e-15)
    assert r[8:12].tobytes() == H.nrm2c(handle, dy).tobytes()
    assert r[6] == H.dot2(handle, dy, dy)[2]
    for a in (dx, dy, dout):
        a.free()
