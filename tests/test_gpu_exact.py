"""GPU exact dot product and exact sum (long accumulator) against the oracle's
Kulisch accumulator (itself pinned against math.fsum and fractions), the
generator's tracked exact dot, and rational arithmetic."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu


def _exact_dot(handle, x, y):
    dx, dy = H.DeviceArray.from_numpy(x, handle), H.DeviceArray.from_numpy(y, handle)
    out = np.zeros(2)
    suf = "f64" if x.dtype == np.float64 else "f32"
    H.call("hpb_exact_dot_%s" % suf, handle.ptr, x.size, dx.ptr, dy.ptr, out.ctypes.data)
    dx.free()
    dy.free()
    return float(out[0]), float(out[1])


def _exact_sum(handle, x):
    dx = H.DeviceArray.from_numpy(x, handle)
    out = np.zeros(2)
    H.call("hpb_exact_sum_f64", handle.ptr, x.size, dx.ptr, out.ctypes.data)
    dx.free()
    return float(out[0]), float(out[1])


def _dd_of_fraction(v):
    hi = float(v)
    return hi, float(v - Fraction(hi))


@pytest.mark.parametrize("n", [0, 1, 2, 7, 1000, 65537, 1000003])
def test_exact_dot_random(handle, n):
    r = np.random.default_rng(n + 5)
    x = r.standard_normal(n) * np.exp2(r.integers(-40, 40, n))
    y = r.standard_normal(n)
    say("exact dot random", n)
    got = _exact_dot(handle, x, y)
    assert got == O.exact_dot(x, y) if n else got == (0.0, 0.0)


@pytest.mark.parametrize("log2cond", [13.3, 26.6, 53.2, 79.7, 106.3])
def test_exact_dot_gendot(handle, log2cond):
    """Ill-conditioned pairs: the generator's exact dot (256-bit tracked) and the
    Kulisch oracle agree with the device."""
    n = 200003
    x, y, limbs, dd = O.gendot(n, 2.0 ** log2cond, 0)
    say("exact dot gendot", log2cond)
    got = _exact_dot(handle, x, y)
    assert got == O.exact_dot(x, y)
    assert got == _dd_of_fraction(O.limbs_to_fraction(limbs))


def test_exact_dot_small_rational(handle):
    r = np.random.default_rng(3)
    for trial in range(20):
        n = int(r.integers(1, 60))
        x = r.standard_normal(n) * np.exp2(r.integers(-300, 300, n))
        y = r.standard_normal(n) * np.exp2(r.integers(-300, 300, n))
        ex = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y))
        assert _exact_dot(handle, x, y) == _dd_of_fraction(ex), trial


def test_exact_sum(handle):
    cases = [
        np.array([1.0, 1e100, 1.0, -1e100]),
        np.array([2.0 ** -1074, 2.0 ** -1074, -2.0 ** -1022, 1.0, -1.0]),
        np.array([1e280, 1e280 * (1 - 2 ** -52), -1e280]),
        np.array([0.1] * 10 + [-1.0]),
    ]
    r = np.random.default_rng(9)
    cases.append(r.standard_normal(300001) * np.exp2(r.integers(-500, 500, 300001)))
    v = r.standard_normal(100000)
    cases.append(np.concatenate([v, -v[::-1], [2.0 ** -1070]]))
    for c in cases:
        say("exact sum", c.size)
        got = _exact_sum(handle, c)
        assert got[0] == math.fsum(c)
        assert got[0] == O.exact_sum(c)
        if c.size < 1000:
            assert got == _dd_of_fraction(sum(Fraction(float(a)) for a in c))


def test_exact_dot_f32(handle):
    r = np.random.default_rng(4)
    for n in (1, 33, 4097, 300007):
        x = (r.standard_normal(n) * np.exp2(r.integers(-60, 60, n))).astype(np.float32)
        y = r.standard_normal(n).astype(np.float32)
        say("exact dot f32", n)
        got = _exact_dot(handle, x, y)
        assert got[0] == O.lib().orc_exact_dot_f32(n, O.ptr(x), O.ptr(y))
        if n < 5000:
            ex = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y))
            assert got == _dd_of_fraction(ex)


def test_exact_domain_errors(handle):
    x = np.ones(1000)
    y = np.ones(1000)
    x[17] = np.inf
    with pytest.raises(H.HPBError) as e:
        _exact_dot(handle, x, y)
    assert e.value.status == -1
    x[17] = 1e300
    y[17] = 1e300
    with pytest.raises(H.HPBError) as e:
        _exact_dot(handle, x, y)
    assert e.value.status == -2
    x[17] = 1e-300
    y[17] = 1e-300
    with pytest.raises(H.HPBError) as e:
        _exact_dot(handle, x, y)
    assert e.value.status == -2
    with pytest.raises(H.HPBError):
        _exact_sum(handle, np.array([1.0, np.nan]))
    # the handle stays usable after a domain error
    x[17] = y[17] = 1.0
    assert _exact_dot(handle, x, y) == (1000.0, 0.0)


def test_exact_dot_async_and_determinism(handle):
    n = 1 << 20
    x, y, limbs, dd = O.gendot(n, 1e20, 5)
    dx, dy = H.DeviceArray.from_numpy(x, handle), H.DeviceArray.from_numpy(y, handle)
    dout = H.DeviceArray(3, np.float64, handle)
    ref = _dd_of_fraction(O.limbs_to_fraction(limbs))
    for _ in range(3):
        H.call("hpb_exact_dot_f64_async", handle.ptr, n, dx.ptr, dy.ptr, dout.ptr)
        v = dout.numpy()
        assert (float(v[0]), float(v[1])) == ref and v[2] == 0.0


def test_exact_dot_2p27_generator(handle):
    """N = 2^27 device-generated pair: the exact dot equals the generator's ground truth."""
    n = 1 << 27
    dx, dy = H.DeviceArray(n, np.float64, handle), H.DeviceArray(n, np.float64, handle)
    limbs = (ctypes.c_uint64 * 4)()
    d3 = (ctypes.c_double * 3)()
    say("exact dot 2^27")
    H.call("hpb_gendot_f64", handle.ptr, n, ctypes.c_double(math.log2(1e16)), 0, n // 2, dx.ptr, dy.ptr, limbs, d3)
    out = np.zeros(2)
    H.call("hpb_exact_dot_f64", handle.ptr, n, dx.ptr, dy.ptr, out.ctypes.data)
    ref = _dd_of_fraction(O.limbs_to_fraction(list(limbs)))
    assert (float(out[0]), float(out[1])) == ref
    assert out[0] == d3[0]
    dx.free()
    dy.free()


def test_exact_dot_unaligned_and_wide_range(handle):
    """Vectors starting 8 bytes past a 16-byte boundary (the scalar-load kernel), and
    products spread over 2^-880..2^880 so that the per-thread expansions overflow into
    the block accumulator many times; both against the Kulisch oracle."""
    r = np.random.default_rng(21)
    n = 300001
    x = r.standard_normal(n) * np.exp2(r.integers(-440, 440, n))
    y = r.standard_normal(n) * np.exp2(r.integers(-440, 440, n))
    ref = O.exact_dot(x, y)
    say("exact dot wide range", n)
    assert _exact_dot(handle, x, y) == ref
    dx = H.DeviceArray.from_numpy(np.concatenate([[0.0], x]), handle)
    dy = H.DeviceArray.from_numpy(np.concatenate([[0.0], y]), handle)
    out = np.zeros(2)
    say("exact dot unaligned", n)
    H.call("hpb_exact_dot_f64", handle.ptr, n, ctypes.c_void_p(dx.addr + 8), ctypes.c_void_p(dy.addr + 8),
           out.ctypes.data)
    assert (float(out[0]), float(out[1])) == ref
    out = np.zeros(2)
    H.call("hpb_exact_sum_f64", handle.ptr, n, ctypes.c_void_p(dx.addr + 8), out.ctypes.data)
    assert float(out[0]) == math.fsum(x) == O.exact_sum(x)
    dx.free()
    dy.free()
