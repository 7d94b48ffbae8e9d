"""Complex dot x^H y on the device: bit-exact against the oracle's statement of
the library's order (orc_zdot_order, pinned in tests/test_oracle_zdot.py) for
complex128 / complex64, compensated and plain, many sizes; the compensated result
on an ill-conditioned pair is exact where the plain one is not."""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 5, 1023, 1024, 4097, 100003, 1 << 20]


def _grid(handle, n):
    tile = 256 * 4
    return max(1, min((n + tile - 1) // tile, handle.num_cu()))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("comp", [True, False])
def test_zdot_bitexact(handle, dtype, comp):
    real = np.float64 if dtype == np.complex128 else np.float32
    suf = "f64" if real == np.float64 else "f32"
    fn = ("hpb_zdot2_" if comp else "hpb_zdot_") + suf
    r = np.random.default_rng(3)
    for n in SIZES:
        say("zdot", np.dtype(dtype).name, comp, n)
        x = ((r.standard_normal(n) + 1j * r.standard_normal(n)) * np.exp2(r.integers(-20, 20, n))).astype(dtype)
        y = (r.standard_normal(n) + 1j * r.standard_normal(n)).astype(dtype)
        dx = H.DeviceArray.from_numpy(x.view(real), handle)
        dy = H.DeviceArray.from_numpy(y.view(real), handle)
        out = np.zeros(6, dtype=real)
        H.call(fn, handle.ptr, n, dx.ptr, dy.ptr, out.ctypes.data)
        exp = O.zdot_order(x, y, comp, (256, 4, _grid(handle, n)))
        assert out.tobytes() == exp.tobytes(), (n, out, exp)
        dx.free()
        dy.free()


def test_zdot_cancellation_and_async(handle):
    n = 1 << 16
    r = np.random.default_rng(4)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    y = r.standard_normal(n) + 1j * r.standard_normal(n)
    x = np.concatenate([x * 1e15, [1.0 + 1j], -x * 1e15])
    y = np.concatenate([y, [2.0 + 0j], y])
    m = x.size
    dx = H.DeviceArray.from_numpy(x.view(np.float64), handle)
    dy = H.DeviceArray.from_numpy(y.view(np.float64), handle)
    out = np.zeros(6)
    H.call("hpb_zdot2_f64", handle.ptr, m, dx.ptr, dy.ptr, out.ctypes.data)
    # conj(1 + i) * 2 = 2 - 2i; the large terms cancel exactly in exact arithmetic
    assert out[4] == 2.0 and out[5] == -2.0
    dout = H.DeviceArray(6, np.float64, handle)
    H.call("hpb_zdot2_f64_async", handle.ptr, m, dx.ptr, dy.ptr, dout.ptr)
    assert np.array_equal(dout.numpy(), out)
