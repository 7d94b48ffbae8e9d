"""GPU parity of the elementwise kernels against the oracle's statement of the
same rounding (oracle.map_op), which tests/test_oracle_map.py pins against
exactly rounded rational arithmetic."""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu

SIZES = [1, 3, 7, 1000, 65536 + 3]

# name -> (oracle op, n scalars, inputs used, output aliases input index or None)
CASES = {
    "fill": ("fill", 1, 0, None),
    "add_mul_scalar": ("add_mul_scalar", 2, 1, 0),
    "scal": ("scal", 1, 1, 0),
    "copy": ("copy", 0, 1, None),
    "abs": ("abs", 0, 1, None),
    "assign_mul": ("scal", 1, 1, None),
    "axpy": ("axpy", 1, 2, 1),
    "axpby": ("axpby", 2, 2, None),
    "axpbypcz": ("axpbypcz", 3, 3, 2),
    "axpbypcwpdz": ("axpbypcwpdz", 4, 4, 3),
    "assign_mul4": ("axpbypcwpdz", 4, 4, None),
    "mul_pointwise": ("mulpw", 2, 2, None),
    "mul_pointwise_inplace": ("mulpw_ip", 2, 2, 0),
    "div_pointwise": ("divpw", 2, 2, None),
    "div_pointwise_inplace": ("divpw_ip", 2, 2, 0),
    "axpbyc": ("axpbyc", 2, 2, None),
}


def _call(name, handle, dtype, n, sc, ins, out):
    suf = "f64" if dtype == np.float64 else "f32"
    c = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    s = [c(v) for v in sc]
    p = [a.ptr for a in ins]
    h = handle.ptr
    fn = "hpb_%s_%s" % (name, suf)
    if name == "fill":
        H.call(fn, h, n, s[0], out.ptr)
    elif name == "add_mul_scalar":
        H.call(fn, h, n, s[0], s[1], out.ptr)
    elif name == "scal":
        H.call(fn, h, n, s[0], out.ptr)
    elif name in ("copy", "abs"):
        H.call(fn, h, n, p[0], out.ptr)
    elif name == "assign_mul":
        H.call(fn, h, n, s[0], p[0], out.ptr)
    elif name == "axpy":
        H.call(fn, h, n, s[0], p[0], out.ptr)
    elif name in ("axpby", "mul_pointwise", "div_pointwise", "axpbyc"):
        H.call(fn, h, n, s[0], p[0], s[1], p[1], out.ptr)
    elif name == "axpbypcz":
        H.call(fn, h, n, s[0], p[0], s[1], p[1], s[2], out.ptr)
    elif name == "axpbypcwpdz":
        H.call(fn, h, n, s[0], p[0], s[1], p[1], s[2], p[2], s[3], out.ptr)
    elif name == "assign_mul4":
        H.call(fn, h, n, s[0], p[0], s[1], p[1], s[2], p[2], s[3], p[3], out.ptr)
    elif name in ("mul_pointwise_inplace", "div_pointwise_inplace"):
        H.call(fn, h, n, out.ptr, s[1], p[1])
    else:
        raise KeyError(name)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(CASES))
def test_map_bitexact(handle, dtype, name):
    op, nsc, nin, alias = CASES[name]
    r = np.random.default_rng(7)
    for n in SIZES:
        for offset in (0, np.dtype(dtype).itemsize):
            say("map", name, np.dtype(dtype).name, n, offset)
            sc = list(r.standard_normal(nsc).astype(dtype))
            ins_h = [(r.standard_normal(n) * 3).astype(dtype) + (2 if k == 1 else 0) for k in range(max(nin, 1))]
            ins_d = [H.DeviceArray.from_numpy(a, handle, offset_bytes=offset) for a in ins_h]
            if alias is None:
                out = H.DeviceArray(n, dtype, handle, offset_bytes=offset)
            else:
                out = ins_d[alias]
            _call(name, handle, dtype, n, sc, ins_d, out)
            got = out.numpy()
            exp = O.map_op(op, n, sc, ins_h[:nin] if nin else [], dtype)
            assert got.tobytes() == exp.tobytes(), (name, n, offset)


def test_swap_valid_set_get(handle):
    n = 12345
    a = np.arange(n, dtype=np.float64)
    b = -a
    da, db = H.DeviceArray.from_numpy(a, handle), H.DeviceArray.from_numpy(b, handle)
    H.call("hpb_swap_f64", handle.ptr, n, da.ptr, db.ptr)
    assert np.array_equal(da.numpy(), b) and np.array_equal(db.numpy(), a)
    ok = ctypes.c_int()
    H.call("hpb_valid_f64", handle.ptr, n, da.ptr, ctypes.byref(ok))
    assert ok.value == 1
    for bad in (np.nan, np.inf, -np.inf):
        H.call("hpb_set_value_f64", handle.ptr, n, 777, ctypes.c_double(bad), da.ptr)
        H.call("hpb_valid_f64", handle.ptr, n, da.ptr, ctypes.byref(ok))
        assert ok.value == 0
        v = ctypes.c_double()
        H.call("hpb_get_value_f64", handle.ptr, n, 777, da.ptr, ctypes.byref(v))
        assert v.value == bad or (np.isnan(bad) and np.isnan(v.value))
    H.call("hpb_set_value_f64", handle.ptr, n, 777, ctypes.c_double(1.5), da.ptr)
    H.call("hpb_valid_f64", handle.ptr, n, da.ptr, ctypes.byref(ok))
    assert ok.value == 1
    with pytest.raises(H.HPBError):
        H.call("hpb_set_value_f64", handle.ptr, n, n, ctypes.c_double(0.0), da.ptr)


def test_random_uniform_matches_host_philox(handle):
    n = 4099
    d = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_random_uniform_f64", handle.ptr, n, 1234, 3, ctypes.c_double(-1.0), ctypes.c_double(1.0), d.ptr)
    v = d.numpy()
    assert v.min() > -1.0 and v.max() <= 1.0
    for i in (0, 1, 77, 4098):
        w = O.philox([i, 0, 3, 0], [1234, 0])
        k = ((w[0] << 32) | w[1]) >> 11
        u = (k + 1) * 2.0 ** -53
        # fma(b - a, u, a) with b - a = 2 and u a multiple of 2^-53: exact
        assert v[i] == 2.0 * u - 1.0


def test_convert(handle):
    n = 1001
    x = np.random.default_rng(1).standard_normal(n)
    dx = H.DeviceArray.from_numpy(x, handle)
    df = H.DeviceArray(n, np.float32, handle)
    dd = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_convert_f64_to_f32", handle.ptr, n, dx.ptr, df.ptr)
    H.call("hpb_convert_f32_to_f64", handle.ptr, n, df.ptr, dd.ptr)
    assert np.array_equal(df.numpy(), x.astype(np.float32))
    assert np.array_equal(dd.numpy(), x.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_map_all_launch_shapes(handle, dtype):
    """Every elementwise launch shape (block, unroll, grid cap, nontemporal
    flag) gives the same bits; sizes cover ragged tails of every tile width."""
    r = np.random.default_rng(11)
    saved = handle.get_tuning(1, dtype)
    try:
        for block, unroll in ((256, 1), (256, 2), (256, 4), (512, 2), (512, 4)):
            for grid in (1, 7, 512):
                for flags in (0, 1):
                    handle.set_tuning(1, dtype, block, unroll, grid, flags)
                    for n in (5, block * unroll * 2 + 3, 300001):
                        say("map shape", np.dtype(dtype).name, block, unroll, grid, flags, n)
                        sc = list(r.standard_normal(4).astype(dtype))
                        ins_h = [r.standard_normal(n).astype(dtype) for _ in range(4)]
                        ins_d = [H.DeviceArray.from_numpy(a, handle) for a in ins_h]
                        out = H.DeviceArray(n, dtype, handle)
                        _call("assign_mul4", handle, dtype, n, sc, ins_d, out)
                        exp = O.map_op("axpbypcwpdz", n, sc, ins_h, dtype)
                        assert out.numpy().tobytes() == exp.tobytes()
                        _call("axpy", handle, dtype, n, sc, ins_d, ins_d[1])
                        exp = O.map_op("axpy", n, sc, ins_h[:2], dtype)
                        assert ins_d[1].numpy().tobytes() == exp.tobytes()
                        for a in ins_d + [out]:
                            a.free()
    finally:
        handle.set_tuning(1, dtype, *saved)
