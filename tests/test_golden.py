"""Golden fixtures of the Dot2 path (tests/golden/, written by
tests/golden/make_goldens.py from the reference's own CPU code built in place and
the oracle's long accumulator).

CPU: the fixture inputs regenerate bit for bit (host generator, seed 0; the
reference's known-answer vectors); the exact values agree with fractions.Fraction
and, for the known-answer vectors, with their closed forms; the oracle's
restatement reproduces every stored output of the reference's code (sequential
Dot2 / naive dot, FP32 Dot2, Sum2 / asum2 and the error bounds) bit for bit.

GPU: the device generator reproduces the inputs, the device exact dot returns the
stored (hi, lo) bit for bit, and the device Dot2 (FP64 and FP32) lies within the
Ogita-Rump-Oishi bound of the stored exact value, as the reference's Dot2 does."""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import high_precision_blas1_amd as H
import oracle as O
from conftest import ROOT, say

GOLD = os.path.join(ROOT, "tests", "golden")
U = 2.0 ** -53


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "dot2_goldens.json")) as f:
        meta = json.load(f)
    arr = np.load(os.path.join(GOLD, "dot2_inputs.npz"), allow_pickle=False)
    cases = []
    for i, name in enumerate(meta["order"]):
        rec = meta["cases"][name]
        cases.append((name, arr["x"][i], arr["y"][i], rec))
    return meta["n"], cases


def _h(s):
    return float.fromhex(s)


def _oro_bound(exact, x, y, u=U):
    n = x.size
    g = (n - 1) * u / (1 - (n - 1) * u)
    return u * abs(exact) + g * g * float(np.dot(np.abs(x), np.abs(y))) * (1 + 1e-12)


def test_inputs_regenerate(gold):
    n, cases = gold
    for name, x, y, _ in cases:
        if name.startswith("gendot_kappa_"):
            k = float(name.rsplit("_", 1)[1])
            hx, hy, _, _ = O.gendot(n, k, 0)
            assert hx.tobytes() == x.tobytes() and hy.tobytes() == y.tobytes(), name
        elif name == "const_7e-4":
            assert np.all(x == 7.0e-4) and np.all(y == 7.0e-4)
        else:
            u = np.zeros(n)
            u[1234], u[n // 2] = -120.0, 120.0
            assert x.tobytes() == u.tobytes() and y.tobytes() == u.tobytes()


def test_exact_values(gold):
    n, cases = gold
    for name, x, y, rec in cases:
        ex = sum((Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y)), Fraction(0))
        hi, lo = _h(rec["exact_hi"]), _h(rec["exact_lo"])
        assert hi == float(ex) and lo == float(ex - Fraction(hi)), name
        assert (hi, lo) == O.exact_dot(x, y), name
        if name == "const_7e-4":
            assert ex == n * Fraction(7.0e-4) ** 2
        if name == "sparse_pm120":
            assert ex == 28800 and O.sum2(x) == 0.0 and O.asum2(x) == 240.0


def test_oracle_reproduces_reference_outputs(gold):
    n, cases = gold
    for name, x, y, rec in cases:
        assert O.dot2(x, y) == _h(rec["ref_dot2"]), name
        assert O.dot_naive(x, y) == _h(rec["ref_dot_naive"]), name
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        assert float(O.lib().orc_dot2_f32(n, O.ptr(x32), O.ptr(y32))) == _h(rec["ref_dot2_f32"]), name
        # the threaded reference with one part is the sequential algorithm
        assert _h(rec["ref_dot2_threaded_1"]) == _h(rec["ref_dot2"]), name
        assert _h(rec["ref_dot_naive_threaded_1"]) == _h(rec["ref_dot_naive"]), name
        assert O.sum2(x) == _h(rec["ref_sum2_threaded_1"]), name
        assert O.asum2(x) == _h(rec["ref_asum2_threaded_1"]), name
        if rec["cond"] != "inf":
            b = O.dot_bounds(n, _h(rec["cond"]), 1, eps=2.0 ** -52)   # the reference uses eps = 2^-52
            assert [float(v).hex() for v in b] == rec["ref_dot_bounds"], name


def test_reference_dot2_within_oro_bound(gold):
    """The stored reference Dot2 meets the bound the device results are held to."""
    n, cases = gold
    for name, x, y, rec in cases:
        ex = Fraction(_h(rec["exact_hi"])) + Fraction(_h(rec["exact_lo"]))
        assert abs(Fraction(_h(rec["ref_dot2"])) - ex) <= Fraction(_oro_bound(float(ex), x, y)), name


@pytest.mark.gpu
def test_device_against_goldens(handle, gold):
    n, cases = gold
    for name, x, y, rec in cases:
        say("golden", name)
        ex = Fraction(_h(rec["exact_hi"])) + Fraction(_h(rec["exact_lo"]))
        if name.startswith("gendot_kappa_"):
            k = float(name.rsplit("_", 1)[1])
            gx, gy = H.DeviceArray(n, np.float64, handle), H.DeviceArray(n, np.float64, handle)
            limbs, d3 = np.zeros(4, dtype=np.uint64), np.zeros(3)
            H.call("hpb_gendot_f64", handle.ptr, n, ctypes.c_double(float(np.log2(k))), 0, n // 2, gx.ptr, gy.ptr,
                   limbs.ctypes.data_as(ctypes.c_void_p), d3.ctypes.data_as(ctypes.c_void_p))
            assert gx.numpy().tobytes() == x.tobytes() and gy.numpy().tobytes() == y.tobytes(), name
            gx.free()
            gy.free()
        dx, dy = H.DeviceArray.from_numpy(x, handle), H.DeviceArray.from_numpy(y, handle)
        out = np.zeros(2)
        H.call("hpb_exact_dot_f64", handle.ptr, n, dx.ptr, dy.ptr, out.ctypes.data)
        assert (float(out[0]), float(out[1])) == (_h(rec["exact_hi"]), _h(rec["exact_lo"])), name
        g = H.dot2(handle, dx, dy)
        assert abs(Fraction(float(g[2])) - ex) <= Fraction(_oro_bound(float(ex), x, y)), name
        dx.free()
        dy.free()
        # FP32 Dot2 on the rounded inputs, against their own exact value and bound
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        ex32 = sum((Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x32, y32)), Fraction(0))
        fx, fy = H.DeviceArray.from_numpy(x32, handle), H.DeviceArray.from_numpy(y32, handle)
        g32 = H.dot2(handle, fx, fy)
        b32 = _oro_bound(float(ex32), x32.astype(np.float64), y32.astype(np.float64), 2.0 ** -24)
        assert abs(Fraction(float(g32[2])) - ex32) <= Fraction(b32), name
        assert abs(Fraction(_h(rec["ref_dot2_f32"])) - ex32) <= Fraction(b32), name
        fx.free()
        fy.free()
