"""Fused multi-dot / multi-update kernels (Gram-Schmidt building blocks).

* compensated multidot: every column within the Ogita-Rump-Oishi bound of the
  exact dot and within twice that bound of the single-vector Dot2;
* plain multidot: within gamma_n |V_j|.|w| of the exact dot;
* compensated multiaxpy: each element faithful to the exact y + sum a_j V_j
  (rational arithmetic), plain multiaxpy equals the sequential fma chain;
* unaligned pointers, k = 1..64 and ragged lengths.
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


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.addr for a in arrs])


@pytest.mark.parametrize("k", [1, 3, 8, 9, 17, 51, 64])
@pytest.mark.parametrize("n", [1, 1000, 100003])
def test_multidot(handle, k, n):
    say("multidot", k, n)
    r = np.random.default_rng(k * 1000 + n)
    w = r.standard_normal(n)
    Vh = [r.standard_normal(n) * np.exp2(r.integers(-10, 10, n)) for _ in range(k)]
    off = 8 if k % 2 else 0
    dw = H.DeviceArray.from_numpy(w, handle, offset_bytes=off)
    dV = [H.DeviceArray.from_numpy(v, handle, offset_bytes=off) for v in Vh]
    for comp in (1, 0):
        out = np.zeros(4 * k)
        H.call("hpb_multidot_f64", handle.ptr, k, _ptrs(dV), n, dw.ptr, comp, out.ctypes.data_as(ctypes.c_void_p))
        for j in range(k):
            ex = O.exact_dot(Vh[j], w)
            ad = float(np.dot(np.abs(Vh[j]), np.abs(w)))
            g = (n - 1) * U / (1 - (n - 1) * U) if n > 1 else 0.0
            res = out[4 * j + 2]
            if comp:
                b = U * abs(ex[0]) + g * g * ad * 1.001 + 1e-300
                assert abs(F(res) - (F(ex[0]) + F(ex[1]))) <= F(b)
                single = H.dot2(handle, dV[j], dw)[2]
                assert abs(res - single) <= 2 * b
            else:
                assert abs(res - ex[0]) <= n * U * ad * 1.01 + 1e-300


@pytest.mark.parametrize("kc,grid,nt", [(4, 256, 0), (4, 4096, 1), (8, 512, 1), (8, 4096, 0)])
def test_multidot_tuned_shapes(handle, kc, grid, nt):
    """Every launch shape the autotuner may pick (kind 3): columns per chunk, grid,
    nontemporal loads -- the compensated results stay within the ORO bound."""
    n, k = 300007, 13
    say("multidot shape", kc, grid, nt)
    r = np.random.default_rng(kc + grid + nt)
    w = r.standard_normal(n)
    Vh = [r.standard_normal(n) for _ in range(k)]
    dw = H.DeviceArray.from_numpy(w, handle)
    dV = [H.DeviceArray.from_numpy(v, handle) for v in Vh]
    saved = handle.get_tuning(3, np.float64)
    try:
        handle.set_tuning(3, np.float64, 256, kc, grid, nt)
        out = np.zeros(4 * k)
        H.call("hpb_multidot_f64", handle.ptr, k, _ptrs(dV), n, dw.ptr, 1, out.ctypes.data_as(ctypes.c_void_p))
    finally:
        handle.set_tuning(3, np.float64, *saved)
    g = (n - 1) * U / (1 - (n - 1) * U)
    for j in range(k):
        ex = O.exact_dot(Vh[j], w)
        b = U * abs(ex[0]) + g * g * float(np.dot(np.abs(Vh[j]), np.abs(w))) * 1.001
        assert abs(F(out[4 * j + 2]) - (F(ex[0]) + F(ex[1]))) <= F(b), j
    with pytest.raises(H.HPBError):
        handle.set_tuning(3, np.float64, 512, 8, grid, 0)
    with pytest.raises(H.HPBError):
        handle.set_tuning(3, np.float64, 256, 2, grid, 0)


def test_multiaxpy(handle):
    n, k = 301, 11
    r = np.random.default_rng(4)
    y0 = r.standard_normal(n)
    Vh = [r.standard_normal(n) for _ in range(k)]
    a = r.standard_normal(k)
    dV = [H.DeviceArray.from_numpy(v, handle, offset_bytes=8) for v in Vh]
    for comp in (1, 0):
        dy = H.DeviceArray.from_numpy(y0, handle, offset_bytes=8)
        H.call("hpb_multiaxpy_f64", handle.ptr, k, a.ctypes.data_as(ctypes.c_void_p), _ptrs(dV), n, dy.ptr, comp)
        got = dy.numpy()
        for i in range(n):
            if comp:
                ex = F(y0[i]) + sum(F(a[j]) * F(Vh[j][i]) for j in range(k))
                rr = float(ex)
                assert got[i] == rr or abs(F(got[i]) - ex) <= F(np.spacing(abs(rr)))
            else:
                s = y0[i]
                for j in range(k):
                    s = float(F(a[j]) * F(Vh[j][i]) + F(s))  # fma
                assert got[i] == s
