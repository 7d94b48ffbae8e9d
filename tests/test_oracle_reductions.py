"""Pins the oracle's reductions.

* Sequential Dot2 / naive dot / Sum2 (dot_product.h:96-121, sums.hpp:163-176) are
  compared bit for bit with the reference's own code compiled in place
  (oracle/_ref, built from source/common/*.h when /root/reference exists).
* The exact dot / sum (replacing dot_product_gmp.hpp:108-130) is checked against
  math.fsum, which is correctly rounded (Shewchuk), over products split exactly
  by Dekker's algorithm, and against fractions.Fraction.
* Known-answer tests from the reference drivers: constant x = y = 7e-4
  (test_threads.cpp:50-64) and the +-120 sparse vector
  (test_vector_operations.cpp:126-128).
* Error bounds restated from error_bounds.hpp are compared with the reference's.
* Results satisfy the Ogita-Rump-Oishi bound
  |res - x.y| <= u|x.y| + gamma_{n-1}^2 |x|.|y| against the exact value.
"""
import math
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as O

U = 2.0 ** -53


def dekker(a, b):
    f = 134217729.0
    def split(v):
        c = f * v
        hi = c - (c - v)
        return hi, v - hi
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    e = al * bl - (((p - ah * bh) - al * bh) - ah * bl)
    return p, e


def fsum_dot(x, y):
    p, e = dekker(x, y)
    return math.fsum(np.concatenate([p, e]).tolist())


def gamma(n, u=U):
    return n * u / (1 - n * u)


def oro_ok(res, exact_dd, x, y):
    n = x.size
    ex = F(exact_dd[0]) + F(exact_dd[1])  # within u^2 |x.y| of the exact value
    ad = float(np.dot(np.abs(x), np.abs(y))) * (1 + 1e-10)
    bound = U * abs(float(ex)) + gamma(n - 1) ** 2 * ad
    return abs(F(float(res)) - ex) <= F(bound) * (1 + F(1, 10 ** 9))


def test_exact_dot_vs_fsum(rng):
    for n in (1, 2, 10, 1000, 100000):
        x = rng.standard_normal(n) * np.exp2(rng.integers(-40, 40, n))
        y = rng.standard_normal(n) * np.exp2(rng.integers(-40, 40, n))
        hi, lo = O.exact_dot(x, y)
        assert hi == fsum_dot(x, y)
        if n <= 1000:
            ex = sum((F(a) * F(b) for a, b in zip(x.tolist(), y.tolist())), F(0))
            assert hi == float(ex)
            assert lo == float(ex - F(hi))


def test_exact_sum_vs_fsum(rng):
    for n in (1, 7, 5000, 200000):
        x = rng.standard_normal(n) * np.exp2(rng.integers(-60, 60, n))
        assert O.exact_sum(x) == math.fsum(x.tolist())
    # catastrophic cancellation and subnormals
    x = np.array([1e308, 1.0, -1e308, 2.0 ** -1074, 2.0 ** -1074, -1.0])
    assert O.exact_sum(x) == 2.0 ** -1073


def test_kat_constant_7e4():
    # test_threads.cpp:50-64: x_j = y_j = 7e-4, exact value N * fl(7e-4)^2.
    n = 100003
    v = 7e-4
    x = np.full(n, v)
    ex = n * F(v) * F(v)
    hi, lo = O.exact_dot(x, x)
    assert hi == float(ex) and lo == float(ex - F(hi))
    d2 = O.dot2(x, x)
    assert abs(F(d2) - ex) <= F(U) * abs(ex) + F(gamma(n - 1) ** 2) * ex
    assert d2 == float(ex)


def test_kat_sparse_pm120():
    # test_vector_operations.cpp:126-128: zeros with -120 at 1234 and +120 at N/2.
    n = 1 << 16
    x = np.zeros(n)
    x[1234] = -120.0
    x[n // 2] = 120.0
    assert O.sum2(x) == 0.0
    assert O.exact_sum(x) == 0.0
    assert O.asum2(x) == 240.0


def _ref_or_skip():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return R


def test_dot2_bitexact_vs_reference_build(rng):
    R = _ref_or_skip()
    for n in (1, 3, 1000, 65537):
        x = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))
        y = rng.standard_normal(n)
        assert O.dot2(x, y) == R.ref_dot2(n, O.ptr(x), O.ptr(y))
        assert O.dot_naive(x, y) == R.ref_dot_naive(n, O.ptr(x), O.ptr(y))
        xf, yf = x.astype(np.float32), y.astype(np.float32)
        assert O.lib().orc_dot2_f32(n, O.ptr(xf), O.ptr(yf)) == R.ref_dot2_f32(n, O.ptr(xf), O.ptr(yf))
    x, y, limbs, dd = O.gendot(1 << 14, 1e24, 7)
    assert O.dot2(x, y) == R.ref_dot2(x.size, O.ptr(x), O.ptr(y))


def test_sum2_bitexact_vs_reference_threaded_one_part(rng):
    # threaded_reduction::sum with one part is sequential Sum2 then an exact merge.
    R = _ref_or_skip()
    for n in (1, 17, 100000):
        x = rng.standard_normal(n) * np.exp2(rng.integers(-50, 50, n))
        assert O.sum2(x) == R.ref_sum2_threaded(n, O.ptr(x), 1)
        assert O.asum2(x) == R.ref_asum2_threaded(n, O.ptr(x), 1)
        assert O.lib().orc_sum2_threaded(n, O.ptr(x), 1) == R.ref_sum2_threaded(n, O.ptr(x), 1)


def test_threaded_vs_reference(rng):
    R = _ref_or_skip()
    n = 1 << 18
    x, y, limbs, dd = O.gendot(n, 1e20, 3)
    ex = O.exact_dot(x, y)
    for parts in (1, 2, 5, 8):
        a = O.lib().orc_dot2_threaded(n, O.ptr(x), O.ptr(y), parts)
        b = R.ref_dot2_threaded(n, O.ptr(x), O.ptr(y), parts)
        assert oro_ok(a, ex, x, y) and oro_ok(b, ex, x, y)
        if parts == 1:
            assert a == b
        na = O.lib().orc_dot_naive_threaded(n, O.ptr(x), O.ptr(y), parts)
        nb = R.ref_dot_naive_threaded(n, O.ptr(x), O.ptr(y), parts)
        if parts <= 2:
            assert na == nb or parts == 2


def test_bounds_vs_reference():
    R = _ref_or_skip()
    for n, cond, k in ((1 << 20, 1e16, 24), (1000, 1e8, 7), (1 << 27, 1e32, 24), (5, 3.0, 1)):
        for fn_o, fn_r in ((O.dot_bounds, R.ref_dot_bounds), (O.sum_bounds, R.ref_sum_bounds)):
            mine = fn_o(n, cond, k, eps=2.0 ** -52)
            ref = np.zeros(6)
            fn_r(n, cond, k, O.ptr(ref))
            np.testing.assert_array_equal(mine, ref)


def test_bounds_closed_forms():
    # gamma_n = n eps / (1 - n eps); ORO dot bound = eps + 0.5 gamma_n^2 cond.
    b = O.dot_bounds(1000, 1e10, 10, eps=U)
    g = 1000 * U / (1 - 1000 * U)
    assert b[4] == pytest.approx(U + 0.5 * g * g * 1e10, rel=1e-15)
    assert b[0] == pytest.approx(0.5 * g * 1e10, rel=1e-15)


@pytest.mark.parametrize("cond", [1e4, 1e8, 1e16, 1e24, 1e32])
def test_dot2_within_oro_bound(cond):
    n = 1 << 16
    x, y, limbs, dd = O.gendot(n, cond, 0)
    ex = O.exact_dot(x, y)
    exf = O.limbs_to_fraction(limbs)
    assert ex[0] == float(exf) and ex[1] == float(exf - F(ex[0]))
    for res in (O.dot2(x, y), O.lib().orc_dot2_refgpu_order(n, O.ptr(x), O.ptr(y), 0),
                O.hpb_order_reduce("dot2", x, y, (256, 4, 16, 2))[2]):
        assert oro_ok(res, ex, x, y)
    # the naive dot loses all digits at high condition numbers
    if cond >= 1e20:
        assert O.rel_err(O.dot_naive(x, y), ex) > 1e-4


def test_sum_variants_exact_on_integers(rng):
    # Integer-valued data below 2^53 sums exactly in every order: any dropped or
    # duplicated element shows.
    x = rng.integers(-1000, 1000, 50001).astype(np.float64)
    ex = float(x.sum(dtype=np.int64)) if False else float(int(np.sum(x.astype(np.int64))))
    assert O.sum2(x) == ex
    assert O.lib().orc_sum_naive(x.size, O.ptr(x)) == ex
    assert O.lib().orc_sum_kahan(x.size, O.ptr(x)) == ex
    assert O.lib().orc_dot2_refgpu_order(x.size, O.ptr(x), O.ptr(x), 1) == ex
    for cfg in ((256, 4, 1, 2), (256, 4, 7, 2), (512, 2, 33, 1), (64, 1, 5, 1), (256, 8, 4096, 2)):
        assert O.hpb_order_reduce("sum2", x, cfg=cfg)[2] == ex
        assert O.hpb_order_reduce("sum", x, cfg=cfg)[2] == ex
        assert O.hpb_order_reduce("asum", x, cfg=cfg)[2] == float(np.abs(x).sum())
        y = np.ones_like(x)
        assert O.hpb_order_reduce("dot2", x, y, cfg=cfg)[2] == ex


def test_kahan_bound(rng):
    n = 100000
    x = rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))
    ex = F(math.fsum(x.tolist()))
    ax = float(np.abs(x).sum())
    k = O.lib().orc_sum_kahan(n, O.ptr(x))
    assert abs(F(k) - ex) <= F(2 * U * ax + n * U * U * ax) + F(U) * abs(ex)


OPS_ALL = ["dot2", "dot", "sum2", "sum", "asum2", "asum", "nrm2c", "nrm2"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("op", OPS_ALL)
def test_order_reduce_every_op_exact_on_integers(rng, dtype, op):
    """The library-order emulation for every op and both precisions: small integers
    make every product and partial sum exact (below 2^24 in FP32), so the result
    is the integer one in any order, and a dropped, repeated or misindexed element
    changes it.  Covers the FP32 instantiation and the vector widths 1, 2 and 4."""
    n = 50001
    x = rng.integers(-8, 9, n).astype(dtype)
    y = rng.integers(-8, 9, n).astype(dtype)
    y[n // 3] = 3                      # the planted change below must matter
    xi, yi = x.astype(np.int64), y.astype(np.int64)
    want = {"dot2": int(xi @ yi), "dot": int(xi @ yi), "sum2": int(xi.sum()), "sum": int(xi.sum()),
            "asum2": int(np.abs(xi).sum()), "asum": int(np.abs(xi).sum()), "nrm2c": int(xi @ xi),
            "nrm2": int(xi @ xi)}[op]
    assert abs(want) < 2 ** 24
    binary = op in ("dot2", "dot")
    for cfg in ((256, 4, 1, 1), (256, 4, 7, 2), (512, 2, 33, 4), (64, 1, 5, 1), (256, 8, 4096, 2),
                (512, 1, 3, 4), (128, 2, 200, 2)):
        out = O.hpb_order_reduce(op, x, y if binary else None, cfg=cfg)
        assert out.dtype == dtype
        assert out[2] == want, (op, cfg, out, want)
        if O.REDUCE_OPS[op] in (0, 2, 4, 6):
            assert out[1] == 0          # exact sums leave no compensation
    # a planted mistake shows: one element changed by one
    x2 = x.copy()
    x2[n // 3] += 1
    assert O.hpb_order_reduce(op, x2, y if binary else None, cfg=(256, 4, 7, 2))[2] != want


def test_order_nrm2_scaling_rule():
    """oracle.hpb_order_nrm2: unchanged inside the safe range, exact power-of-two
    scaling outside it, LAPACK-style results at the extremes."""
    import scipy.linalg.blas as B
    r = np.random.default_rng(11)
    base = r.uniform(-1, 1, 20011)
    cfg = (512, 4, 64, 2)
    # inside the range: identical to the unscaled library order
    for op in ("nrm2c", "nrm2"):
        a = O.hpb_order_nrm2(op, base, cfg)
        b = O.hpb_order_reduce(op, base, cfg=cfg)
        assert a[:3].tobytes() == b.tobytes() and a[3] == np.sqrt(b[2])
    # outside it: the norm of the unscaled data times the power-of-two scale, bit
    # for bit (same significands), and within 2 ulps of the scaled BLAS nrm2
    n0 = O.hpb_order_nrm2("nrm2c", base, cfg)[3]
    for e in (500, 700, 1000, -500, -700, -1000):
        v = np.ldexp(base, e)
        got = O.hpb_order_nrm2("nrm2c", v, cfg)
        assert got[3] == np.ldexp(n0, e), e
        assert abs(got[3] / B.dnrm2(v) - 1) <= 4.5e-16, e
    # decimal scales change significands: compare with BLAS
    for sc in (1e200, 1e-200, 1e300, 1e-300):
        got = O.hpb_order_nrm2("nrm2c", base * sc, cfg)[3]
        assert abs(got / B.dnrm2(base * sc) - 1) <= 4.5e-16, sc
    # FP32
    f = base.astype(np.float32)
    f0 = O.hpb_order_nrm2("nrm2c", f, (512, 4, 64, 4))[3]
    for e in (40, 60, -40, -60):
        v = np.ldexp(f, e).astype(np.float32)
        got = O.hpb_order_nrm2("nrm2c", v, (512, 4, 64, 4))[3]
        assert got == np.float32(np.ldexp(np.float64(f0), e)), e
        assert abs(float(got) / float(B.snrm2(v)) - 1) <= 2.5e-7, e
    # closed form: 4^j copies of c have norm |c| 2^j
    for e in (700, -700, 0):
        v = np.full(1 << 16, 0.75 * 2.0 ** e)
        assert O.hpb_order_nrm2("nrm2c", v, cfg)[3] == 0.75 * 2.0 ** e * 256
    # special values: inf without NaN, NaN with one; zero and empty vectors
    v = np.ones(1000)
    v[3] = -np.inf
    assert O.hpb_order_nrm2("nrm2c", v, cfg)[3] == np.inf
    v[9] = np.nan
    assert np.isnan(O.hpb_order_nrm2("nrm2c", v, cfg)[3])
    assert O.hpb_order_nrm2("nrm2c", np.zeros(10), cfg)[3] == 0.0
    assert O.hpb_order_nrm2("nrm2c", np.zeros(0), cfg)[3] == 0.0
