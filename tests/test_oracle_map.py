"""Pins oracle.map_op (the rounding the elementwise kernels document) against
exactly rounded rational arithmetic: fma(a, x, t) must equal the nearest double
to a*x + t, and the compensated axpby must be faithful (within one ulp of the
exact a*x + b*y) and equal to the nearest double unless the products cancel."""
from fractions import Fraction as F

import numpy as np

import oracle as O


def rn(q):
    return float(q)  # Fraction -> nearest double, ties to even


def test_fma_forms_exact(rng):
    n = 400
    x, y, w, z = (rng.standard_normal(n) for _ in range(4))
    a, b, c, d = rng.standard_normal(4)
    got = O.map_op("axpy", n, [a], [x, y])
    assert all(got[i] == rn(F(a) * F(x[i]) + F(y[i])) for i in range(n))
    got = O.map_op("axpby", n, [a, b], [x, y])
    assert all(got[i] == rn(F(a) * F(x[i]) + F(b * y[i])) for i in range(n))
    got = O.map_op("axpbypcz", n, [a, b, c], [x, y, z])
    for i in range(n):
        t = rn(F(a) * F(x[i]) + F(b * y[i]))
        assert got[i] == rn(F(c) * F(z[i]) + F(t))
    got = O.map_op("axpbypcwpdz", n, [a, b, c, d], [x, y, w, z])
    for i in range(n):
        t = rn(F(a) * F(x[i]) + F(b * y[i]))
        t = rn(F(c) * F(w[i]) + F(t))
        assert got[i] == rn(F(d) * F(z[i]) + F(t))
    got = O.map_op("add_mul_scalar", n, [a, b], [x])
    assert all(got[i] == rn(F(b) * F(x[i]) + F(a)) for i in range(n))


def test_axpbyc_faithful(rng):
    n = 2000
    x = rng.standard_normal(n)
    y = x * (1 + rng.standard_normal(n) * 1e-12)  # near cancellation
    a, b = 1.7, -1.7
    got = O.map_op("axpbyc", n, [a, b], [x, y])
    hits = 0
    for i in range(n):
        ex = F(a) * F(x[i]) + F(b) * F(y[i])
        r = rn(ex)
        if got[i] == r:
            hits += 1
        else:
            assert abs(F(got[i]) - ex) <= abs(F(np.spacing(r)))
    assert hits > 0.95 * n
    # plain a*x + b*y loses everything here
    plain = O.map_op("axpby", n, [a, b], [x, y])
    assert np.mean(plain != got) > 0.5


def test_simple_ops(rng):
    n = 100
    x = rng.standard_normal(n)
    y = rng.standard_normal(n) + 3
    assert np.array_equal(O.map_op("abs", n, [], [x]), np.abs(x))
    assert np.array_equal(O.map_op("copy", n, [], [x]), x)
    assert np.array_equal(O.map_op("fill", n, [2.5], []), np.full(n, 2.5))
    assert np.array_equal(O.map_op("scal", n, [3.0], [x]), 3.0 * x)
    assert np.array_equal(O.map_op("mulpw", n, [2.0, 3.0], [x, y]), (2.0 * x) * (3.0 * y))
    assert np.array_equal(O.map_op("divpw", n, [2.0, 3.0], [x, y]), (2.0 * x) / (3.0 * y))
    assert np.array_equal(O.map_op("mulpw_ip", n, [0.0, 3.0], [x, y]), x * (3.0 * y))
    assert np.array_equal(O.map_op("divpw_ip", n, [0.0, 3.0], [x, y]), x / (3.0 * y))
