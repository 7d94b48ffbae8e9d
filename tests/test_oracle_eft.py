"""Pins the oracle's error-free transformations (dot_product.h:50-66) by the exact
identities a + b = s + t and a*b = p + e, checked in rational arithmetic."""
import ctypes
from fractions import Fraction as F

import numpy as np

import oracle as O


def _cases(rng, n=3000):
    a = rng.standard_normal(n) * np.exp2(rng.integers(-300, 300, n))
    b = rng.standard_normal(n) * np.exp2(rng.integers(-300, 300, n))
    adv = [(1.0, 1e-16), (1.0, -1e-16), (1e100, 1.0), (-1e100, 1e100), (0.0, -0.0), (2.0 ** -1074, 2.0 ** -1074),
           (1.7976931348623157e308, -1.7976931348623157e308), (0.1, 0.2), (1.0, 2.0 ** -53), (1.0, 3 * 2.0 ** -54),
           (np.nextafter(1.0, 2.0), -1.0)]
    return list(zip(a, b)) + adv


def test_two_sum_exact(rng):
    L = O.lib()
    t = ctypes.c_double()
    for a, b in _cases(rng):
        s = L.orc_two_sum(float(a), float(b), ctypes.byref(t))
        assert F(s) + F(t.value) == F(float(a)) + F(float(b))
        assert s == float(a) + float(b)


def test_two_prod_exact(rng):
    L = O.lib()
    e = ctypes.c_double()
    a = rng.standard_normal(3000) * np.exp2(rng.integers(-200, 200, 3000))
    b = rng.standard_normal(3000) * np.exp2(rng.integers(-200, 200, 3000))
    pairs = list(zip(a, b)) + [(0.1, 0.1), (1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53), (3.0, 1.0 / 3.0), (-7e-4, 7e-4)]
    for x, y in pairs:
        p = L.orc_two_prod(float(x), float(y), ctypes.byref(e))
        assert F(p) + F(e.value) == F(float(x)) * F(float(y))
        assert p == float(x) * float(y)
