"""The oracle's complex dot x^H y in the library's order (orc_zdot_order), pinned by
closed forms, integer data (exact for every launch shape, so a dropped or repeated
element shows), numpy's vdot, and exact rational arithmetic within the
compensated bound."""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def _exact(x, y):
    re = sum(Fraction(float(a.real)) * Fraction(float(b.real)) + Fraction(float(a.imag)) * Fraction(float(b.imag))
             for a, b in zip(x, y))
    im = sum(Fraction(float(a.real)) * Fraction(float(b.imag)) - Fraction(float(a.imag)) * Fraction(float(b.real))
             for a, b in zip(x, y))
    return re, im


def test_closed_form():
    x = np.array([1 + 2j])
    y = np.array([3 + 4j])
    for comp in (True, False):
        o = O.zdot_order(x, y, comp, (64, 1, 1))
        assert o[4] == 11.0 and o[5] == -2.0


@pytest.mark.parametrize("cfg", [(64, 1, 1), (256, 4, 3), (128, 2, 7), (512, 4, 40)])
def test_integer_data_exact(cfg):
    r = np.random.default_rng(1)
    n = 50001
    x = r.integers(-1000, 1000, n) + 1j * r.integers(-1000, 1000, n)
    y = r.integers(-1000, 1000, n) + 1j * r.integers(-1000, 1000, n)
    ex = np.sum(np.conj(x.astype(np.complex128)) * y)   # exact: integers < 2^53
    for comp in (True, False):
        o = O.zdot_order(x.astype(np.complex128), y.astype(np.complex128), comp, cfg)
        assert o[4] == ex.real and o[5] == ex.imag


def test_vs_vdot_and_exact_bound():
    r = np.random.default_rng(2)
    n = 3000
    x = (r.standard_normal(n) + 1j * r.standard_normal(n)) * np.exp2(r.integers(-30, 30, n))
    y = r.standard_normal(n) + 1j * r.standard_normal(n)
    o = O.zdot_order(x, y, False, (256, 4, 5))
    v = np.vdot(x, y)
    assert abs(o[4] - v.real) <= 1e-10 * np.sum(np.abs(x) * np.abs(y))
    assert abs(o[5] - v.imag) <= 1e-10 * np.sum(np.abs(x) * np.abs(y))
    oc = O.zdot_order(x, y, True, (256, 4, 5))
    re, im = _exact(x, y)
    absum = float(sum(abs(Fraction(float(a.real)) * Fraction(float(b.real))) + abs(Fraction(float(a.imag)) *
                      Fraction(float(b.imag))) for a, b in zip(x, y)))
    u = 2.0 ** -53
    g = 2 * n * u / (1 - 2 * n * u)
    for got, ex in ((oc[4], re), (oc[5], im)):
        assert abs(Fraction(float(got)) - ex) <= Fraction(u) * abs(ex) + Fraction(g * g * absum * 4)


def test_ill_conditioned_compensated_beats_plain():
    # cancellation: x^H y with a huge term and its negative
    x = np.array([1e16 + 0j, 1.0 + 1j, -1e16 + 0j])
    y = np.array([1.0 + 1j, 1.0 + 0j, 1.0 + 1j])
    o = O.zdot_order(x, y, True, (64, 1, 1))
    # conj(x) y = (1e16 + 1e16 i) + (1 - i) + (-1e16 - 1e16 i) = 1 - i
    assert o[4] == 1.0 and o[5] == -1.0
