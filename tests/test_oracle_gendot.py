"""Pins the host GenDot generator (ORO Alg. 6.1 with an exactly tracked dot; the
reference's variant is generate_vector_pair.hpp:119-182) and the Philox stream.

* Philox4x32-10 against the Random123 known-answer vectors.
* The tracked exact dot equals the dot recomputed in rational arithmetic.
* Invariants: every entry is a multiple of 2^-53; head exponents lie in
  [0, b/2 + 1]; the achieved condition number is at least the target; the same
  seed reproduces the same bits and another seed does not.
"""
from fractions import Fraction as F

import numpy as np
import pytest

import oracle as O


def test_philox_kat():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("cond", [1e4, 1e16, 1e32])
@pytest.mark.parametrize("n,n_tail", [(3000, 1500), (2050, 1), (1000, 999), (5000, 300)])
def test_exact_dot_tracked(cond, n, n_tail):
    x, y, limbs, dd = O.gendot(n, cond, 11, n_tail)
    ex = sum((F(a) * F(b) for a, b in zip(x.tolist(), y.tolist())), F(0))
    assert O.limbs_to_fraction(limbs) == ex
    assert dd[0] == float(ex)
    ab = sum((abs(F(a) * F(b)) for a, b in zip(x.tolist(), y.tolist())), F(0))
    assert dd[1] == float(ab)


def test_invariants():
    n = 1 << 15
    for cond in (1e4, 1e8, 1e16, 1e24, 1e32):
        x, y, limbs, dd = O.gendot(n, cond, 0)
        hb = 0.5 * np.log2(cond)
        for v in (x, y):
            assert np.all(np.isfinite(v))
            assert np.all(v * 2.0 ** 53 == np.rint(v * 2.0 ** 53))
        head = x[: n // 2]
        ex = np.floor(np.log2(np.abs(head)))
        assert ex.max() <= np.rint(hb) + 1 and np.all(np.abs(head) < 2.0 ** (np.rint(hb) + 2))
        assert np.all(x != 0)
        # the construction cancels the head: achieved kappa >= target
        assert dd[2] >= cond
        assert abs(dd[0]) < 4.0


def test_seed_reproducible():
    a = O.gendot(4096, 1e16, 5)
    b = O.gendot(4096, 1e16, 5)
    c = O.gendot(4096, 1e16, 6)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert not np.array_equal(a[0], c[0])


def test_rejects_bad_arguments():
    with pytest.raises(ValueError):
        O.gendot(1, 1e16, 0)
    with pytest.raises(ValueError):
        O.gendot(1 << 20, 1e40, 0)
