"""Writes the golden fixtures of the Dot2 path: dot2_inputs.npz and dot2_goldens.json.

Inputs (N = 4096 each):
* host GenDot pairs (oracle.gendot, seed 0) with target condition numbers 1e4, 1e8,
  1e16, 1e24 and 1e32;
* the reference's known-answer vectors: x = y = 7e-4 (test_threads.cpp:50-64) and
  the sparse vector with -120 at index 1234 and +120 at N/2
  (test_vector_operations.cpp:126-128), paired with itself.

Outputs per input, all doubles written as float.hex strings:
* the exact dot product: the oracle's long accumulator (checked here against
  fractions.Fraction) as the nearest double and the rounded remainder;
* the reference's own CPU results, from its dot_product.h, threaded_reduction.h and
  error_bounds.hpp compiled in place (oracle/_ref, built by oracle/build_oracle.py):
  sequential Dot2 and naive dot (FP64), sequential Dot2 in FP32 on the inputs rounded
  to FP32, threaded Dot2 / naive dot / Sum2 / asum2 of x with one part (deterministic
  there; several parts merge in a mutex-dependent order), and the Dot2 error bounds
  for (N, the achieved condition number, one block).

Run from the repository root in the container that holds the reference sources:
    python tests/golden/make_goldens.py
"""
import json
import os
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

N = 4096
KAPPAS = [1e4, 1e8, 1e16, 1e24, 1e32]


def inputs():
    cases = []
    for k in KAPPAS:
        x, y, _, dd = O.gendot(N, k, 0)
        cases.append(("gendot_kappa_%g" % k, x, y))
    c = np.full(N, 7.0e-4)
    cases.append(("const_7e-4", c, c.copy()))
    u = np.zeros(N)
    u[1234] = -120.0
    u[N // 2] = 120.0
    cases.append(("sparse_pm120", u, u.copy()))
    return cases


def exact_fraction(x, y):
    return sum((Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y)), Fraction(0))


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run oracle/build_oracle.py in the container with the reference")
    names, xs, ys, out = [], [], [], {"n": N, "cases": {}}
    for name, x, y in inputs():
        hi, lo = O.exact_dot(x, y)
        ex = exact_fraction(x, y)
        # the long accumulator rounds once: hi = RN(exact), lo = RN(exact - hi)
        assert hi == float(ex) and lo == float(ex - Fraction(hi)), name
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        abs_dot = float(np.dot(np.abs(x), np.abs(y)))
        cond = 2.0 * abs_dot / abs(hi) if hi != 0.0 else float("inf")
        bounds = np.zeros(6)
        if np.isfinite(cond):
            R.ref_dot_bounds(N, cond, 1, O.ptr(bounds))
        rec = {
            "exact_hi": hi.hex(), "exact_lo": lo.hex(),
            "ref_dot2": R.ref_dot2(N, O.ptr(x), O.ptr(y)).hex(),
            "ref_dot_naive": R.ref_dot_naive(N, O.ptr(x), O.ptr(y)).hex(),
            "ref_dot2_f32": float(R.ref_dot2_f32(N, O.ptr(x32), O.ptr(y32))).hex(),
            "ref_dot2_threaded_1": R.ref_dot2_threaded(N, O.ptr(x), O.ptr(y), 1).hex(),
            "ref_dot_naive_threaded_1": R.ref_dot_naive_threaded(N, O.ptr(x), O.ptr(y), 1).hex(),
            "ref_sum2_threaded_1": R.ref_sum2_threaded(N, O.ptr(x), 1).hex(),
            "ref_asum2_threaded_1": R.ref_asum2_threaded(N, O.ptr(x), 1).hex(),
            "cond": cond.hex() if np.isfinite(cond) else "inf",
            "ref_dot_bounds": [float(v).hex() for v in bounds],
        }
        out["cases"][name] = rec
        names.append(name)
        xs.append(x)
        ys.append(y)
    out["order"] = names
    np.savez_compressed(os.path.join(HERE, "dot2_inputs.npz"), x=np.stack(xs), y=np.stack(ys))
    with open(os.path.join(HERE, "dot2_goldens.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote %d cases" % len(names))


if __name__ == "__main__":
    main()
