"""Build recipe for the CPU oracle (test infrastructure).

* ``oracle/liboracle.so``   -- the C restatement in ``hpb_oracle.c`` (gcc).
* ``oracle/_ref/libhpbref.so`` -- the reference's own header-only CPU code
  (``source/common/dot_product.h``, ``source/common/threaded_reduction.h``,
  ``source/high_prec/error_bounds.hpp``) compiled from where it lies under
  /root/reference through the wrapper ``ref_wrapper.cpp``.  It needs ROCm's
  ``thrust/complex.h``, hence hipcc; nothing is copied out of the reference.
  Built only when /root/reference exists.

Usage: ``python oracle/build_oracle.py`` (also called by ``__graft_entry__.build``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/source"


def build_oracle(verbose=True):
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-Wall", "-Wno-unused-function", os.path.join(HERE, "hpb_oracle.c"), "-o", out,
           "-lm", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build_ref(verbose=True):
    """Compile the reference's CPU Dot2/Sum2/bounds headers; returns None if absent."""
    if not os.path.isdir(REF_SRC):
        if verbose:
            print("reference sources absent: oracle/_ref not rebuilt", flush=True)
        return None
    os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
    out = os.path.join(HERE, "_ref", "libhpbref.so")
    # -O2 (the reference Makefile uses -g only, Makefile:9,13): the CPU baseline is
    # timed optimised so that it is not handicapped.
    cmd = ["hipcc", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
           "-I", REF_SRC, os.path.join(HERE, "ref_wrapper.cpp"), "-o", out, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


if __name__ == "__main__":
    build_oracle()
    build_ref()
    sys.exit(0)
