"""Build the gfx950 library ``high_precision_blas1_amd/libhpb.so`` in-tree with hipcc.

Every source under ``csrc/`` is compiled separately (in parallel) and linked into
one shared object exporting the C ABI declared in ``include/hpb.h``.

Flags:
  --offload-arch=gfx950          MI355X only, no other targets
  -ffp-contract=off              every fused multiply-add is written explicitly;
                                 TwoSum/TwoProd and the documented rounding of the
                                 elementwise kernels must not be re-associated
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 instructions (v_pk_*_f32)
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "libhpb.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", os.path.join(PKG, "cpp"),
          "-Wno-unused-command-line-argument"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src, verbose):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
TEST_BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_dropin")


def build_tests(verbose=True):
    """The drop-in C++ test program (tests/cpp/test_dropin.cpp), host code linked
    against libhpb.so; run by tests/test_gpu_dropin.py on the GPU."""
    os.makedirs(os.path.dirname(TEST_BIN), exist_ok=True)
    tmp = TEST_BIN + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-Werror", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "cpp"),
           TEST_SRC, "-o", tmp, "-L", PKG, "-lhpb", "-Wl,-rpath,$ORIGIN/../../../high_precision_blas1_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("test program build failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, TEST_BIN)
    return TEST_BIN


if __name__ == "__main__":
    build()
    build_tests()
    sys.exit(0)
