"""Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The solver templates (GMRES, BiCGStab(L), BiCGStab, the monitor) on the host
  test double of the VectorOperations concept: the program of
  tests/test_solvers_cpu.py, built with -fsanitize=address,undefined and
  -fno-sanitize-recover, must run clean and print the same results as the plain
  build.
* The library's MatrixMarket reader and writer (csrc/hpb_mm.cpp, with the
  library's status plumbing csrc/hpb_capi.cpp and what it links to) compiled with
  the sanitizers on the host side only (hipcc -Xarch_host), driven through good
  and malformed files.
No GPU code is instrumented and nothing here needs a GPU.
"""
import os
import subprocess

import pytest

from conftest import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
PKG = os.path.join(ROOT, "high_precision_blas1_amd")


def _parse(out):
    return {f[0]: f[1:] for f in (line.split() for line in out.splitlines()) if f}


def _build_solvers(exe, extra):
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-Wall", "-Werror"] + extra +
                          ["-I", os.path.join(PKG, "cpp"), "-I", os.path.join(ROOT, "tests", "cpp"),
                           os.path.join(ROOT, "tests", "cpp", "test_solvers_cpu.cpp"), "-o", exe])


def test_solver_templates_sanitized(tmp_path):
    plain, san = str(tmp_path / "plain"), str(tmp_path / "san")
    _build_solvers(plain, [])
    _build_solvers(san, SAN)
    a = subprocess.run([plain], capture_output=True, text=True, timeout=300, env=ENV)
    b = subprocess.run([san], capture_output=True, text=True, timeout=600, env=ENV)
    assert a.returncode == 0, a.stderr[-3000:]
    assert b.returncode == 0, b.stderr[-3000:]
    assert "runtime error" not in b.stderr and "AddressSanitizer" not in b.stderr, b.stderr[-3000:]
    # the instrumented build computes the same thing
    assert _parse(a.stdout) == _parse(b.stdout)


def test_matrix_market_sanitized(tmp_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    objs = []
    # hpb_capi.cpp's hpb_create refers to the no-allocation set-up in hpb_multi.hip
    for src in ("hpb_mm.cpp", "hpb_capi.cpp", "hpb_multi.hip"):
        obj = str(tmp_path / (src + ".o"))
        # host side instrumented only; the device pass of these files is empty
        subprocess.check_call([hipcc, "-x", "hip", "-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
                               "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                               "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer",
                               "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
                               "-c", os.path.join(PKG, "csrc", src), "-o", obj])
        objs.append(obj)
    exe = str(tmp_path / "mm")
    # the driver is plain host C++: g++, instrumented, linked with the objects above
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc))), "lib")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror"] + SAN +
                          ["-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "test_mm_sanitized.cpp")] + objs +
                          ["-L", rocm_lib, "-lamdhip64", "-Wl,-rpath," + rocm_lib, "-o", exe])
    d = tmp_path / "files"
    d.mkdir()
    p = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=300, env=ENV)
    assert p.returncode == 0, p.stderr[-3000:]
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-3000:]
    res = _parse(p.stdout)
    ok = {"general": (3, 5), "symmetric": (4, 6), "skew": (3, 4), "integer": (2, 3), "pattern": (3, 3),
          "duplicates": (2, 4), "empty": (5, 0), "big": (150000, 600000), "big_split": (150000, 600000)}
    for name, (rows, nnz) in ok.items():
        st, r, k, _ = res[name]
        assert (int(st), int(r), int(k)) == (0, rows, nnz), name
    assert res["big"][3] == res["big_split"][3]   # the same matrix through either parser
    for name in ("bad_banner", "bad_index", "zero_index", "short", "array", "complex", "garbage_value", "no_size",
                 "huge_size"):
        assert int(res[name][0]) != 0, name
    assert int(res["missing"][0]) != 0
