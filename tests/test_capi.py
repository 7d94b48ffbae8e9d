"""The C ABI library loads and exports exactly what include/hpb.h declares.

CPU-only: no HIP compute call is made (the container may have no GPU)."""
import ctypes
import os
import subprocess

import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "high_precision_blas1_amd", "libhpb.so")


def _exports():
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    return {l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("hpb_")}


def test_library_built():
    assert os.path.exists(LIB), "libhpb.so missing: run __graft_entry__.build()"


def test_every_declared_symbol_exported():
    from high_precision_blas1_amd._lib import parse_header
    decl = set(parse_header())
    assert len(decl) > 50
    exp = _exports()
    assert decl - exp == set(), "declared but not exported: %s" % sorted(decl - exp)
    assert exp - decl == set(), "exported but not declared: %s" % sorted(exp - decl)


def test_bindings_load():
    import high_precision_blas1_amd as H
    L = H.lib()
    assert L.hpb_version().decode().startswith("hpb")
    assert isinstance(L.hpb_last_error(), bytes)
    protos = H.prototypes()
    for name in protos:
        assert getattr(L, name).argtypes is not None


def test_invalid_arguments_fail_loudly_without_device_work():
    import high_precision_blas1_amd as H
    L = H.lib()
    # null handle / null output: rejected before any HIP call
    assert L.hpb_dot2_f64(None, 10, None, None, None) == -1
    assert "null" in L.hpb_last_error().decode()
    assert L.hpb_create(None, None) == -1
    with pytest.raises(H.HPBError):
        H.call("hpb_set_tuning", None, 0, 0, 256, 4, 16, 0)


def test_header_is_c_compatible(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "hpb.h"\nint main(void){return 0;}\n')
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           str(src), "-o", str(tmp_path / "t")])
