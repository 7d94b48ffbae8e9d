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
    from high_precision_blas1_amd._lib import missing_symbols
    assert missing_symbols() == []
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


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors of hpb_solver_params / hpb_solver_result have the C layout."""
    from high_precision_blas1_amd.solvers import SolverParams, SolverResult
    lines = []
    for cname, py in (("hpb_solver_params", SolverParams), ("hpb_solver_result", SolverResult)):
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src = tmp_path / "l.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "hpb.h"\nint main(void){\n%s\nreturn 0;}\n'
                   % "\n".join(lines))
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(tmp_path / "l")])
    got = dict(l.split() for l in subprocess.check_output([str(tmp_path / "l")], text=True).splitlines())
    for cname, py in (("hpb_solver_params", SolverParams), ("hpb_solver_result", SolverResult)):
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
