"""The solver templates (cpp/lin_solvers) compiled with g++ against a host test
double of the VectorOperations concept and run on the reference's 1-D upwind
convection-diffusion problem (source/linear_operator.hpp:28-90): every variant
(GMRES MGS/CGS2/selective CGS, BiCGStab(L), plain BiCGStab with and without the
preconditioner, preconditioned or plain residual, with and without residual
recomputation) converges to the tolerance, agrees with a direct tridiagonal
solve, and the failure paths (iteration cap, NaN input) stop with the documented
flags."""
import os
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    d = tmp_path_factory.mktemp("solvers")
    exe = str(d / "t")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-Wall", "-Werror",
                           "-I", os.path.join(ROOT, "high_precision_blas1_amd", "cpp"),
                           "-I", os.path.join(ROOT, "tests", "cpp"),
                           os.path.join(ROOT, "tests", "cpp", "test_solvers_cpu.cpp"), "-o", exe])
    out = subprocess.check_output([exe], text=True, timeout=120)
    res = {}
    for line in out.splitlines():
        f = line.split()
        res[f[0]] = f[1:]
    return res


def test_all_variants_converge(results):
    solves = {k: v for k, v in results.items()
              if ("gmres" in k or "bicgstab" in k) and "_minres_" not in k and "_identity_" not in k}
    assert len(solves) == 21 + 24 + 6
    for name, (conv, iters, rel, err) in solves.items():
        assert conv == "1", name
        assert float(rel) <= 1.5e-10, (name, rel)
        assert float(err) < 1e-7, (name, err)
        assert int(iters) > 0


def test_preconditioning_and_restart_iteration_counts(results):
    # identical arithmetic up to the Gram-Schmidt variant: MGS and CGS2 need the
    # same number of iterations on this well-conditioned problem (within one cycle)
    for tag in ("plain", "high", "multi"):
        a = int(results["%s_gmres_mgs_prec0" % tag][1])
        b = int(results["%s_gmres_cgs2_prec0" % tag][1])
        c = int(results["%s_gmres_dgks_prec0" % tag][1])
        assert abs(a - b) <= 30 and abs(c - b) <= 30


def test_failure_paths(results):
    for tag in ("plain", "high", "multi"):
        conv, iters, valid = results["%s_cap" % tag]
        assert conv == "0" and int(iters) >= 10 and valid == "1"
        conv, iters, valid = results["%s_nan" % tag]
        assert conv == "0" and valid == "0"


def test_min_resid_norm_iterate_returned(results):
    # with out_min_resid_norm the solvers copy the best checked iterate back into x:
    # its true residual equals the smallest residual norm the monitor saw (the
    # monitor's norms are the true ones here: no preconditioner, true residuals
    # at the checks of GMRES; recursive ones for BiCGStab, equal up to rounding)
    for tag in ("plain", "high", "multi"):
        for solver in ("bicgstab", "bicgstabl", "gmres"):
            mn, tr, last = (float(v) for v in results["%s_minres_%s" % (tag, solver)])
            assert mn <= last * (1 + 1e-12), (tag, solver, mn, last)
            assert abs(tr - mn) <= 1e-6 * mn, (tag, solver, tr, mn)


def test_lucky_breakdown_on_identity(results):
    # BiCGStab(L) on the identity: the first BiCG step gives x = b exactly, then the
    # denominators vanish; the solver stops without touching x and, the true residual
    # being zero, reports convergence rather than a breakdown
    for tag in ("plain", "high", "multi"):
        for L in (1, 2, 4):
            conv, iters, err, brk = results["%s_identity_bicgstabl%d" % (tag, L)]
            assert conv == "1" and brk == "0", (tag, L, conv, brk)
            assert 1 <= int(iters) <= L and float(err) == 0.0, (tag, L, iters, err)
