"""The precision form at its full size: FP32-storage inner GMRES(30) and
BiCGStab(2) with compensated FP32 dots inside FP64 iterative refinement, on the
128^3 first-order upwind convection-diffusion matrix (2.1 M rows) and the 256^3
anisotropic diffusion matrix (k_z / k_x = 1e4, 16.8 M rows), against the plain
FP64 compensated solve of the same solver.

For every mixed-precision solve:
* it reports convergence;
* the true relative residual ||b - A x|| / ||b||, recomputed here with compensated
  FP64 rows (hpb_csr_residualc_f64) and compensated norms, is at most 1e-12 and
  agrees with the one the solver reports (the reference reports this quantity
  after its solves, source/test_csr_linear_solver.cpp:308-317);
* the solution differs from the FP64 solve of the same solver by no more than
  ten times what the same settings gave on MI355X (1.5e-11 / 4.2e-12 on
  convection-diffusion, 1.9e-9 / 1.7e-8 on the anisotropic matrix; the error of
  either solve against x* is about 1e-8 there).

The inner-solver settings are the ones the benchmark prices (inner tolerance
1e-4 / 1e-2, compensated FP32 rows on the anisotropic matrix, reliable updates
at delta 0.1, restarted BiCGStab(2) refinement at 1e-2).
"""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
import high_precision_blas1_amd.solvers as S
from conftest import say

pytestmark = pytest.mark.gpu


def _nrm(handle, n, v):
    o = np.zeros(4)
    H.call("hpb_nrm2c_f64", handle.ptr, n, v.ptr, o.ctypes.data_as(ctypes.c_void_p))
    return float(o[3])


def _true_rel(handle, A, b, x):
    r = H.DeviceArray(A.rows, np.float64, handle)
    H.call("hpb_csr_residualc_f64", handle.ptr, A.rows, A.rows, A.nnz, A.rp.ptr, A.col.ptr, A.val.ptr, b.ptr, x.ptr,
           r.ptr)
    out = _nrm(handle, A.rows, r) / _nrm(handle, A.rows, b)
    r.free()
    return out


def _rel_diff(handle, n, a, b):
    e = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_axpbyc_f64", handle.ptr, n, ctypes.c_double(1.0), a.ptr, ctypes.c_double(-1.0), b.ptr, e.ptr)
    out = _nrm(handle, n, e) / _nrm(handle, n, b)
    e.free()
    return out


def _zeros(handle, n):
    x = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_fill_f64", handle.ptr, n, ctypes.c_double(0.0), x.ptr)
    return x


G = dict(solver=S.GMRES, high_precision=1, restart=30)
B = dict(solver=S.BICGSTABL, high_precision=1, basis=2)

CASES = {
    "convdiff128": dict(
        nx=128, coef=S.CONVDIFF7,
        fp64={"gmres": S.params(rel_tol=1e-12, max_iters=20000, **G),
              "bicgstab": S.params(rel_tol=1e-12, max_iters=20000, resid_recalc_freq=0, **B)},
        mixed=[("gmres", S.params(rel_tol=1e-4, max_iters=5000, **G), 1.5e-10),
               ("bicgstab", S.params(rel_tol=0.1, max_iters=20000, refinement=1, **B), 4.2e-11),
               ("bicgstab", S.params(rel_tol=1e-2, max_iters=20000, refinement=0, resid_recalc_freq=0, **B),
                4.2e-11)]),
    "aniso256": dict(
        nx=256, coef=S.ANISO7,
        fp64={"gmres": S.params(rel_tol=1e-12, max_iters=20000, **G),
              "bicgstab": S.params(rel_tol=1e-12, max_iters=20000, resid_recalc_freq=0, **B)},
        mixed=[("gmres", S.params(rel_tol=1e-2, max_iters=5000, compensated_spmv=1, **G), 1.9e-8),
               ("bicgstab", S.params(rel_tol=0.1, max_iters=40000, refinement=1, **B), 1.7e-7),
               ("bicgstab", S.params(rel_tol=1e-2, max_iters=40000, refinement=0, resid_recalc_freq=0, **B),
                1.7e-7)]),
}


@pytest.mark.parametrize("problem", ["convdiff128", "aniso256"])
def test_mixed_precision_full_size(handle, problem):
    c = CASES[problem]
    nx = c["nx"]
    say("config 3 full size:", problem)
    A = S.DeviceCSR.stencil7(handle, nx, nx, nx, c["coef"])
    n = A.rows
    xs = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_random_uniform_f64", handle.ptr, n, 2, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), xs.ptr)
    b = H.DeviceArray(n, np.float64, handle)
    A.spmv(xs, b, compensated=True)
    ref = {}
    for fam, p in c["fp64"].items():
        say("  fp64", fam)
        x = _zeros(handle, n)
        res = S.solve(handle, A, b, x, p)
        say("  ", res)
        assert res["converged"] == 1, res
        tr = _true_rel(handle, A, b, x)
        assert tr <= 1e-12 and abs(tr - res["true_rel_resid"]) <= 1e-3 * tr + 1e-16, (tr, res)
        ref[fam] = x
    for fam, p, max_diff in c["mixed"]:
        say("  mixed", fam, "refinement", p.refinement, "inner tol", p.rel_tol)
        x = _zeros(handle, n)
        res = S.solve_ir(handle, A, b, x, p, rel_tol=1e-12, max_outer=40)
        say("  ", res)
        assert res["converged"] == 1, res
        tr = _true_rel(handle, A, b, x)
        assert tr <= 1e-12, (tr, res)
        assert abs(tr - res["true_rel_resid"]) <= 1e-3 * tr + 1e-16, (tr, res)
        d = _rel_diff(handle, n, x, ref[fam])
        say("   relative difference from the FP64 solve: %.3g" % d)
        assert d <= max_diff, (d, max_diff)
        x.free()
    for v in list(ref.values()) + [xs, b]:
        v.free()
    A.free()
