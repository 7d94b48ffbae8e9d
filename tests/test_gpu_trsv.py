"""Sparse triangular solves (level-scheduled) and ILU(0) on the device.

* Triangular solves against scipy.sparse.linalg.spsolve_triangular (lower with a
  unit diagonal, upper with the stored diagonal), plain and compensated rows,
  aliased x == b, FP64 and FP32; level counts against the closed form for a
  lexicographic 2-D 5-point pattern (2m - 1 wavefronts).
* ILU(0) by its defining property: (L U)_ij = a_ij on the pattern of A; for a
  tridiagonal matrix (no fill-in) L U = A exactly, so one preconditioned solve
  is exact.
* Preconditioned GMRES / BiCGStab(L) / BiCGStab through the C ABI converge to the
  true residual and need fewer iterations than without the preconditioner.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import high_precision_blas1_amd as H
import high_precision_blas1_amd.solvers as S
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu


def _dev_csr(handle, M, dtype=np.float64):
    M = sp.csr_matrix(M)
    M.sort_indices()
    return (H.DeviceArray.from_numpy(M.indptr.astype(np.int32), handle),
            H.DeviceArray.from_numpy(M.indices.astype(np.int32), handle),
            H.DeviceArray.from_numpy(M.data.astype(dtype), handle), M)


def _analysis(handle, uplo, M, rp, ci):
    t = ctypes.c_void_p()
    H.call("hpb_csr_trsv_analysis", handle.ptr, uplo, M.shape[0], M.nnz, rp.ptr, ci.ptr, ctypes.byref(t))
    return t


def _random_matrix(n, density, seed):
    r = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csr")
    M.data = r.uniform(-1, 1, M.nnz)
    return (M + sp.diags(r.uniform(2, 3, n) * np.sign(r.uniform(-1, 1, n)))).tocsr()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("comp", [0, 1])
def test_triangular_solves_vs_scipy(handle, dtype, comp):
    n = 3000
    M = _random_matrix(n, 0.002, 1)
    rp, ci, va, Ms = _dev_csr(handle, M, dtype)
    b = np.random.default_rng(2).standard_normal(n).astype(dtype)
    suf = "f64" if dtype == np.float64 else "f32"
    c = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    tol = 1e-11 if dtype == np.float64 else 2e-4
    for uplo in (0, 1):
        say("trsv", suf, comp, uplo)
        t = _analysis(handle, uplo, Ms, rp, ci)
        db = H.DeviceArray.from_numpy(b, handle)
        dx = H.DeviceArray(n, dtype, handle)
        H.call("hpb_csr_trsv_" + suf, handle.ptr, t, va.ptr, c(2.0), db.ptr, dx.ptr, comp)
        Md = Ms.astype(np.float64)
        if uplo == 0:
            T = sp.tril(Md, -1) + sp.identity(n)
        else:
            T = sp.triu(Md)
        ref = spla.spsolve_triangular(T.tocsr(), 2.0 * b.astype(np.float64), lower=(uplo == 0))
        got = dx.numpy().astype(np.float64)
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref)))
        # x aliasing b
        H.call("hpb_csr_trsv_" + suf, handle.ptr, t, va.ptr, c(2.0), db.ptr, db.ptr, comp)
        assert np.array_equal(db.numpy(), dx.numpy())
        H.call("hpb_trsv_destroy", t)


def test_levels_closed_form_and_errors(handle):
    m = 37
    A = O.stencil7(m, m, 1, (0.0, -1.0, -1.0, 4.0, -1.0, -1.0, 0.0))
    rp, ci, va, Ms = _dev_csr(handle, A)
    for uplo in (0, 1):
        t = _analysis(handle, uplo, Ms, rp, ci)
        nl = ctypes.c_int()
        H.call("hpb_csr_trsv_levels", t, ctypes.byref(nl))
        assert nl.value == 2 * m - 1
        H.call("hpb_trsv_destroy", t)
    # an upper triangle without its diagonal is structurally singular
    N = sp.csr_matrix(np.triu(np.ones((5, 5)), 1) + np.tril(np.ones((5, 5)), -1))
    rp2, ci2, va2, Ns = _dev_csr(handle, N)
    t = ctypes.c_void_p()
    with pytest.raises(H.HPBError):
        H.call("hpb_csr_trsv_analysis", handle.ptr, 1, 5, Ns.nnz, rp2.ptr, ci2.ptr, ctypes.byref(t))
    # ILU(0) needs every diagonal entry
    t = _analysis(handle, 0, Ns, rp2, ci2)
    with pytest.raises(H.HPBError):
        H.call("hpb_csr_ilu0_f64", handle.ptr, t, va2.ptr)
    H.call("hpb_trsv_destroy", t)


def test_analysis_first_error_in_dependency_order(handle):
    """The analysis checks rows in parallel ranges, yet names the error of the first
    bad row in dependency order (ascending rows for the lower triangle, descending for
    the upper), as one scan in that order does."""
    n = 300000
    T = sp.diags([np.full(n - 1, -1.0), np.full(n, 4.0), np.full(n - 1, -1.0)], [-1, 0, 1], format="csr")
    rp = T.indptr.astype(np.int32).copy()
    ci = T.indices.astype(np.int32).copy()
    lo_bad, hi_bad = 100, 250000
    a = rp[lo_bad]
    ci[a], ci[a + 1] = ci[a + 1], ci[a]          # row 100: columns out of order
    ci[rp[hi_bad + 1] - 1] = n + 5                # row 250000: a column out of range
    drp = H.DeviceArray.from_numpy(rp, handle)
    dci = H.DeviceArray.from_numpy(ci, handle)
    t = ctypes.c_void_p()
    for uplo, msg in ((0, "sorted"), (2, "sorted"), (1, "out of range"), (3, "out of range")):
        say("analysis errors", uplo)
        with pytest.raises(H.HPBError) as e:
            H.call("hpb_csr_trsv_analysis", handle.ptr, uplo, n, int(rp[-1]), drp.ptr, dci.ptr, ctypes.byref(t))
        assert msg in str(e.value), (uplo, str(e.value))
    # the same pattern without the two faults analyses, with 2 levels per 1 row
    good_ci = H.DeviceArray.from_numpy(T.indices.astype(np.int32), handle)
    t = _analysis(handle, 0, T, drp, good_ci)
    nl = ctypes.c_int()
    H.call("hpb_csr_trsv_levels", t, ctypes.byref(nl))
    assert nl.value == n
    H.call("hpb_trsv_destroy", t)


def _ilu0(handle, M):
    rp, ci, va, Ms = _dev_csr(handle, M)
    t = _analysis(handle, 0, Ms, rp, ci)
    H.call("hpb_csr_ilu0_f64", handle.ptr, t, va.ptr)
    H.call("hpb_trsv_destroy", t)
    F = sp.csr_matrix((va.numpy(), Ms.indices, Ms.indptr), shape=Ms.shape)
    L = sp.tril(F, -1) + sp.identity(M.shape[0])
    U = sp.triu(F)
    return Ms, L.tocsr(), U.tocsr()


@pytest.mark.parametrize("case", ["poisson2d", "convdiff3d", "random"])
def test_ilu0_defining_property(handle, case):
    if case == "poisson2d":
        M = O.stencil7(30, 30, 1, (0.0, -1.0, -1.0, 4.0, -1.0, -1.0, 0.0))
    elif case == "convdiff3d":
        M = O.stencil7(12, 11, 10, S.CONVDIFF7)
    else:
        M = _random_matrix(2000, 0.003, 5)
        M = (M + sp.diags(np.full(2000, 10.0))).tocsr()   # diagonally dominant: no zero pivot
    say("ilu0", case)
    Ms, L, U = _ilu0(handle, M)
    P = (L @ U).tocsr()
    Ms = Ms.tocoo()
    lu_on_pattern = np.asarray(P[Ms.row, Ms.col]).ravel()
    scale = np.max(np.abs(Ms.data))
    assert np.max(np.abs(lu_on_pattern - Ms.data)) <= 1e-12 * scale


def test_ilu0_exact_for_tridiagonal(handle):
    n = 500
    r = np.random.default_rng(3)
    M = sp.diags([r.uniform(-1, 0, n - 1), r.uniform(3, 4, n), r.uniform(-1, 0, n - 1)], [-1, 0, 1]).tocsr()
    Ms, L, U = _ilu0(handle, M)
    assert np.max(np.abs((L @ U - Ms).toarray())) <= 1e-14 * 4


@pytest.mark.parametrize("solver,arg", [(S.GMRES, 30), (S.BICGSTABL, 2), (S.BICGSTAB, 0)])
def test_ilu0_preconditioned_solvers(handle, solver, arg):
    M = O.stencil7(16, 14, 12, S.CONVDIFF7)
    A = S.DeviceCSR.from_scipy(handle, M)
    b = M @ np.random.default_rng(4).uniform(-1, 1, M.shape[0])
    say("ilu0 solve", solver, arg)
    its = []
    for pc in (S.PRECOND_NONE, S.PRECOND_ILU0):
        db = H.DeviceArray.from_numpy(b, handle)
        dx = H.DeviceArray.from_numpy(np.zeros_like(b), handle)
        p = S.params(solver=solver, restart=arg, basis=arg, rel_tol=1e-11, max_iters=5000, preconditioner=pc,
                     resid_recalc_freq=1)
        res = S.solve(handle, A, db, dx, p)
        assert res["converged"] == 1, res
        x = dx.numpy()
        # the preconditioned residual meets 1e-11; the true one is within the
        # preconditioner's conditioning of it
        assert np.linalg.norm(b - M @ x) <= 1e-9 * np.linalg.norm(b)
        its.append(res["iters"])
    assert its[1] < its[0], its
    A.free()


def test_ilu0_mixed_precision_inner(handle):
    M = O.stencil7(20, 20, 20, S.CONVDIFF7)
    A = S.DeviceCSR.from_scipy(handle, M)
    b = M @ np.random.default_rng(6).uniform(-1, 1, M.shape[0])
    db = H.DeviceArray.from_numpy(b, handle)
    dx = H.DeviceArray.from_numpy(np.zeros_like(b), handle)
    inner = S.params(solver=S.BICGSTABL, basis=2, rel_tol=1e-5, max_iters=2000, preconditioner=S.PRECOND_ILU0)
    res = S.solve_ir(handle, A, db, dx, inner, rel_tol=1e-12, max_outer=30)
    assert res["converged"] == 1 and res["true_rel_resid"] <= 1e-12
    A.free()
