"""Batched BiCGStab(2) over many 2-D log-normal diffusion systems.

* The block-diagonal builder equals, block by block, the single-system builder
  hpb_csr_diffusion2d_f64 fed with the same cell field (bit for bit), so the
  closed-form row offsets and the on-the-fly coefficients are right.
* Every system converges to the requested true relative residual, the solution
  matches x* to the accuracy the residual implies, and the graph-replayed and
  directly launched cycles give identical iterates (same kernels, same order).
* Iteration counts agree with the numpy restatement of the reference's
  BiCGStab(L) (oracle/krylov.py) within a few cycles for small systems.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import high_precision_blas1_amd as H
from conftest import say
from oracle import krylov as K

pytestmark = pytest.mark.gpu

GOLD = 0x9E3779B97F4A7C15


def make_batch(handle, sides, seed=3, rhs_seed=7):
    b = ctypes.c_void_p()
    arr = (ctypes.c_int * len(sides))(*sides)
    H.call("hpb_batch_create_diffusion2d", handle.ptr, len(sides), arr, ctypes.c_double(1.0), seed, rhs_seed,
           ctypes.byref(b))
    return b


def batch_host_csr(handle, b):
    rows, nnz, nt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    H.call("hpb_batch_info", b, ctypes.byref(rows), ctypes.byref(nnz), ctypes.byref(nt))
    rp, col, val, rhs = (ctypes.c_void_p() for _ in range(4))
    H.call("hpb_batch_csr", b, ctypes.byref(rp), ctypes.byref(col), ctypes.byref(val), ctypes.byref(rhs))
    out = []
    for ptr, n, dt in ((rp, rows.value + 1, np.int32), (col, nnz.value, np.int32), (val, nnz.value, np.float64),
                       (rhs, rows.value, np.float64)):
        a = np.empty(n, dtype=dt)
        H.call("hpb_memcpy_d2h", handle.ptr, a.ctypes.data_as(ctypes.c_void_p), ptr, a.nbytes)
        out.append(a)
    return out


def single_system(handle, m, seed):
    kap = H.DeviceArray(m * m, np.float64, handle)
    H.call("hpb_lognormal_field_f64", handle.ptr, m * m, ctypes.c_double(1.0), seed, kap.ptr)
    rows, nnz = ctypes.c_int64(), ctypes.c_int64()
    H.call("hpb_stencil_nnz", m, m, 1, ctypes.byref(rows), ctypes.byref(nnz))
    rp = H.DeviceArray(rows.value + 1, np.int32, handle)
    col = H.DeviceArray(nnz.value, np.int32, handle)
    val = H.DeviceArray(nnz.value, np.float64, handle)
    H.call("hpb_csr_diffusion2d_f64", handle.ptr, m, kap.ptr, rp.ptr, col.ptr, val.ptr)
    return rp.numpy(), col.numpy(), val.numpy()


def test_builder_matches_single_system_builder(handle):
    sides = [5, 64, 33, 2, 100]
    say("batch builder", sides)
    b = make_batch(handle, sides)
    rp, col, val, rhs = batch_host_csr(handle, b)
    row0 = nz0 = 0
    for s, m in enumerate(sides):
        srp, scol, sval = single_system(handle, m, (3 + s * GOLD) % (1 << 64))
        n = m * m
        assert np.array_equal(rp[row0:row0 + n + 1] - rp[row0], srp)
        nnz = srp[-1]
        assert np.array_equal(col[nz0:nz0 + nnz] - row0, scol)
        assert val[nz0:nz0 + nnz].tobytes() == sval.tobytes()
        row0 += n
        nz0 += nnz
    H.call("hpb_batch_destroy", b)


@pytest.mark.parametrize("use_graph,recalc", [(1, 1), (0, 1), (1, 0), (1, 3), (2, 1), (2, 0), (2, 3)])
def test_batch_solve_small(handle, use_graph, recalc):
    """recalc: true residual every cycle (1), only to confirm convergence (0), every
    third cycle (3); convergence is always declared on the true residual."""
    r = np.random.default_rng(3)
    sides = list(r.integers(8, 48, 37)) + [2, 3]
    say("batch solve", len(sides), use_graph, recalc)
    b = make_batch(handle, sides)
    H.call("hpb_batch_set_resid_recalc_freq", b, recalc)
    ns = len(sides)
    it = np.zeros(ns, dtype=np.int32)
    rr = np.zeros(ns)
    cv = np.zeros(ns, dtype=np.int32)
    secs, cyc = ctypes.c_double(), ctypes.c_int()
    H.call("hpb_batch_solve_bicgstab2", b, ctypes.c_double(1e-10), 4000, use_graph, 4,
           it.ctypes.data_as(ctypes.c_void_p), rr.ctypes.data_as(ctypes.c_void_p), cv.ctypes.data_as(ctypes.c_void_p),
           ctypes.byref(secs), ctypes.byref(cyc))
    assert cv.all(), (rr, it)
    assert rr.max() <= 1e-10
    rows = int(sum(m * m for m in sides))
    x = np.empty(rows)
    xs = np.empty(rows)
    H.call("hpb_batch_solution", b, x.ctypes.data_as(ctypes.c_void_p), xs.ctypes.data_as(ctypes.c_void_p))
    rp, col, val, rhs = batch_host_csr(handle, b)
    A = sp.csr_matrix((val, col, rp), shape=(rows, rows))
    row0 = 0
    for s, m in enumerate(sides):
        n = m * m
        sl = slice(row0, row0 + n)
        As = A[sl, sl]
        bs = rhs[sl]
        assert np.linalg.norm(bs - As @ x[sl]) <= 1.5e-10 * np.linalg.norm(bs)
        if s < 6:
            xo, ito, _ = K.bicgstabl(As, bs, np.zeros(n), 2, 1e-10, 4000, recalc=recalc)
            # same recurrences, different rounding (compensated dots, Gram-form MR
            # part, unnormalised shadow residual): BiCGStab's count drifts a little;
            # more on the recursive residual, which the true residual no longer
            # resets every cycle (and the batch confirms convergence on the true
            # residual, where the restatement stops on the recursive one)
            slack = 0.1 if recalc == 1 else 0.25
            assert abs(int(it[s]) - ito) <= max(12, slack * ito), (s, it[s], ito)
        row0 += n
    H.call("hpb_batch_destroy", b)


@pytest.mark.parametrize("recalc", [1, 3, 0])
def test_graph_direct_persistent_identical(handle, recalc):
    """Graph replay (1), direct launches (0) and the one persistent launch (2) run
    the same phase bodies in the same order: identical iteration counts and
    solutions, bit for bit."""
    sides = [30, 17, 45]
    say("batch modes identical", sides, recalc)
    outs = []
    for g in (1, 0, 2):
        b = make_batch(handle, sides)
        H.call("hpb_batch_set_resid_recalc_freq", b, recalc)
        it = np.zeros(3, dtype=np.int32)
        H.call("hpb_batch_solve_bicgstab2", b, ctypes.c_double(1e-10), 2000, g, 3, it.ctypes.data_as(ctypes.c_void_p),
               None, None, None, None)
        x = np.empty(sum(m * m for m in sides))
        H.call("hpb_batch_solution", b, x.ctypes.data_as(ctypes.c_void_p), None)
        outs.append((it.copy(), x))
        H.call("hpb_batch_destroy", b)
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0])
        assert outs[0][1].tobytes() == o[1].tobytes()


def test_persistent_cycle_cap(handle):
    """The persistent launch stops at the cycle cap like the host loop does: with
    max_iters = 10 (5 cycles) both run exactly 5 cycles and leave the same iterate."""
    sides = [40, 25, 33]
    say("batch persistent cycle cap", sides)
    outs = []
    for g in (0, 2):
        b = make_batch(handle, sides)
        it = np.zeros(3, dtype=np.int32)
        cv = np.zeros(3, dtype=np.int32)
        cyc = ctypes.c_int()
        H.call("hpb_batch_solve_bicgstab2", b, ctypes.c_double(1e-12), 10, g, 8, it.ctypes.data_as(ctypes.c_void_p),
               None, cv.ctypes.data_as(ctypes.c_void_p), None, ctypes.byref(cyc))
        assert cyc.value == 5 and not cv.any(), (g, cyc.value, cv)
        x = np.empty(sum(m * m for m in sides))
        H.call("hpb_batch_solution", b, x.ctypes.data_as(ctypes.c_void_p), None)
        outs.append((it.copy(), x))
        H.call("hpb_batch_destroy", b)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1].tobytes() == outs[1][1].tobytes()


def _solve(handle, b, ns, rel_tol=1e-10, max_iters=6000, use_graph=1, check_every=4):
    it = np.zeros(ns, dtype=np.int32)
    rr = np.zeros(ns)
    cv = np.zeros(ns, dtype=np.int32)
    H.call("hpb_batch_solve_bicgstab2", b, ctypes.c_double(rel_tol), max_iters, use_graph, check_every,
           it.ctypes.data_as(ctypes.c_void_p), rr.ctypes.data_as(ctypes.c_void_p), cv.ctypes.data_as(ctypes.c_void_p),
           None, None)
    return it, rr, cv


def test_jacobi_preconditioning(handle):
    """Jacobi right preconditioning: same true residual test on the original
    systems, fewer iterations on the heterogeneous coefficients, and x solves
    A x = b (checked on the host with the original matrix)."""
    r = np.random.default_rng(5)
    sides = [int(v) for v in r.integers(20, 70, 12)]
    ns = len(sides)
    say("batch jacobi", sides)
    b = make_batch(handle, sides)
    it0, rr0, cv0 = _solve(handle, b, ns)
    H.call("hpb_batch_set_jacobi", b, 1)
    it1, rr1, cv1 = _solve(handle, b, ns)
    it1p, rr1p, _ = _solve(handle, b, ns, use_graph=2)   # persistent launch, same iterates
    assert np.array_equal(it1p, it1) and np.array_equal(rr1p, rr1)
    assert cv0.all() and cv1.all()
    assert rr1.max() <= 1e-10
    assert it1.sum() < it0.sum(), (it0, it1)
    rows = sum(m * m for m in sides)
    x = np.empty(rows)
    H.call("hpb_batch_solution", b, x.ctypes.data_as(ctypes.c_void_p), None)
    rp, col, val, rhs = batch_host_csr(handle, b)
    A = sp.csr_matrix((val, col, rp), shape=(rows, rows))
    row0 = 0
    for m in sides:
        sl = slice(row0, row0 + m * m)
        assert np.linalg.norm(rhs[sl] - A[sl, sl] @ x[sl]) <= 1.5e-10 * np.linalg.norm(rhs[sl])
        row0 += m * m
    # switching back re-records the cycle: the unpreconditioned iterates again
    H.call("hpb_batch_set_jacobi", b, 0)
    it2, rr2, cv2 = _solve(handle, b, ns)
    assert np.array_equal(it2, it0) and np.array_equal(rr2, rr0)
    H.call("hpb_batch_destroy", b)


def test_tolerance_change_and_uneven_convergence(handle):
    """A second solve with another tolerance re-records the captured cycle; with
    very uneven sizes the small systems finish first and the rest still converge
    (the tile list shrinks as systems finish)."""
    sides = [2, 3, 90, 8, 61, 4]
    ns = len(sides)
    say("batch tolerance change", sides)
    b = make_batch(handle, sides)
    it_a, rr_a, cv_a = _solve(handle, b, ns, rel_tol=1e-6)
    it_b, rr_b, cv_b = _solve(handle, b, ns, rel_tol=1e-11)
    assert cv_a.all() and cv_b.all()
    assert rr_a.max() <= 1e-6 and rr_b.max() <= 1e-11
    assert (it_b >= it_a).all() and it_b.sum() > it_a.sum()
    H.call("hpb_batch_destroy", b)


def test_recalc_frequency_change(handle):
    """Changing the residual recalculation frequency between solves of one batch:
    every setting converges on the true residual, the captured cycles of each kind
    are recorded when first needed, and returning to a frequency gives the same
    iterates again."""
    sides = [12, 40, 25, 7]
    ns = len(sides)
    say("batch recalc change", sides)
    b = make_batch(handle, sides)
    res = {}
    for freq in (1, 0, 4, 1, 0):
        H.call("hpb_batch_set_resid_recalc_freq", b, freq)
        it, rr, cv = _solve(handle, b, ns)
        assert cv.all() and rr.max() <= 1e-10, (freq, it, rr)
        if freq in res:
            assert np.array_equal(res[freq][0], it) and np.array_equal(res[freq][1], rr), freq
        res[freq] = (it, rr)
    with pytest.raises(H.HPBError):
        H.call("hpb_batch_set_resid_recalc_freq", b, -1)
    H.call("hpb_batch_destroy", b)
