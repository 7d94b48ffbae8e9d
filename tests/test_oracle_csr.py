"""Pins the oracle's CSR products (oracle.csr_spmv) and the scipy stencil builder.

* Integer-valued matrices and vectors: every summation order is exact, so the
  row products must equal scipy's A @ x exactly for every lanes-per-row split.
* Poisson 7-point: A*1 equals 6 minus the number of in-box neighbours per row
  (closed form); the stencil builder's nnz matches the closed-form count.
* Compensated rows are faithful against exact rational row sums, and the plain
  rows satisfy the gamma_k bound.
"""
from fractions import Fraction as F

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O

U = 2.0 ** -53


def test_stencil_closed_forms():
    nx, ny, nz = 5, 4, 3
    A = O.stencil7(nx, ny, nz, (-1, -1, -1, 6, -1, -1, -1))
    n = nx * ny * nz
    assert A.nnz == n + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    missing = (i == 0).astype(int) + (i == nx - 1) + (j == 0) + (j == ny - 1) + (k == 0) + (k == nz - 1)
    assert np.array_equal(O.csr_spmv(A, np.ones(n)), missing.astype(np.float64))
    assert (A != A.T).nnz == 0


@pytest.mark.parametrize("lpr", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("comp", [False, True])
def test_integer_data_exact(lpr, comp, rng):
    A = sp.random(300, 280, density=0.05, format="csr", random_state=3)
    A.data = rng.integers(-50, 50, A.nnz).astype(np.float64)
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    x = rng.integers(-100, 100, 280).astype(np.float64)
    assert np.array_equal(O.csr_spmv(A, x, lpr=lpr, comp=comp), A @ x)
    b = rng.integers(-100, 100, 300).astype(np.float64)
    assert np.array_equal(O.csr_spmv(A, x, lpr=lpr, comp=comp, b=b), b - A @ x)
    y0 = rng.integers(-9, 9, 300).astype(np.float64)
    assert np.array_equal(O.csr_spmv(A, x, alpha=2.0, beta=-3.0, y=y0, lpr=lpr, comp=comp), 2 * (A @ x) - 3 * y0)


def test_compensated_rows_faithful(rng):
    n = 400
    A = sp.random(n, n, density=0.03, format="csr", random_state=5)
    A.data = rng.standard_normal(A.nnz) * np.exp2(rng.integers(-30, 30, A.nnz))
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    x = rng.standard_normal(n)
    yc = O.csr_spmv(A, x, lpr=4, comp=True)
    yp = O.csr_spmv(A, x, lpr=4, comp=False)
    for r in range(n):
        lo, hi = A.indptr[r], A.indptr[r + 1]
        ex = sum((F(float(A.data[k])) * F(float(x[A.indices[k]])) for k in range(lo, hi)), F(0))
        ab = float(sum(abs(F(float(A.data[k])) * F(float(x[A.indices[k]]))) for k in range(lo, hi)))
        m = hi - lo
        g = m * U / (1 - m * U)
        assert abs(F(yc[r]) - ex) <= F(U) * abs(ex) + F(g * g * ab * 1.01)
        assert abs(F(yp[r]) - ex) <= F(g * ab * 1.01)
