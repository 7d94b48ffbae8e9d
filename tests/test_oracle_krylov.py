"""Pins the numpy restatement of the reference's GMRES(m)/MGS, BiCGStab(L) and
plain BiCGStab (oracle/krylov.py) against scipy's direct sparse solver on small nonsymmetric
problems (textbook special cases: A = I converges in one step; a diagonal
matrix with k distinct eigenvalues in k GMRES steps)."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from oracle import krylov as K

CONVDIFF = (-26.0, -51.0, -101.0, 181.0, -1.0, -1.0, -1.0)


def test_gmres_and_bicgstabl_vs_direct():
    A = O.stencil7(8, 7, 6, CONVDIFF).astype(np.float64)
    rng = np.random.default_rng(2)
    xs = rng.uniform(-1, 1, A.shape[0])
    b = A @ xs
    xd = spla.spsolve(A.tocsc(), b)
    x, it, hist = K.gmres_mgs(A, b, np.zeros_like(b), 20, 1e-12, 2000)
    assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b) and np.allclose(x, xd, atol=1e-9)
    for L in (1, 2, 4):
        x, it, hist = K.bicgstabl(A, b, np.zeros_like(b), L, 1e-12, 4000)
        assert np.linalg.norm(b - A @ x) <= 1.01e-12 * np.linalg.norm(b)
        assert np.allclose(x, xd, atol=1e-9)
    for rc in (0, 1, 5):
        x, it, hist = K.bicgstab(A, b, np.zeros_like(b), 1e-12, 4000, recalc=rc)
        assert hist[-1][1] <= 1e-12 and np.allclose(x, xd, atol=1e-9)
        if rc == 1:   # every residual is the true one
            assert np.linalg.norm(b - A @ x) <= 1.01e-12 * np.linalg.norm(b)


def test_special_cases():
    n = 50
    b = np.linspace(1, 2, n)
    x, it, _ = K.gmres_mgs(sp.identity(n, format="csr"), b, np.zeros(n), 10, 1e-14, 100)
    assert it == 1 and np.allclose(x, b)
    d = np.repeat([1.0, 2.0, 5.0, 7.0], n // 4 + 1)[:n]
    x, it, _ = K.gmres_mgs(sp.diags(d).tocsr(), b, np.zeros(n), 10, 1e-13, 100, early_exit=True)
    assert it == 4 and np.allclose(x, b / d)
    # One BiCGStab step from x0 = 0 is x1 = alpha r0 + omega s, s = r0 - alpha A r0,
    # where alpha makes s orthogonal to the shadow residual r0 (Petrov-Galerkin)
    # and omega minimises ||s - omega A s||, i.e. r1 = s - omega A s is orthogonal
    # to A s.  Recover alpha, omega from x1 = (alpha + omega) r0 - alpha omega A r0
    # and check both conditions.
    rng = np.random.default_rng(11)
    M = np.eye(6) * 4 + rng.standard_normal((6, 6))
    bb = rng.standard_normal(6)
    x1, it, _ = K.bicgstab(sp.csr_matrix(M), bb, np.zeros(6), 1e-30, 1)
    assert it == 1
    c, *_ = np.linalg.lstsq(np.stack([bb, M @ bb], axis=1), x1, rcond=None)
    roots = np.roots([1.0, -c[0], -c[1]])
    ok = False
    for alpha, omega in ((roots[0], roots[1]), (roots[1], roots[0])):
        sv = bb - alpha * (M @ bb)
        r1 = sv - omega * (M @ sv)
        if abs(np.dot(sv, bb)) < 1e-10 * np.dot(bb, bb) and abs(np.dot(r1, M @ sv)) < 1e-10 * np.dot(bb, bb):
            ok = True
    assert ok
