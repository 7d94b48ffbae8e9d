"""GPU CSR: structured-matrix builders bit-equal to independent scipy
construction, SpMV / residual bit-exact against the oracle's row order for every
lanes-per-row split, and size-independent checks at 256^3 (closed-form A*1,
exact power-of-two linearity)."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import high_precision_blas1_amd as H
import oracle as O
from conftest import say

pytestmark = pytest.mark.gpu

POISSON = (-1.0, -1.0, -1.0, 6.0, -1.0, -1.0, -1.0)
# first-order upwind convection-diffusion, velocity (1, 0.5, 0.25), cell Peclet 100,
# scaled by h^2/nu: lower neighbour -1 - Pe_d, upper -1, diagonal 6 + sum Pe_d
CONVDIFF = (-26.0, -51.0, -101.0, 181.0, -1.0, -1.0, -1.0)
ANISO = (-1e4, -1.0, -1.0, 4.0 + 2e4, -1.0, -1.0, -1e4)


class DevCSR:
    def __init__(self, handle, rows, nnz):
        self.rows, self.nnz = rows, nnz
        self.rp = H.DeviceArray(rows + 1, np.int32, handle)
        self.col = H.DeviceArray(nnz, np.int32, handle)
        self.val = H.DeviceArray(nnz, np.float64, handle)

    def free(self):
        for a in (self.rp, self.col, self.val):
            a.free()

    def host(self):
        return sp.csr_matrix((self.val.numpy(), self.col.numpy(), self.rp.numpy()), shape=(self.rows, self.rows))


def build_stencil(handle, nx, ny, nz, coef):
    rows, nnz = ctypes.c_int64(), ctypes.c_int64()
    H.call("hpb_stencil_nnz", nx, ny, nz, ctypes.byref(rows), ctypes.byref(nnz))
    A = DevCSR(handle, rows.value, nnz.value)
    c = np.array(coef, dtype=np.float64)
    H.call("hpb_csr_stencil7_f64", handle.ptr, nx, ny, nz, c.ctypes.data_as(ctypes.c_void_p), A.rp.ptr, A.col.ptr,
           A.val.ptr)
    return A


def upload_csr(handle, M):
    A = DevCSR.__new__(DevCSR)
    A.rows, A.nnz = M.shape[0], M.nnz
    A.rp = H.DeviceArray.from_numpy(M.indptr.astype(np.int32), handle)
    A.col = H.DeviceArray.from_numpy(M.indices.astype(np.int32), handle)
    A.val = H.DeviceArray.from_numpy(M.data.astype(np.float64), handle)
    return A


def spmv(handle, A, x, alpha=1.0, beta=0.0, y=None, lpr=0, comp=False):
    if y is None:
        y = H.DeviceArray(A.rows, np.float64, handle)
    H.call("hpb_csr_spmv_lpr_f64", handle.ptr, lpr, int(comp), A.rows, A.rows, A.nnz, A.rp.ptr, A.col.ptr,
           A.val.ptr, ctypes.c_double(alpha), x.ptr, ctypes.c_double(beta), y.ptr)
    return y


@pytest.mark.parametrize("dims", [(7, 5, 3), (16, 16, 16), (33, 17, 1), (1, 1, 9), (64, 1, 1)])
@pytest.mark.parametrize("coef", [POISSON, CONVDIFF, ANISO])
def test_stencil_builder_matches_scipy(handle, dims, coef):
    say("stencil", dims, coef)
    A = build_stencil(handle, *dims, coef)
    ref = O.stencil7(*dims, coef)
    got = A.host()
    assert np.array_equal(got.indptr, ref.indptr)
    assert np.array_equal(got.indices, ref.indices)
    assert got.data.tobytes() == ref.data.astype(np.float64).tobytes()
    A.free()


def test_diffusion2d_builder(handle):
    m = 40
    kap = H.DeviceArray(m * m, np.float64, handle)
    H.call("hpb_lognormal_field_f64", handle.ptr, m * m, ctypes.c_double(1.0), 3, kap.ptr)
    k = kap.numpy()
    assert np.all(k > 0)
    rows, nnz = ctypes.c_int64(), ctypes.c_int64()
    H.call("hpb_stencil_nnz", m, m, 1, ctypes.byref(rows), ctypes.byref(nnz))
    A = DevCSR(handle, rows.value, nnz.value)
    H.call("hpb_csr_diffusion2d_f64", handle.ptr, m, kap.ptr, A.rp.ptr, A.col.ptr, A.val.ptr)
    got = A.host()
    # independent host assembly from the same field
    K = k.reshape(m, m)  # [j, i]
    def face(a, b):
        return 2.0 * a * b / (a + b)
    rows_l, cols_l, vals_l = [], [], []
    for j in range(m):
        for i in range(m):
            r = i + m * j
            kc = K[j, i]
            fs = face(kc, K[j - 1, i]) if j > 0 else kc
            fw = face(kc, K[j, i - 1]) if i > 0 else kc
            fe = face(kc, K[j, i + 1]) if i < m - 1 else kc
            fn = face(kc, K[j + 1, i]) if j < m - 1 else kc
            ent = []
            if j > 0: ent.append((r - m, -fs))
            if i > 0: ent.append((r - 1, -fw))
            ent.append((r, ((fs + fw) + fe) + fn))
            if i < m - 1: ent.append((r + 1, -fe))
            if j < m - 1: ent.append((r + m, -fn))
            for c_, v_ in ent:
                rows_l.append(r); cols_l.append(c_); vals_l.append(v_)
    ref = sp.csr_matrix((vals_l, (rows_l, cols_l)), shape=(m * m, m * m))
    ref.sort_indices()
    assert np.array_equal(got.indices, ref.indices) and np.array_equal(got.indptr, ref.indptr)
    assert got.data.tobytes() == np.asarray(ref.data, dtype=np.float64).tobytes()
    assert abs(got - got.T).max() == 0.0
    big = H.DeviceArray(1 << 20, np.float64, handle)
    H.call("hpb_lognormal_field_f64", handle.ptr, 1 << 20, ctypes.c_double(1.0), 3, big.ptr)
    g = np.log(big.numpy())
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01


def default_lpr(M):
    """The library's lanes per row for lpr = 0: about two entries per lane."""
    mean = M.nnz / M.shape[0]
    for lim, lanes in ((1.5, 1), (8, 2), (12, 4), (24, 8), (48, 16), (96, 32)):
        if mean <= lim:
            return lanes
    return 64


@pytest.mark.parametrize("lpr", [0, 1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("comp", [False, True])
def test_spmv_bitexact(handle, lpr, comp):
    r = np.random.default_rng(lpr + 10 * comp)
    M = sp.random(3000, 3000, density=0.004, format="csr", random_state=lpr)
    M.data = r.standard_normal(M.nnz) * np.exp2(r.integers(-20, 20, M.nnz))
    M = M.tocsr()
    M.sort_indices()
    # staged tiles (up to 4 lanes per row): a bidiagonal matrix (about 2 entries per
    # row: one lane's tiles of 256 rows fit), and rows of uneven length with empty
    # rows, so that some 128-row tiles are staged and others read directly
    B2 = sp.diags([r.standard_normal(2500), r.standard_normal(2499)], [0, -1], format="csr")
    lens = np.where(np.arange(3000) < 700, 20, r.integers(0, 4, 3000))
    U = sp.lil_matrix((3000, 3000))
    for i, ln in enumerate(lens):
        cols = r.choice(3000, size=int(ln), replace=False)
        U[i, cols] = r.standard_normal(int(ln))
    U = U.tocsr()
    for Z in (B2, U):
        Z.sort_indices()
    cases = [M, O.stencil7(12, 10, 9, CONVDIFF), B2, U]
    for Mh in cases:
        say("spmv", lpr, comp, Mh.shape, Mh.nnz)
        A = upload_csr(handle, Mh)
        x = r.standard_normal(Mh.shape[0])
        dx = H.DeviceArray.from_numpy(x, handle)
        eff = lpr if lpr else default_lpr(Mh)
        y = spmv(handle, A, dx, lpr=lpr, comp=comp).numpy()
        assert y.tobytes() == O.csr_spmv(Mh, x, lpr=eff, comp=comp).tobytes()
        y0 = r.standard_normal(Mh.shape[0])
        dy = H.DeviceArray.from_numpy(y0, handle)
        spmv(handle, A, dx, alpha=-0.5, beta=2.0, y=dy, lpr=lpr, comp=comp)
        assert dy.numpy().tobytes() == O.csr_spmv(Mh, x, -0.5, 2.0, y0, lpr=eff, comp=comp).tobytes()
        A.free()


def test_residual_bitexact(handle):
    Mh = O.stencil7(20, 20, 20, CONVDIFF)
    A = upload_csr(handle, Mh)
    r = np.random.default_rng(2)
    x, b = r.standard_normal(Mh.shape[0]), r.standard_normal(Mh.shape[0])
    dx, db = H.DeviceArray.from_numpy(x, handle), H.DeviceArray.from_numpy(b, handle)
    out = H.DeviceArray(Mh.shape[0], np.float64, handle)
    for name, comp in (("hpb_csr_residual_f64", False), ("hpb_csr_residualc_f64", True)):
        H.call(name, handle.ptr, A.rows, A.rows, A.nnz, A.rp.ptr, A.col.ptr, A.val.ptr, db.ptr, dx.ptr, out.ptr)
        assert out.numpy().tobytes() == O.csr_spmv(Mh, x, lpr=default_lpr(Mh), comp=comp, b=b).tobytes()


def test_poisson_256_cubed_properties(handle):
    nx = 256
    say("poisson 256^3")
    A = build_stencil(handle, nx, nx, nx, POISSON)
    n = A.rows
    assert n == nx ** 3 and A.nnz == n + 6 * (nx - 1) * nx * nx
    one = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_fill_f64", handle.ptr, n, ctypes.c_double(1.0), one.ptr)
    for comp in (False, True):
        y = spmv(handle, A, one, comp=comp).numpy()
        i = np.arange(n) % nx
        j = (np.arange(n) // nx) % nx
        k = np.arange(n) // (nx * nx)
        missing = ((i == 0).astype(np.int8) + (i == nx - 1) + (j == 0) + (j == nx - 1) + (k == 0) + (k == nx - 1))
        assert np.array_equal(y, missing.astype(np.float64))
    x = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_random_uniform_f64", handle.ptr, n, 1, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), x.ptr)
    y1 = spmv(handle, A, x, comp=True).numpy()
    H.call("hpb_scal_f64", handle.ptr, n, ctypes.c_double(2.0), x.ptr)
    y2 = spmv(handle, A, x, comp=True).numpy()
    assert np.array_equal(y2, 2.0 * y1)
    A.free()


def test_spmv_256_cubed_bitexact(handle):
    """At 256^3 the staged kernel's grid is capped (256 blocks per CU), so blocks form
    several 128-row tiles strided over their XCD's slice: every row still equals the
    oracle's row-order emulation, bit for bit, plain and compensated, 2 and 4 lanes."""
    from types import SimpleNamespace
    nx = 256
    say("spmv 256^3 bit-exact")
    A = build_stencil(handle, nx, nx, nx, CONVDIFF)
    n = A.rows
    Mh = SimpleNamespace(indptr=A.rp.numpy(), indices=A.col.numpy(), data=A.val.numpy(), shape=(n, n))
    x = H.DeviceArray(n, np.float64, handle)
    H.call("hpb_random_uniform_f64", handle.ptr, n, 5, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), x.ptr)
    xh = x.numpy()
    for lpr in (0, 4):
        for comp in (False, True):
            say("spmv 256^3", lpr, comp)
            y = spmv(handle, A, x, lpr=lpr, comp=comp).numpy()
            ref = O.csr_spmv(Mh, xh, lpr=lpr if lpr else 2, comp=comp)
            assert y.tobytes() == ref.tobytes()
    A.free()


def test_spmv_f32_256_cubed_integer_exact(handle):
    """FP32 rows at 256^3 (tiles of 256 rows, several per block): Poisson values and
    integer x are exact in FP32, so y must equal the integer stencil product."""
    nx = 256
    say("spmv f32 256^3 integer")
    A = build_stencil(handle, nx, nx, nx, POISSON)
    n = A.rows
    valf = H.DeviceArray(A.nnz, np.float32, handle)
    H.call("hpb_convert_f64_to_f32", handle.ptr, A.nnz, A.val.ptr, valf.ptr)
    r = np.random.default_rng(31)
    x = r.integers(-8, 9, n).astype(np.float32)
    g = x.reshape(nx, nx, nx).astype(np.int64)   # [k, j, i]
    exact = 6 * g
    exact[1:] -= g[:-1]
    exact[:-1] -= g[1:]
    exact[:, 1:] -= g[:, :-1]
    exact[:, :-1] -= g[:, 1:]
    exact[:, :, 1:] -= g[:, :, :-1]
    exact[:, :, :-1] -= g[:, :, 1:]
    dx = H.DeviceArray.from_numpy(x, handle)
    dy = H.DeviceArray(n, np.float32, handle)
    for name in ("hpb_csr_spmv_f32", "hpb_csr_spmvc_f32"):
        H.call(name, handle.ptr, n, n, A.nnz, A.rp.ptr, A.col.ptr, valf.ptr, ctypes.c_float(1.0), dx.ptr,
               ctypes.c_float(0.0), dy.ptr)
        assert np.array_equal(dy.numpy(), exact.reshape(-1).astype(np.float32)), name
    for a in (valf, dx, dy):
        a.free()
    A.free()


@pytest.mark.parametrize("comp", [False, True])
def test_spmv_f32_integer_exact(handle, comp):
    """FP32 rows (the staged tiles of the mixed-precision inner solves) on integer data:
    every product and partial sum is exact in FP32, so y must equal the integer
    product whatever the order; a dropped, repeated or misplaced entry shows."""
    r = np.random.default_rng(17 + comp)
    cases = [O.stencil7(12, 10, 9, POISSON)]
    M = sp.random(3000, 3000, density=0.003, format="csr", random_state=3)
    M.data = r.integers(-8, 9, M.nnz).astype(np.float64)
    cases.append(M)
    # rows of uneven length (20 entries in the first 700 rows, 0-3 after, some empty):
    # the first 256-row FP32 tiles exceed 2048 entries and read directly, the rest stage
    lens = np.where(np.arange(3000) < 700, 20, r.integers(0, 4, 3000))
    U = sp.lil_matrix((3000, 3000))
    for i, ln in enumerate(lens):
        cols = r.choice(3000, size=int(ln), replace=False)
        U[i, cols] = r.integers(-8, 9, int(ln)).astype(np.float64)
    cases.append(U.tocsr())
    for Mh in cases:
        Mh = Mh.tocsr()
        Mh.sort_indices()
        say("spmv f32 integer", comp, Mh.shape, Mh.nnz)
        rows = Mh.shape[0]
        x = r.integers(-8, 9, rows).astype(np.float32)
        rp = H.DeviceArray.from_numpy(Mh.indptr.astype(np.int32), handle)
        col = H.DeviceArray.from_numpy(Mh.indices.astype(np.int32), handle)
        val = H.DeviceArray.from_numpy(Mh.data.astype(np.float32), handle)
        dx = H.DeviceArray.from_numpy(x, handle)
        dy = H.DeviceArray(rows, np.float32, handle)
        H.call("hpb_csr_spmvc_f32" if comp else "hpb_csr_spmv_f32", handle.ptr, rows, rows, Mh.nnz, rp.ptr, col.ptr,
               val.ptr, ctypes.c_float(1.0), dx.ptr, ctypes.c_float(0.0), dy.ptr)
        exact = (Mh.astype(np.int64) @ x.astype(np.int64)).astype(np.float32)
        assert np.array_equal(dy.numpy(), exact)
        for a in (rp, col, val, dx, dy):
            a.free()
