"""CPU oracle for the compensated BLAS1 hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The GPU library
(``high_precision_blas1_amd``) never imports it and has no CPU fallback.

``lib()`` loads ``liboracle.so`` (the C restatement in ``hpb_oracle.c``);
``ref()`` loads ``_ref/libhpbref.so`` (the reference's own CPU headers compiled
in place) when it was built, else returns None.

Pinning status of each oracle function (see tests/test_oracle_*.py):
  two_sum, two_prod ........ exact identities in rational arithmetic
  dot2, dot_naive, sum2 .... bit-exact vs the reference build (_ref) + exact bounds
  sum_kahan ................ Kahan's bound |err| <= 2u sum|x| + O(n u^2) sum|x|; exact sums
  threaded dot2/sum2 ....... bit-exact vs _ref at 1 part; ORO bound otherwise
  refgpu-order Dot2/Sum2 ... equals sequential Dot2 when the tree is trivial;
                             ORO block bound; exact on integer-valued inputs
  error bounds ............. value-exact vs _ref error_bounds.hpp
  exact dot / sum .......... math.fsum (correctly rounded) and fractions.Fraction
  philox4x32_10 ............ Random123 known-answer vectors
  gendot ................... exact dot recomputed with Fraction; grid/range invariants
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_sz = ctypes.c_size_t


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def lib():
    """Load (building on first use if needed) the oracle's shared library."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        from oracle.build_oracle import build_oracle
        build_oracle(verbose=False)
    L = ctypes.CDLL(path)
    d = ctypes.c_double
    _sig(L, "orc_two_sum", d, [d, d, _dp])
    _sig(L, "orc_two_prod", d, [d, d, _dp])
    for nm in ("orc_dot_naive", "orc_dot2"):
        _sig(L, nm, d, [_sz, _dp, _dp])
    for nm in ("orc_dot2_f32", "orc_dot_naive_f32"):
        _sig(L, nm, ctypes.c_float, [_sz, _fp, _fp])
    _sig(L, "orc_sum2_f32", ctypes.c_float, [_sz, _fp])
    for nm in ("orc_sum_naive", "orc_sum_kahan", "orc_sum2", "orc_asum2", "orc_exact_sum"):
        _sig(L, nm, d, [_sz, _dp])
    for nm in ("orc_dot2_threaded", "orc_dot_naive_threaded"):
        _sig(L, nm, d, [_sz, _dp, _dp, ctypes.c_int])
    for nm in ("orc_sum2_threaded", "orc_asum2_threaded"):
        _sig(L, nm, d, [_sz, _dp, ctypes.c_int])
    _sig(L, "orc_dot2_refgpu_order", d, [_sz, _dp, _dp, ctypes.c_int])
    _sig(L, "orc_dot_bounds", None, [_sz, d, _sz, d, _dp])
    _sig(L, "orc_sum_bounds", None, [_sz, d, _sz, d, _dp])
    _sig(L, "orc_exact_dot", d, [_sz, _dp, _dp, _dp])
    _sig(L, "orc_exact_dot_f32", d, [_sz, _fp, _fp])
    _sig(L, "orc_rel_err_dd", d, [d, d, d])
    _sig(L, "orc_philox4x32_10", None, [_u32p, _u32p, _u32p])
    _sig(L, "orc_gendot", ctypes.c_int, [_sz, d, ctypes.c_uint64, _sz, _dp, _dp, _u64p, _dp])
    _sig(L, "orc_acc256_minus", d, [_u64p, d])
    _sig(L, "orc_acc256_round", d, [_u64p])
    _sig(L, "orc_hpb_order_reduce", ctypes.c_int,
         [ctypes.c_int, _sz, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp])
    _sig(L, "orc_hpb_order_reduce_f32", ctypes.c_int,
         [ctypes.c_int, _sz, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp])
    _sig(L, "orc_hpb_order_nrm2", ctypes.c_int,
         [ctypes.c_int, _sz, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp])
    _sig(L, "orc_hpb_order_nrm2_f32", ctypes.c_int,
         [ctypes.c_int, _sz, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp])
    _sig(L, "orc_zdot_order", ctypes.c_int,
         [ctypes.c_int, _sz, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp])
    _sig(L, "orc_zdot_order_f32", ctypes.c_int,
         [ctypes.c_int, _sz, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp])
    vp = ctypes.c_void_p
    _sig(L, "orc_map", None, [ctypes.c_int, _sz, vp, vp, vp, vp, vp, vp])
    _sig(L, "orc_map_f32", None, [ctypes.c_int, _sz, vp, vp, vp, vp, vp, vp])
    _sig(L, "orc_csr_spmv", None, [ctypes.c_int, vp, vp, vp, vp, d, d, vp, vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int])
    _LIB = L
    return L


def ref():
    """The reference's CPU code built by build_oracle.build_ref(), or None."""
    global _REF
    if _REF is not None:
        return _REF
    path = os.path.join(HERE, "_ref", "libhpbref.so")
    if not os.path.exists(path):
        return None
    R = ctypes.CDLL(path)
    d = ctypes.c_double
    for nm in ("ref_dot2", "ref_dot_naive"):
        _sig(R, nm, d, [_sz, _dp, _dp])
    _sig(R, "ref_dot2_f32", ctypes.c_float, [_sz, _fp, _fp])
    for nm in ("ref_dot2_threaded", "ref_dot_naive_threaded"):
        _sig(R, nm, d, [_sz, _dp, _dp, ctypes.c_int])
    for nm in ("ref_sum2_threaded", "ref_asum2_threaded"):
        _sig(R, nm, d, [_sz, _dp, ctypes.c_int])
    _sig(R, "ref_dot_bounds", None, [_sz, d, _sz, _dp])
    _sig(R, "ref_sum_bounds", None, [_sz, d, _sz, _dp])
    _REF = R
    return R


def ptr(a):
    """ctypes pointer to a contiguous numpy array."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.uint64:
        return a.ctypes.data_as(_u64p)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(_u32p)
    raise TypeError(a.dtype)


# -- thin numpy-level helpers -------------------------------------------------

def dot2(x, y):
    return lib().orc_dot2(x.size, ptr(x), ptr(y))


def dot_naive(x, y):
    return lib().orc_dot_naive(x.size, ptr(x), ptr(y))


def sum2(x):
    return lib().orc_sum2(x.size, ptr(x))


def asum2(x):
    return lib().orc_asum2(x.size, ptr(x))


def exact_dot(x, y):
    """Exact x.y as (nearest double, residual double)."""
    out = np.zeros(2)
    lib().orc_exact_dot(x.size, ptr(x), ptr(y), ptr(out))
    return float(out[0]), float(out[1])


def exact_sum(x):
    return lib().orc_exact_sum(x.size, ptr(x))


def rel_err(res, exact_dd):
    return lib().orc_rel_err_dd(float(res), float(exact_dd[0]), float(exact_dd[1]))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(o))
    return [int(v) for v in o]


def dot_bounds(n, cond, k_blocks, eps=2.0 ** -53):
    o = np.zeros(6)
    lib().orc_dot_bounds(n, cond, k_blocks, eps, ptr(o))
    return o


def sum_bounds(n, cond, k_blocks, eps=2.0 ** -53):
    o = np.zeros(6)
    lib().orc_sum_bounds(n, cond, k_blocks, eps, ptr(o))
    return o


def gendot(n, cond, seed, n_tail=None):
    """Host GenDot pair: returns x, y, exact limbs (uint64[4]), (dot, sum|xy|, kappa)."""
    if n_tail is None:
        n_tail = n // 2
    x = np.empty(n)
    y = np.empty(n)
    limbs = np.zeros(4, dtype=np.uint64)
    dd = np.zeros(3)
    rc = lib().orc_gendot(n, float(np.log2(cond)), seed, n_tail, ptr(x), ptr(y), ptr(limbs), ptr(dd))
    if rc != 0:
        raise ValueError("orc_gendot failed with code %d" % rc)
    return x, y, limbs, dd


def limbs_to_fraction(limbs):
    """Exact value of a 256-bit two's complement fixed-point number with LSB 2^-106."""
    from fractions import Fraction
    v = 0
    for i in range(4):
        v |= int(limbs[i]) << (64 * i)
    if v >> 255:
        v -= 1 << 256
    return Fraction(v, 1 << 106)


def ori_bound(n, x, y, u=2.0 ** -53):
    """Acceptance bound u|x.y| + gamma_{n-1}^2 |x|.|y| of a compensated dot (floats, generous)."""
    g = (n - 1) * u / (1 - (n - 1) * u)
    ad = float(np.dot(np.abs(x), np.abs(y)))
    return g * g * ad


REDUCE_OPS = {"dot2": 0, "dot": 1, "sum2": 2, "sum": 3, "asum2": 4, "asum": 5, "nrm2c": 6, "nrm2": 7}


def hpb_order_reduce(op, x, y=None, cfg=(256, 4, 1, 2)):
    """(s, c, s+c) of reduction ``op`` in the GPU library's order for launch shape
    cfg = (block, unroll, grid, vecw)."""
    x = np.ascontiguousarray(x)
    y = x if y is None else np.ascontiguousarray(y)
    out = np.zeros(3, dtype=x.dtype)
    f = lib().orc_hpb_order_reduce if x.dtype == np.float64 else lib().orc_hpb_order_reduce_f32
    rc = f(REDUCE_OPS[op], x.size, ptr(x), ptr(y), int(cfg[0]), int(cfg[1]), int(cfg[2]), int(cfg[3]), ptr(out))
    if rc:
        raise ValueError("bad launch shape %r" % (cfg,))
    return out


def hpb_order_nrm2(op, x, cfg=(256, 4, 1, 2)):
    """(s, c, s+c, norm) of the library's overflow- and underflow-safe norm ``op``
    ("nrm2c" or "nrm2"): unscaled when max |x_i| lies in the safe range, else the
    same order over x_i 2^k, scaled back."""
    x = np.ascontiguousarray(x)
    out = np.zeros(4, dtype=x.dtype)
    f = lib().orc_hpb_order_nrm2 if x.dtype == np.float64 else lib().orc_hpb_order_nrm2_f32
    rc = f(REDUCE_OPS[op], x.size, ptr(x), int(cfg[0]), int(cfg[1]), int(cfg[2]), int(cfg[3]), ptr(out))
    if rc:
        raise ValueError("bad op or launch shape %r %r" % (op, cfg))
    return out


def zdot_order(x, y, comp=True, cfg=(256, 4, 1)):
    """x^H y of complex vectors in the GPU library's order for launch shape
    cfg = (block, unroll, grid): (s_re, c_re, s_im, c_im, re, im)."""
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y)
    real = np.float64 if x.dtype == np.complex128 else np.float32
    xr, yr = x.view(real), y.view(real)
    out = np.zeros(6, dtype=real)
    f = lib().orc_zdot_order if real == np.float64 else lib().orc_zdot_order_f32
    rc = f(1 if comp else 0, x.size, ptr(xr), ptr(yr), int(cfg[0]), int(cfg[1]), int(cfg[2]), ptr(out))
    if rc:
        raise ValueError("bad launch shape %r" % (cfg,))
    return out


MAP_OPS = {"fill": 0, "add_mul_scalar": 1, "scal": 2, "copy": 3, "abs": 4, "axpy": 5, "axpby": 6,
           "axpbypcz": 7, "axpbypcwpdz": 8, "mulpw": 9, "mulpw_ip": 10, "divpw": 11, "divpw_ip": 12,
           "axpbyc": 13}


def map_op(op, n, scalars, inputs, dtype=np.float64):
    """Elementwise op with the library's documented rounding (see hpb_oracle.c)."""
    dtype = np.dtype(dtype)
    s = np.zeros(4, dtype=dtype)
    s[:len(scalars)] = scalars
    ins = [None if a is None else np.ascontiguousarray(a, dtype=dtype) for a in list(inputs) + [None] * (4 - len(inputs))]
    out = np.empty(n, dtype=dtype)
    f = lib().orc_map if dtype == np.float64 else lib().orc_map_f32
    f(MAP_OPS[op], n, s.ctypes.data, *[None if a is None else a.ctypes.data for a in ins], out.ctypes.data)
    return out


def csr_spmv(A, x, alpha=1.0, beta=0.0, y=None, lpr=1, comp=False, b=None):
    """y = alpha A x + beta y (or b - A x when b is given) in the library's row
    order; A is a scipy.sparse CSR matrix with int32 indices."""
    rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    col = np.ascontiguousarray(A.indices, dtype=np.int32)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    rows = A.shape[0]
    out = np.zeros(rows) if y is None else np.array(y, dtype=np.float64)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
    lib().orc_csr_spmv(rows, rp.ctypes.data, col.ctypes.data, val.ctypes.data, x.ctypes.data, alpha, beta,
                       None if bb is None else bb.ctypes.data, out.ctypes.data, int(lpr), int(comp),
                       0 if b is None else 1)
    return out


def stencil7(nx, ny, nz, coef):
    """CSR of a constant 7-point stencil built independently with scipy Kronecker
    products; coef = (z-, y-, x-, diag, x+, y+, z+), lexicographic i fastest."""
    import scipy.sparse as sp
    zm, ym, xm, d, xp, yp, zp = coef
    def one(n, lo, up):
        return sp.diags([np.full(n - 1, lo), np.full(n - 1, up)], [-1, 1], shape=(n, n), format="csr")
    Ix, Iy, Iz = (sp.identity(m, format="csr") for m in (nx, ny, nz))
    A = sp.kron(Iz, sp.kron(Iy, one(nx, xm, xp))) + sp.kron(Iz, sp.kron(one(ny, ym, yp), Ix))
    if nz > 1:
        A = A + sp.kron(one(nz, zm, zp), sp.kron(Iy, Ix))
    A = (A + d * sp.identity(nx * ny * nz)).tocsr()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A
