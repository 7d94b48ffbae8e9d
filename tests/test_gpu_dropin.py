"""The drop-in C++ layer on the GPU, through a C++ program linked against
libhpb.so (tests/cpp/test_dropin.cpp, built by build()).

The program calls every member of hpb::hip_vector_operations (the reference's
VectorOperations concept, source/common/gpu_vector_operations.h:86-538) once in
each precision mode, the reduction objects hip_reduction_ogita / hip_reduction
(gpu_reduction_ogita.h:36-96, gpu_reduction.h:36-90), the exact dot object
(dot_product_gmp.hpp:38-164), the vector-pair generator
(generate_vector_pair.hpp:40-216) and the CSR matrix's products and triangular
solves, plain and transposed (gpu_matrix.h:142-375), for FP64 and FP32.  Every
result is checked here:
* elementwise results bit for bit against oracle.map_op;
* reductions bit for bit against the library-order emulation
  (oracle.hpb_order_reduce / hpb_order_nrm2) for the handle's launch shape;
* the exact dot against the Kulisch oracle; the generator's pair bit for bit
  against the host restatement of the generator with the same seed and tail;
* matrix products bit for bit against the oracle's row-order product (FP64;
  FP32 within its rounding bound), the transposed product against the product of
  the transpose that scipy builds; triangular solves, plain and transposed,
  against scipy's dense triangular solve.
"""
import math
import os
import subprocess
from fractions import Fraction as F

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle as O
from conftest import ROOT, say

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "bin", "test_dropin")
N = 100003


def _matrix(m, seed):
    """2-D 5-point pattern on m x m with random values and a dominant diagonal."""
    r = np.random.default_rng(seed)
    n = m * m
    A = sp.lil_matrix((n, n))
    for i in range(m):
        for j in range(m):
            k = i * m + j
            A[k, k] = 8.0 + r.uniform()
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < m and 0 <= jj < m:
                    A[k, ii * m + jj] = r.uniform(-1, 1)
    return A.tocsr()


def _lpr(M):
    """The library's lanes per row: about two entries per lane."""
    mean = M.nnz / M.shape[0]
    for lim, lanes in ((1.5, 1), (8, 2), (12, 4), (24, 8), (48, 16), (96, 32)):
        if mean <= lim:
            return lanes
    return 64


@pytest.fixture(scope="module")
def run(tmp_path_factory, handle):
    d = tmp_path_factory.mktemp("dropin")
    ind, outd = d / "in", d / "out"
    ind.mkdir()
    outd.mkdir()
    r = np.random.default_rng(2024)
    data = {k: r.standard_normal(N) * np.exp2(r.integers(-6, 6, N)) for k in "xyzw"}
    data["z"] = np.where(np.abs(data["z"]) < 1e-3, 1.0, data["z"])     # divisor
    for k, v in data.items():
        v.astype(np.float64).tofile(str(ind / ("%s.f64" % k)))
    A = _matrix(24, 5)
    A.indptr.astype(np.int32).tofile(str(ind / "rp.i32"))
    A.indices.astype(np.int32).tofile(str(ind / "ci.i32"))
    A.data.astype(np.float64).tofile(str(ind / "val.f64"))
    assert os.path.exists(EXE), "build() makes %s" % EXE
    say("running", EXE)
    p = subprocess.run([EXE, str(ind), str(outd)], capture_output=True, text=True, timeout=300)
    say(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0, p.stderr

    def get(name):
        for ext, dt in (("f64", np.float64), ("f32", np.float32), ("i64", np.int64)):
            f = outd / ("%s.%s" % (name, ext))
            if f.exists():
                return np.fromfile(str(f), dtype=dt)
        raise KeyError(name)

    return data, A, get, outd


@pytest.mark.parametrize("tag,dt", [("d", np.float64), ("s", np.float32)])
def test_vector_operations_elementwise(run, tag, dt):
    data, _, get, _ = run
    x, y, z, w = (data[k].astype(dt) for k in "xyzw")
    M = lambda op, sc, ins: O.map_op(op, N, [dt(v) for v in sc], ins, dt)   # noqa: E731
    want = {
        "assign_scalar": M("fill", [1.5], []),
        "add_mul_scalar": M("add_mul_scalar", [0.75, -2.0], [x]),
        "scale": M("scal", [-3.5], [x]),
        "assign": x,
        "swap_a": y,
        "swap_b": x,
        "make_abs_copy": np.abs(x),
        "make_abs": np.abs(y),
        "assign_mul1": M("scal", [1.25], [x]),
        "add_mul1": M("axpy", [-0.5], [x, y]),
        "assign_mul2": M("axpby", [1.25, -0.75], [x, y]),
        "add_mul2": M("axpby", [1.25, -0.75], [x, y]),
        "add_mul3": M("axpbypcz", [1.25, -0.75, 0.5], [x, y, z]),
        "mul_pointwise3": M("mulpw", [1.5, -2.0], [x, y]),
        "mul_pointwise2": M("mulpw_ip", [0.0, -2.0], [x, y]),
        "div_pointwise3": M("divpw", [1.5, -2.0], [x, z]),
        "div_pointwise2": M("divpw_ip", [0.0, -2.0], [x, z]),
        "add_mul4": M("axpbypcwpdz", [1.25, -0.75, 0.5, 2.0], [x, y, w, z]),
        "assign_mul4": M("axpbypcwpdz", [1.25, -0.75, 0.5, 2.0], [x, y, z, w]),
        "assign_mul_compensated": M("axpbyc", [1.25, -0.75], [x, y]),
    }
    for name, exp in want.items():
        got = get("%s_%s" % (tag, name))
        assert got.dtype == dt and got.tobytes() == exp.astype(dt).tobytes(), name
    # single entries and validity
    sv = x.copy()
    sv[7] = 42.0
    sv[2] = -1.0
    assert get(tag + "_set_value_at_point").tobytes() == sv.tobytes()
    assert get(tag + "_get_value_at_point")[0] == 42.0
    assert get(tag + "_valid_before")[0] == 1 and get(tag + "_valid_after_nan")[0] == 0
    # host staging
    assert get(tag + "_view").tobytes() == x.tobytes()
    assert get(tag + "_set_from_buffer").tobytes() == (dt(2) * x).tobytes()
    txt = open(os.path.join(str(run[3]), tag + "_debug_view.txt")).read().split()
    assert len(txt) == N and float(txt[5]) == float(x[5])
    assert get(tag + "_size")[0] == N and get(tag + "_l2_size")[0] == int(math.sqrt(N))
    assert get(tag + "_device_location")[0] == 1 and get(tag + "_rank1_freed")[0] == 1
    # seeded random fill: (0, 1] and (a, b], the same stream for both calls
    r01, rab = get(tag + "_assign_random01"), get(tag + "_assign_random_ab")
    assert r01.min() > 0 and r01.max() <= 1 and rab.min() > -1 and rab.max() <= 1
    if dt == np.float64:
        for i in (0, 1, 77, N - 1):
            wv = O.philox([i, 0, 3, 0], [1234, 0])
            u = ((((wv[0] << 32) | wv[1]) >> 11) + 1) * 2.0 ** -53
            assert r01[i] == u and rab[i] == 2.0 * u - 1.0


@pytest.mark.parametrize("tag,dt", [("d", np.float64), ("s", np.float32)])
def test_vector_operations_reductions(run, handle, tag, dt):
    data, _, get, _ = run
    x, y, z, w = (data[k].astype(dt) for k in "xyzw")
    cfg2 = handle.reduce_config(dt, N, True, True)
    cfg1 = handle.reduce_config(dt, N, True, False)
    for p, comp in (("sp", False), ("hp", True)):
        assert get("%s_%s_precision" % (tag, p))[0] == (1 if comp else 0)
        dot = O.hpb_order_reduce("dot2" if comp else "dot", x, y, cfg2)
        asum = O.hpb_order_reduce("asum2" if comp else "asum", x, None, cfg1)
        nrm = O.hpb_order_nrm2("nrm2c" if comp else "nrm2", x, cfg1)
        g = lambda k: get("%s_%s_%s" % (tag, p, k))   # noqa: E731
        assert g("scalar_prod")[0] == dot[2] and g("scalar_prod_ptr")[0] == dot[2]
        assert g("absolute_sum")[0] == asum[2] and g("absolute_sum_ptr")[0] == asum[2]
        assert g("norm")[0] == nrm[3] and g("norm_ptr")[0] == nrm[3]
        assert g("norm_l2")[0] == nrm[3] / np.sqrt(dt(N))
        # ||(x, v)|| with one more Dot2 step on the unrounded pair
        v = dt(3.25)
        pp = v * v
        e = dt(F(float(v)) ** 2 - F(float(pp)))
        ss = nrm[0] + pp
        bs = ss - nrm[0]
        t = (pp - bs) + (nrm[0] - (ss - bs))
        assert g("norm_rank1")[0] == np.sqrt(ss + (nrm[1] + (e + t)))
        dev8 = g("dev8")
        assert dev8[:4].tobytes() == np.concatenate([dot, [dot[2]]]).astype(dt).tobytes()
        assert dev8[4:].tobytes() == nrm.tobytes()
        ny = O.hpb_order_nrm2("nrm2c" if comp else "nrm2", y, cfg1)[3]
        assert g("normalize")[0] == ny
        assert g("normalized").tobytes() == O.map_op("scal", N, [dt(1) / ny], [y], dt).tobytes()
        # fused multi-dot and multi-update: within the compensated / plain bounds of the exact values
        md = g("multi_scalar_prod")
        u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
        for j, vj in enumerate((x, y, z)):
            ex = O.exact_dot(vj.astype(np.float64), w.astype(np.float64))
            ad = float(np.dot(np.abs(vj.astype(np.float64)), np.abs(w.astype(np.float64))))
            tol = (u * abs(ex[0]) + (N * u) ** 2 * ad) if comp else (u * abs(ex[0]) + N * u * ad)
            assert abs(float(md[j]) - ex[0]) <= 2 * tol, (p, j)
        mam = g("multi_add_mul").astype(np.float64)
        ref = (w.astype(np.float64) + 0.5 * x.astype(np.float64) - 1.25 * y.astype(np.float64)
               + 2.0 * z.astype(np.float64))
        scale = np.abs(w) + 0.5 * np.abs(x) + 1.25 * np.abs(y) + 2.0 * np.abs(z)
        assert np.all(np.abs(mam - ref) <= 4 * u * scale.astype(np.float64))


@pytest.mark.parametrize("tag,dt", [("d", np.float64), ("s", np.float32)])
def test_reduction_objects_and_exact_dot(run, handle, tag, dt):
    data, _, get, _ = run
    x, y = data["x"].astype(dt), data["y"].astype(dt)
    cfg2 = handle.reduce_config(dt, N, True, True)
    cfg1 = handle.reduce_config(dt, N, True, False)
    g = lambda k: get("%s_%s" % (tag, k))[0]   # noqa: E731
    assert g("ogita_sum") == O.hpb_order_reduce("sum2", x, None, cfg1)[2]
    assert g("ogita_asum") == O.hpb_order_reduce("asum2", x, None, cfg1)[2]
    d2 = O.hpb_order_reduce("dot2", x, y, cfg2)
    assert g("ogita_dot") == d2[2] and g("ogita_dot_s") == d2[0] and g("ogita_dot_c") == d2[1]
    assert g("ogita_norm") == O.hpb_order_nrm2("nrm2c", x, cfg1)[3]
    assert g("naive_min") == x.min() and g("naive_max") == x.max()
    assert g("naive_sum") == O.hpb_order_reduce("sum", x, None, cfg1)[2] == g("naive_sum_debug")
    assert g("naive_asum") == O.hpb_order_reduce("asum", x, None, cfg1)[2]
    assert g("naive_dot") == O.hpb_order_reduce("dot", x, y, cfg2)[2]
    ex = O.exact_dot(x.astype(np.float64), y.astype(np.float64))
    assert g("exact_hi") == ex[0] and g("exact_lo") == ex[1]
    assert g("exact_dot_host") == dt(ex[0]) and g("exact_dot_dev") == dt(ex[0])
    exf = F(ex[0]) + F(ex[1])
    ap = F(float(d2[2]))
    assert float(g("err_abs")) == pytest.approx(float(abs(ap - exf)), rel=1e-6, abs=1e-300)
    assert float(g("err_rel")) == pytest.approx(float(abs((ap - exf) / exf)), rel=1e-6, abs=1e-300)
    assert g("err_abs_T") == dt(abs(float(d2[2]) - ex[0]))
    assert g("sum_two") == dt(1) + dt(1e-30)


@pytest.mark.parametrize("tag,dt", [("d", np.float64), ("s", np.float32)])
def test_generator_object(run, tag, dt):
    _, _, get, _ = run
    g = lambda k: get("%s_%s" % (tag, k))   # noqa: E731
    seed, tail = int(g("gen_seed")[0]), int(g("gen_tail")[0])
    cond = 1e16 if dt == np.float64 else 1e6
    assert seed == 5 and int(g("gen2_seed")[0]) == 6
    # the reference's tail rule: min(max(U{10000..20000}, log2(cond)/2), round(N/2))
    assert 10000 <= tail <= 20000
    hx, hy, limbs, dd = O.gendot(N, cond, seed, n_tail=tail)
    gx, gy = g("gen_x"), g("gen_y")
    assert gx.tobytes() == hx.astype(dt).tobytes() and gy.tobytes() == hy.astype(dt).tobytes()
    if dt == np.float64:
        assert g("gen_limbs").view(np.uint64).tobytes() == limbs.tobytes()
    ex = O.exact_dot(gx.astype(np.float64), gy.astype(np.float64))
    assert g("gen_exact_hi")[0] == ex[0] and g("gen_exact_lo")[0] == ex[1]
    ab = O.exact_dot(np.abs(gx).astype(np.float64), np.abs(gy).astype(np.float64))
    assert g("gen_kappa")[0] == dt(ab[0] / abs(ex[0])) == g("gen_kappa_again")[0]
    if dt == np.float64:
        assert g("gen_kappa")[0] >= cond
    assert np.isfinite(g("gen2_kappa")[0])


@pytest.mark.parametrize("tag,dt", [("d", np.float64), ("s", np.float32)])
def test_matrix_products_and_solves(run, tag, dt):
    data, A, get, _ = run
    n = A.shape[0]
    g = lambda k: get("%s_%s" % (tag, k))   # noqa: E731
    assert g("mat_rows")[0] == n and g("mat_cols")[0] == n
    Ad = A.astype(dt)
    x, y = data["x"][:n].astype(dt), data["y"][:n].astype(dt)
    # plain and transposed products: the transposed one must equal the plain
    # product of the explicitly transposed matrix, row by row in the library order
    At = sp.csr_matrix(Ad.T)
    At.sort_indices()
    for name, M in (("mat_axpy", Ad), ("mat_axpy_t", At)):
        if dt == np.float64:
            exp = O.csr_spmv(M, x, alpha=1.5, beta=-0.5, y=y, lpr=_lpr(M))
            assert g(name).tobytes() == exp.tobytes(), name
        else:
            exp = 1.5 * (M.astype(np.float64) @ x.astype(np.float64)) - 0.5 * y.astype(np.float64)
            sc = 1.5 * (abs(M.astype(np.float64)) @ np.abs(x.astype(np.float64))) + 0.5 * np.abs(y)
            assert np.all(np.abs(g(name) - exp) <= 8 * 2.0 ** -24 * sc), name
    D = Ad.toarray().astype(np.float64)
    L = np.tril(D, -1) + np.eye(n)
    U = np.triu(D)
    b = 2.0 * y.astype(np.float64)
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    for name, M in (("mat_lower", L), ("mat_upper", U), ("mat_lower_t", L.T), ("mat_upper_t", U.T)):
        got = g(name).astype(np.float64)
        # backward error of a substitution: |M got - b| <= c u (|M| |got| + |b|) row by row
        res = np.abs(M @ got - b)
        assert np.all(res <= 16 * u * (np.abs(M) @ np.abs(got) + np.abs(b))), name
        ref = sla.solve_triangular(M, b, lower=name in ("mat_lower", "mat_upper_t"))
        rt = 1e-3 if dt == np.float32 else 1e-9
        assert np.allclose(got, ref, rtol=rt, atol=rt * np.max(np.abs(ref))), name
