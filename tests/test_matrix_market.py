"""MatrixMarket reader/writer of the library (host code, no GPU), pinned against
scipy.io.mmread: general, symmetric, skew-symmetric, integer and pattern files,
comments, exact round trip of doubles, duplicates kept in file order, and the
reference's error cases (bad banner, index out of range, unsupported kinds)."""
import ctypes
import io

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import high_precision_blas1_amd as H
import oracle as O


def read(path):
    L = H.lib()
    rows, cols, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    rp, ci, va = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    H.call("hpb_mm_read_csr", str(path).encode(), ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz),
           ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va))
    n = rows.value
    k = nnz.value
    a_rp = np.ctypeslib.as_array(ctypes.cast(rp, ctypes.POINTER(ctypes.c_int)), (n + 1,)).copy()
    a_ci = np.ctypeslib.as_array(ctypes.cast(ci, ctypes.POINTER(ctypes.c_int)), (max(k, 1),))[:k].copy()
    a_va = np.ctypeslib.as_array(ctypes.cast(va, ctypes.POINTER(ctypes.c_double)), (max(k, 1),))[:k].copy()
    for p in (rp, ci, va):
        L.hpb_mm_free(p)
    return sp.csr_matrix((a_va, a_ci, a_rp), shape=(n, cols.value)), a_rp, a_ci, a_va


def write(path, M):
    M = M.tocsr()
    M.sort_indices()
    rp = np.ascontiguousarray(M.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(M.indices, dtype=np.int32)
    va = np.ascontiguousarray(M.data, dtype=np.float64)
    H.call("hpb_mm_write_csr", str(path).encode(), M.shape[0], M.shape[1], M.nnz, rp.ctypes.data_as(ctypes.c_void_p),
           ci.ctypes.data_as(ctypes.c_void_p), va.ctypes.data_as(ctypes.c_void_p))


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_against_scipy(tmp_path, rng, symmetry, field):
    if field == "pattern" and symmetry == "skew-symmetric":
        pytest.skip("pattern skew-symmetric is not a MatrixMarket combination")
    n = 40
    M = sp.random(n, n, density=0.1, format="coo", random_state=1)
    M.data = rng.standard_normal(M.nnz)
    if field == "integer":
        M.data = np.round(M.data * 100)
    if symmetry == "symmetric":
        M = sp.tril(M + M.T)
    if symmetry == "skew-symmetric":
        M = sp.tril(M - M.T, -1)
    p = tmp_path / "m.mtx"
    scipy.io.mmwrite(str(p), M, field=field, symmetry=symmetry)
    got, rp, ci, va = read(p)
    ref = scipy.io.mmread(str(p)).tocsr()
    ref.sort_indices()
    assert (got != ref).nnz == 0
    assert np.all(np.diff(rp) >= 0) and rp[0] == 0
    for r in range(n):
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) >= 0)


def test_round_trip_exact(tmp_path):
    M = O.stencil7(6, 5, 4, (-26.0, -51.0, -101.0, 181.0, -1.0, -1.0, -1.0)).astype(np.float64)
    M.data = M.data * np.pi / 7.0
    p = tmp_path / "a.mtx"
    write(p, M)
    got, rp, ci, va = read(p)
    assert np.array_equal(rp, M.indptr) and np.array_equal(ci, M.indices) and va.tobytes() == M.data.tobytes()


def _big(rng, n=3000, density=0.03):
    M = sp.random(n, n, density=density, format="coo", random_state=7)
    M.data = rng.standard_normal(M.nnz) * np.exp2(rng.integers(-30, 30, M.nnz))
    return M


def test_parallel_read_against_scipy(tmp_path, rng):
    """A file of several MB is parsed in parallel chunks (threads split at line ends):
    the CSR equals scipy's reading of the same file, entry for entry."""
    M = _big(rng)
    p = tmp_path / "big.mtx"
    scipy.io.mmwrite(str(p), M, field="real", symmetry="general", precision=17)
    assert p.stat().st_size > 6 << 20
    got, rp, ci, va = read(p)
    ref = scipy.io.mmread(str(p)).tocsr()
    ref.sort_indices()
    assert np.array_equal(rp, ref.indptr) and np.array_equal(ci, ref.indices)
    assert va.tobytes() == ref.data.tobytes()


def test_parallel_write_round_trip(tmp_path, rng):
    """The writer formats row ranges in parallel: a large matrix round-trips bit for
    bit, and the file equals the one-line-per-entry %.17g text."""
    M = _big(rng).tocsr()
    M.sort_indices()
    p = tmp_path / "w.mtx"
    write(p, M)
    got, rp, ci, va = read(p)
    assert np.array_equal(rp, M.indptr) and np.array_equal(ci, M.indices) and va.tobytes() == M.data.tobytes()
    lines = p.read_text().splitlines()
    assert lines[1] == "%d %d %d" % (M.shape[0], M.shape[1], M.nnz) and len(lines) == M.nnz + 2
    rows = np.repeat(np.arange(M.shape[0]), np.diff(M.indptr))
    for k in (0, 1, M.nnz // 3, M.nnz // 2, M.nnz - 1):
        assert lines[k + 2] == "%d %d %.17g" % (rows[k] + 1, M.indices[k] + 1, M.data[k])


def test_parallel_write_skewed_rows(tmp_path, rng):
    """Row ranges of about equal entry counts when the entries sit in a few rows: one
    row of 300 000 entries among empty rows, written in parallel and read back."""
    n = 400000
    cols = np.sort(rng.choice(n, 300000, replace=False)).astype(np.int32)
    rows = np.full(cols.size, 123456, dtype=np.int32)
    rows[:10] = 7
    rows[-5:] = n - 1
    M = sp.csr_matrix((rng.standard_normal(cols.size), (rows, cols)), shape=(n, n))
    M.sort_indices()
    p = tmp_path / "skew.mtx"
    write(p, M)
    got, rp, ci, va = read(p)
    assert np.array_equal(rp, M.indptr) and np.array_equal(ci, M.indices) and va.tobytes() == M.data.tobytes()


def test_irregular_lines_read_as_tokens(tmp_path, rng):
    """Lines the chunk parser does not take (an entry split over two lines, a '+'
    sign, trailing text after the last entry) send the data through the token parser:
    the matrix is the one of the regular file."""
    M = _big(rng)
    p = tmp_path / "r.mtx"
    scipy.io.mmwrite(str(p), M, field="real", symmetry="general", precision=17)
    want = read(p)
    lines = p.read_text().split("\n")
    head = next(i for i, ln in enumerate(lines) if ln and not ln.startswith("%")) + 1
    k = head + len(lines) // 2
    i, j, v = lines[k].split()
    variants = [
        lines[:k] + ["%s %s" % (i, j), v] + lines[k + 1:],
        lines[:k] + ["%s %s +%s" % (i, j, v.lstrip("+"))] + lines[k + 1:] if not v.startswith("-") else None,
        lines + ["this is not an entry"],
    ]
    for n, ls in enumerate(variants):
        if ls is None:
            continue
        q = tmp_path / ("r%d.mtx" % n)
        q.write_text("\n".join(ls))
        got = read(q)
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2]) and got[3].tobytes() == want[3].tobytes(), n


def test_comments_and_duplicates_kept(tmp_path):
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% a comment\n%\n3 3 4\n2 2 1.5\n1 3 2.0\n2 2 -0.5\n"
                 "3 1 7\n")
    got, rp, ci, va = read(p)
    assert list(rp) == [0, 1, 3, 4]
    assert list(ci) == [2, 1, 1, 0]
    assert list(va) == [2.0, 1.5, -0.5, 7.0]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "storage"),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "data type"),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", "symmetry"),
    ("%%NotMarket matrix coordinate real general\n1 1 1\n1 1 1\n", "banner"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "index"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", "index"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n", "EOF"),
])
def test_errors(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(H.HPBError) as e:
        read(p)
    assert msg in str(e.value)


def test_cpp_reader_header(tmp_path):
    """The C++ csr::matrix_market_reader compiles with g++ against the C ABI and
    reports the file's dimensions and entries (host-only calls)."""
    import os
    import subprocess
    from conftest import ROOT
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 2\n2 1 -1\n3 3 4\n")
    src = tmp_path / "t.cpp"
    src.write_text('#include "csr/matrix_market_reader.h"\n#include <cstdio>\nint main(int, char** a){'
                   'csr::matrix_market_reader<double> r; r.read_file(a[1]); int n, m; r.get_matrix_dim(n, m);'
                   'std::printf("%d %d %lld", n, m, (long long)r.nnz());'
                   'for (long long k = 0; k < r.nnz(); ++k) std::printf(" %g", r.values()[k]); return 0;}\n')
    lib = os.path.join(ROOT, "high_precision_blas1_amd")
    exe = tmp_path / "t"
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(lib, "cpp"), str(src), "-L", lib, "-lhpb",
                           "-Wl,-rpath," + lib, "-o", str(exe)])
    out = subprocess.check_output([str(exe), str(p)], text=True).split()
    assert out == ["3", "3", "4", "2", "-1", "-1", "4"]
