// hpb_mm.cpp -- MatrixMarket coordinate files <-> zero-based int32 CSR (host).
//
// Same file semantics as csr::matrix_market_reader (source/common/csr/matrix_market_reader.h:246-430):
//   banner "%%MatrixMarket matrix coordinate <real|integer|pattern> <general|symmetric|skew-symmetric>",
//   '%' comment lines, then "rows cols entries", then 1-based "i j [value]" lines;
//   pattern entries get the value 1; symmetric files get every off-diagonal entry
//   mirrored (the mirrored copies are appended, then all entries are stably sorted
//   by (row, column), so duplicates stay in file order and are kept, not summed);
//   indices are checked against [1, rows] x [1, cols].
// Differences: skew-symmetric files are supported (mirrored with the opposite
// sign; the reference throws), "array" storage and complex/hermitian files are
// rejected with a message, and the row pointer is built with 64-bit counts.
// The writer emits "coordinate real general" with 17 significant digits, which
// round-trips every double exactly.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hpb_internal.hpp"

namespace {

struct entry {
    int64_t r, c;
    double v;
};

std::string lower(std::string s)
{
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

}  // namespace

extern "C" {

int hpb_mm_read_csr(const char* path, int* out_rows, int* out_cols, int64_t* out_nnz, int** out_row_ptr,
                    int** out_col_ind, double** out_val)
{
    HPB_REQUIRE(path && out_rows && out_cols && out_nnz && out_row_ptr && out_col_ind && out_val, HPB_ERR_INVALID,
                "hpb_mm_read_csr: null argument");
    *out_row_ptr = nullptr;
    *out_col_ind = nullptr;
    *out_val = nullptr;
    std::ifstream fin(path);
    HPB_REQUIRE(fin.is_open(), HPB_ERR_INVALID, std::string("hpb_mm_read_csr: cannot open ") + path);
    std::string line;
    HPB_REQUIRE((bool)std::getline(fin, line), HPB_ERR_INVALID, "hpb_mm_read_csr: empty file");
    std::istringstream bs(line);
    std::string tag, obj, storage, type, sym;
    bs >> tag >> obj >> storage >> type >> sym;
    obj = lower(obj);
    storage = lower(storage);
    type = lower(type);
    sym = lower(sym);
    HPB_REQUIRE(tag == "%%MatrixMarket" && obj == "matrix", HPB_ERR_INVALID, "hpb_mm_read_csr: invalid MatrixMarket banner");
    HPB_REQUIRE(storage == "coordinate", HPB_ERR_UNSUPPORTED,
                "hpb_mm_read_csr: unsupported MatrixMarket storage format [" + storage + "]");
    HPB_REQUIRE(type == "real" || type == "integer" || type == "pattern", HPB_ERR_UNSUPPORTED,
                "hpb_mm_read_csr: unsupported MatrixMarket data type [" + type + "]");
    HPB_REQUIRE(sym == "general" || sym == "symmetric" || sym == "skew-symmetric", HPB_ERR_UNSUPPORTED,
                "hpb_mm_read_csr: unsupported MatrixMarket symmetry [" + sym + "]");
    // skip comments
    while (std::getline(fin, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        break;
    }
    int64_t rows = 0, cols = 0, ne = 0;
    {
        std::istringstream ss(line);
        HPB_REQUIRE((bool)(ss >> rows >> cols >> ne), HPB_ERR_INVALID, "hpb_mm_read_csr: error while reading sizes");
    }
    HPB_REQUIRE(rows > 0 && cols > 0 && ne >= 0 && rows < (1ll << 31) && cols < (1ll << 31), HPB_ERR_INVALID,
                "hpb_mm_read_csr: sizes out of range");
    std::vector<entry> e;
    e.reserve((size_t)ne);
    const bool pattern = type == "pattern";
    for (int64_t k = 0; k < ne; ++k) {
        entry x{0, 0, 1.0};
        if (!(fin >> x.r >> x.c)) HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_mm_read_csr: unexpected EOF while reading entries");
        if (!pattern && !(fin >> x.v)) HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_mm_read_csr: unexpected EOF while reading entries");
        HPB_REQUIRE(x.r >= 1 && x.c >= 1, HPB_ERR_INVALID, "hpb_mm_read_csr: found invalid index (index < 1)");
        HPB_REQUIRE(x.r <= rows && x.c <= cols, HPB_ERR_INVALID, "hpb_mm_read_csr: found invalid index (index > size)");
        x.r -= 1;
        x.c -= 1;
        e.push_back(x);
    }
    if (sym != "general") {
        const double sgn = sym == "skew-symmetric" ? -1.0 : 1.0;
        const size_t n0 = e.size();
        for (size_t k = 0; k < n0; ++k)
            if (e[k].r != e[k].c) e.push_back(entry{e[k].c, e[k].r, sgn * e[k].v});
    }
    std::stable_sort(e.begin(), e.end(), [](const entry& a, const entry& b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    const int64_t nnz = (int64_t)e.size();
    HPB_REQUIRE(nnz < (1ll << 31), HPB_ERR_INVALID, "hpb_mm_read_csr: more than 2^31-1 entries (int32 indices)");
    int* rp = (int*)std::calloc((size_t)rows + 1, sizeof(int));
    int* ci = (int*)std::malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    double* va = (double*)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    if (!rp || !ci || !va) {
        std::free(rp);
        std::free(ci);
        std::free(va);
        HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_mm_read_csr: out of host memory");
    }
    std::vector<int64_t> cnt((size_t)rows + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) {
        ci[k] = (int)e[k].c;
        va[k] = e[k].v;
        cnt[e[k].r + 1]++;
    }
    for (int64_t r = 0; r < rows; ++r) cnt[r + 1] += cnt[r];
    for (int64_t r = 0; r <= rows; ++r) rp[r] = (int)cnt[r];
    *out_rows = (int)rows;
    *out_cols = (int)cols;
    *out_nnz = nnz;
    *out_row_ptr = rp;
    *out_col_ind = ci;
    *out_val = va;
    return HPB_OK;
}

void hpb_mm_free(void* host_ptr) { std::free(host_ptr); }

int hpb_mm_write_csr(const char* path, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const double* val)
{
    HPB_REQUIRE(path && row_ptr && (nnz == 0 || (col_ind && val)) && rows > 0 && cols > 0, HPB_ERR_INVALID,
                "hpb_mm_write_csr: bad argument");
    HPB_REQUIRE(row_ptr[0] == 0 && row_ptr[rows] == nnz, HPB_ERR_INVALID, "hpb_mm_write_csr: row_ptr inconsistent with nnz");
    FILE* f = std::fopen(path, "w");
    HPB_REQUIRE(f, HPB_ERR_INVALID, std::string("hpb_mm_write_csr: cannot open ") + path + ": " + std::strerror(errno));
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %lld\n", rows, cols, (long long)nnz);
    for (int r = 0; r < rows; ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) std::fprintf(f, "%d %d %.17g\n", r + 1, col_ind[k] + 1, val[k]);
    const bool ok = std::fclose(f) == 0;
    HPB_REQUIRE(ok, HPB_ERR_INVALID, "hpb_mm_write_csr: write failed");
    return HPB_OK;
}

}  // extern "C"
