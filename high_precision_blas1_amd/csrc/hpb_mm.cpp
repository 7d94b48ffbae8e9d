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
// Host threads: the reader parses the data lines in chunks split at line ends and
// joins them in file order (any line it does not recognise sends the whole data
// section through the token parser instead); the writer formats row ranges in
// parallel and writes them in order.  Files read and written are the same as with
// one thread.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
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

// Host threads for parsing and formatting (at most 16, one below 2 MB of text).
unsigned text_threads(size_t bytes)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t by_size = bytes / (2u << 20);
    return (unsigned)std::max<size_t>(1, std::min<size_t>({(size_t)hw, (size_t)16, by_size}));
}

bool blank(char ch) { return ch == ' ' || ch == '\t' || ch == '\r'; }

// Parses the data lines in [b, e): one entry "i j [v]" per line, blank lines skipped.
// Returns false at the first line of another form (a split entry, a '+' sign, a word,
// a value out of range, a comment), for which the caller parses the whole data
// section as a token stream instead, so that every file reads as the token parser
// reads it.
bool parse_lines(const char* b, const char* e, bool pattern, std::vector<entry>& out)
{
    const char* p = b;
    while (p < e) {
        while (p < e && blank(*p)) ++p;
        if (p < e && *p == '\n') {
            ++p;
            continue;
        }
        if (p >= e) break;
        entry x{0, 0, 1.0};
        auto num = [&](auto& v) {
            if (p >= e || !((*p >= '0' && *p <= '9') || *p == '.' || *p == '-')) return false;
            const auto r = std::from_chars(p, e, v);
            if (r.ec != std::errc()) return false;
            p = r.ptr;
            return p >= e || blank(*p) || *p == '\n';
        };
        auto sep = [&]() {
            const char* q = p;
            while (p < e && blank(*p)) ++p;
            return p > q && p < e && *p != '\n';
        };
        if (!num(x.r) || !sep() || !num(x.c)) return false;
        if (!pattern && (!sep() || !num(x.v))) return false;
        while (p < e && blank(*p)) ++p;
        if (p < e && *p != '\n') return false;
        out.push_back(x);
    }
    return true;
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
    // the whole file in memory: the header is read line by line, the data lines in
    // parallel chunks
    std::string text;
    {
        FILE* fin = std::fopen(path, "rb");
        HPB_REQUIRE(fin, HPB_ERR_INVALID, std::string("hpb_mm_read_csr: cannot open ") + path);
        std::fseek(fin, 0, SEEK_END);
        const long sz = std::ftell(fin);
        std::fseek(fin, 0, SEEK_SET);
        text.resize(sz > 0 ? (size_t)sz : 0);
        const size_t got = sz > 0 ? std::fread(&text[0], 1, (size_t)sz, fin) : 0;
        std::fclose(fin);
        HPB_REQUIRE(got == text.size(), HPB_ERR_INVALID, std::string("hpb_mm_read_csr: cannot read ") + path);
    }
    size_t pos = 0;
    auto next_line = [&](std::string& line) {
        if (pos >= text.size()) return false;
        size_t nl = text.find('\n', pos);
        if (nl == std::string::npos) nl = text.size();
        line.assign(text, pos, nl - pos);
        pos = nl < text.size() ? nl + 1 : nl;
        return true;
    };
    std::string line;
    HPB_REQUIRE(next_line(line), HPB_ERR_INVALID, "hpb_mm_read_csr: empty file");
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
    while (next_line(line)) {
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
    const bool pattern = type == "pattern";
    // data lines: T chunks split at line ends, parsed in parallel and joined in file
    // order; the first ne entries are the matrix, as for the token parser
    std::vector<entry> e;
    bool lines_ok = true;
    {
        const char* base = text.data();
        const size_t n0 = pos, n1 = text.size();
        const unsigned T = text_threads(n1 - n0);
        std::vector<size_t> cut(T + 1, n1);
        cut[0] = n0;
        for (unsigned t = 1; t < T; ++t) {
            size_t c = n0 + (n1 - n0) * t / T;
            c = std::max(c, cut[t - 1]);
            const size_t nl = text.find('\n', c);
            cut[t] = nl == std::string::npos ? n1 : nl + 1;
        }
        std::vector<std::vector<entry>> part(T);
        std::vector<char> ok(T, 1);
        auto work = [&](unsigned t) {
            part[t].reserve((size_t)(ne / T + 16));
            ok[t] = parse_lines(base + cut[t], base + cut[t + 1], pattern, part[t]) ? 1 : 0;
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < T; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
        size_t tot = 0;
        for (unsigned t = 0; t < T; ++t) {
            lines_ok = lines_ok && ok[t];
            tot += part[t].size();
        }
        if (lines_ok && tot >= (size_t)ne) {
            e.reserve((size_t)ne);
            for (unsigned t = 0; t < T && e.size() < (size_t)ne; ++t) {
                const size_t take = std::min(part[t].size(), (size_t)ne - e.size());
                e.insert(e.end(), part[t].begin(), part[t].begin() + (ptrdiff_t)take);
            }
        } else {
            lines_ok = false;
        }
    }
    auto check = [&](entry& x) {   // 1-based indices in range, made 0-based
        HPB_REQUIRE(x.r >= 1 && x.c >= 1, HPB_ERR_INVALID, "hpb_mm_read_csr: found invalid index (index < 1)");
        HPB_REQUIRE(x.r <= rows && x.c <= cols, HPB_ERR_INVALID, "hpb_mm_read_csr: found invalid index (index > size)");
        x.r -= 1;
        x.c -= 1;
        return HPB_OK;
    };
    if (lines_ok) {
        for (entry& x : e) {   // in file order: the first bad entry names the error
            const int rc = check(x);
            if (rc) return rc;
        }
    } else {   // the token parser over the data section, each entry checked as read
        e.clear();
        e.reserve((size_t)ne);
        std::istringstream din(text.substr(pos));
        for (int64_t k = 0; k < ne; ++k) {
            entry x{0, 0, 1.0};
            if (!(din >> x.r >> x.c)) HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_mm_read_csr: unexpected EOF while reading entries");
            if (!pattern && !(din >> x.v)) HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_mm_read_csr: unexpected EOF while reading entries");
            const int rc = check(x);
            if (rc) return rc;
            e.push_back(x);
        }
    }
    if (sym != "general") {
        const double sgn = sym == "skew-symmetric" ? -1.0 : 1.0;
        const size_t n0 = e.size();
        for (size_t k = 0; k < n0; ++k)
            if (e[k].r != e[k].c) e.push_back(entry{e[k].c, e[k].r, sgn * e[k].v});
    }
    // stable order by (row, column): a stable counting sort by row, then each row
    // stably by column (duplicates keep their file order); files already in that
    // order (as the writer's) are left as they are
    const int64_t nnz = (int64_t)e.size();
    bool sorted = true;
    for (int64_t k = 1; k < nnz && sorted; ++k)
        sorted = e[k - 1].r < e[k].r || (e[k - 1].r == e[k].r && e[k - 1].c <= e[k].c);
    if (!sorted) {
        std::vector<int64_t> start((size_t)rows + 1, 0);
        for (const entry& x : e) start[(size_t)x.r + 1]++;
        for (int64_t r = 0; r < rows; ++r) start[r + 1] += start[r];
        std::vector<entry> s2(e.size());
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (const entry& x : e) s2[(size_t)fill[(size_t)x.r]++] = x;
        for (int64_t r = 0; r < rows; ++r) {
            entry* a = s2.data() + start[r];
            const int64_t len = start[r + 1] - start[r];
            if (len > 32) {
                std::stable_sort(a, a + len, [](const entry& u, const entry& v) { return u.c < v.c; });
                continue;
            }
            for (int64_t i = 1; i < len; ++i) {   // insertion sort: stable, short rows
                const entry x = a[i];
                int64_t j = i;
                while (j > 0 && a[j - 1].c > x.c) {
                    a[j] = a[j - 1];
                    --j;
                }
                a[j] = x;
            }
        }
        e.swap(s2);
    }
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
    // the entry lines formatted in T row ranges of about equal entry counts, in
    // parallel, and written in order
    const unsigned T = text_threads((size_t)nnz * 32);
    std::vector<int> rcut(T + 1, rows);
    rcut[0] = 0;
    for (unsigned t = 1; t < T; ++t) {
        const int64_t want = nnz * t / T;
        rcut[t] = (int)(std::lower_bound(row_ptr, row_ptr + rows + 1, (int)want) - row_ptr);
        rcut[t] = std::min(std::max(rcut[t], rcut[t - 1]), rows);
    }
    std::vector<std::string> out(T);
    auto work = [&](unsigned t) {
        std::string& o = out[t];
        o.reserve((size_t)(row_ptr[rcut[t + 1]] - row_ptr[rcut[t]]) * 32);
        char buf[96];
        for (int r = rcut[t]; r < rcut[t + 1]; ++r)
            for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                const int len = std::snprintf(buf, sizeof buf, "%d %d %.17g\n", r + 1, col_ind[k] + 1, val[k]);
                o.append(buf, (size_t)len);
            }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    bool wrote = true;
    for (unsigned t = 0; t < T; ++t) wrote = wrote && std::fwrite(out[t].data(), 1, out[t].size(), f) == out[t].size();
    const bool ok = (std::fclose(f) == 0) && wrote;
    HPB_REQUIRE(ok, HPB_ERR_INVALID, "hpb_mm_write_csr: write failed");
    return HPB_OK;
}

}  // extern "C"
