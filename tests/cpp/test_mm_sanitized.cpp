// MatrixMarket reader / writer of the library (host code) under AddressSanitizer and
// UndefinedBehaviorSanitizer: reads and writes well-formed files of every kind the
// reader accepts and every malformed file it must reject, so that buffer overruns,
// leaks, use-after-free or undefined arithmetic in the parser show up as a
// sanitizer report (and a non-zero exit) rather than as a wrong matrix.
// usage: test_mm_sanitized <dir>; prints one line per case: name status rows nnz
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "hpb.h"

static int read_case(const std::string& dir, const char* name, const char* text)
{
    const std::string path = dir + "/" + name + ".mtx";
    {
        std::ofstream f(path);
        f << text;
    }
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    int *rp = nullptr, *ci = nullptr;
    double* va = nullptr;
    const int st = hpb_mm_read_csr(path.c_str(), &rows, &cols, &nnz, &rp, &ci, &va);
    double sum = 0.0;
    if (st == 0) {
        for (int64_t k = 0; k < nnz; ++k) sum += va[k] * (double)(ci[k] + 1);
        sum += (double)rp[rows];
        // write it back and read the copy
        const std::string out = dir + "/" + name + "_copy.mtx";
        if (hpb_mm_write_csr(out.c_str(), rows, cols, nnz, rp, ci, va) != 0) return 3;
        int r2 = 0, c2 = 0;
        int64_t n2 = 0;
        int *rp2 = nullptr, *ci2 = nullptr;
        double* va2 = nullptr;
        if (hpb_mm_read_csr(out.c_str(), &r2, &c2, &n2, &rp2, &ci2, &va2) != 0) return 4;
        for (int64_t k = 0; k < n2; ++k)
            if (va2[k] != va[k] || ci2[k] != ci[k]) return 5;
        hpb_mm_free(rp2);
        hpb_mm_free(ci2);
        hpb_mm_free(va2);
        hpb_mm_free(rp);
        hpb_mm_free(ci);
        hpb_mm_free(va);
    }
    std::printf("%s %d %d %lld %.17g\n", name, st, rows, (long long)nnz, sum);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    struct tcase {
        const char* name;
        const char* text;
    };
    const tcase cases[] = {
        {"general", "%%MatrixMarket matrix coordinate real general\n% c\n3 4 5\n1 1 1.5\n3 4 -2\n2 2 0.25\n1 3 7\n3 1 1e-300\n"},
        {"symmetric", "%%MatrixMarket matrix coordinate real symmetric\n4 4 4\n1 1 2\n2 1 -1\n4 3 3\n4 4 5\n"},
        {"skew", "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1\n3 2 -4\n"},
        {"integer", "%%MatrixMarket matrix coordinate integer general\n2 2 3\n1 1 3\n2 1 -7\n2 2 11\n"},
        {"pattern", "%%MatrixMarket matrix coordinate pattern general\n3 3 3\n1 2\n2 3\n3 1\n"},
        {"duplicates", "%%MatrixMarket matrix coordinate real general\n2 2 4\n1 1 1\n1 1 2\n2 2 3\n1 1 4\n"},
        {"empty", "%%MatrixMarket matrix coordinate real general\n5 5 0\n"},
        {"bad_banner", "%%MatrixMarket tensor coordinate real general\n2 2 1\n1 1 1\n"},
        {"bad_index", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n"},
        {"zero_index", "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n"},
        {"short", "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n"},
        {"array", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n"},
        {"complex", "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n"},
        {"garbage_value", "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 abc\n"},
        {"no_size", "%%MatrixMarket matrix coordinate real general\n"},
        {"huge_size", "%%MatrixMarket matrix coordinate real general\n99999999999 2 1\n1 1 1\n"},
    };
    for (const tcase& c : cases)
        if (int rc = read_case(dir, c.name, c.text)) return rc;
    // files of several MB, parsed by threads in chunks (big), and the same with one
    // entry split over two lines, which sends them through the token parser (big_split)
    {
        const int n = 150000;
        std::string big = "%%MatrixMarket matrix coordinate real general\n150000 150000 600000\n", split;
        char line[96];
        for (int i = 1; i <= n; ++i)
            for (int q = 0; q < 4; ++q) {
                const int j = (int)(((long long)i * (2 * q + 7)) % n) + 1;
                std::snprintf(line, sizeof line, "%d %d %.17g\n", i, j, (q - 1.5) * 1.0e-3 * i);
                if (i == n / 2 && q == 1) {   // "i j\nv\n" in the copy
                    std::string l(line);
                    const size_t sp = l.rfind(' ');
                    split += l.substr(0, sp) + "\n" + l.substr(sp + 1);
                } else {
                    split += line;
                }
                big += line;
            }
        split = big.substr(0, big.find('\n', big.find('\n') + 1) + 1) + split;
        if (int rc = read_case(dir, "big", big.c_str())) return rc;
        if (int rc = read_case(dir, "big_split", split.c_str())) return rc;
    }
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    int *rp = nullptr, *ci = nullptr;
    double* va = nullptr;
    std::printf("missing %d\n", hpb_mm_read_csr((dir + "/does_not_exist.mtx").c_str(), &rows, &cols, &nnz, &rp, &ci,
                                                &va));
    return 0;
}
