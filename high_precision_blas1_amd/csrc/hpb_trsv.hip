// hpb_trsv.hip -- sparse triangular solves and ILU(0) for CSR matrices.
//
// Replaces csr::gpu_matrix::triangular_solve_lower / _upper
// (source/common/csr/gpu_matrix.h:142-320), which run cuSPARSE csrsv2 with a
// buffer-size query, an analysis, a cudaMalloc and the solve on every call:
//   lower: L x = alpha b with the strictly lower triangle and a unit diagonal
//          (CUSPARSE_DIAG_TYPE_UNIT: the stored diagonal is ignored);
//   upper: U x = alpha b with the upper triangle including the stored diagonal.
// The transposed solves of the reference (L^T x = alpha b, U^T x = alpha b) are the
// upper solve with a unit diagonal and the lower solve with the stored diagonal on
// the transpose of the matrix (hpb_csr_transpose_*); the analysis therefore also
// takes those two kinds (uplo 2: lower, stored diagonal; 3: upper, unit diagonal).
// Both read only their triangle, so one CSR holding an incomplete LU factorisation
// (L strictly below, U on and above the diagonal) serves both solves, as in the
// reference's csr::preconditioner (common/csr/preconditioner.h:61-75).
//
// Design: the analysis is done once per pattern and kept in an hpb_trsv_t: the
// level of every row (longest dependency chain, a host pass over row_ptr/col_ind
// in dependency order), the rows grouped by level, and the diagonal positions.
// A solve launches one kernel per level; the rows of a level are independent.
// LPR lanes per row (as in the SpMV) accumulate alpha b_i - sum a_ij x_j, plain
// (fma) or compensated (Dot2 over the row), merged by an xor butterfly.
// x may alias b: each row reads its own b_i before writing x_i.
//
// ILU(0) (new: the reference's gpu_ilu0_preconditioner::apply only copies its
// argument, common/csr/gpu_ilu0_preconditioner.h:62-72): the factors keep the
// pattern of A and satisfy (L U)_ij = a_ij on it.  Row i is finished level by level
// with the lower analysis (it depends on the rows k < i it references): for
// k in row i in increasing order, l_ik = a_ik / u_kk, then a_ij -= l_ik u_kj for
// every j > k present in both rows; one 64-lane block per row, lanes over j.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

struct hpb_trsv_s {
    int uplo = 0;        // 0 lower (unit diagonal), 1 upper, 2 lower (stored diagonal), 3 upper (unit diagonal)
    bool upper = false;
    bool unit = true;
    int rows = 0;
    int64_t nnz = 0;
    const int* rp = nullptr;    // the caller's device arrays (not owned)
    const int* ci = nullptr;
    int* perm = nullptr;        // device: rows ordered by level
    int* diag = nullptr;        // device: position of the diagonal entry of each row, -1 if absent
    std::vector<int> level_ptr; // host: level l holds perm[level_ptr[l] .. level_ptr[l+1])
    int lpr = 1;
    bool all_diag = false;      // every row stores its diagonal entry
};

namespace hpb {
namespace {

constexpr int TB = 256;

template <typename T, int LPR, bool COMP, bool UPPER, bool UNIT>
__global__ __launch_bounds__(TB) void k_trsv(int beg, int end, const int* __restrict__ perm,
                                             const int* __restrict__ rp, const int* __restrict__ ci,
                                             const T* __restrict__ val, const int* __restrict__ diag, T alpha,
                                             const T* b, T* x)
{
    const int lane = threadIdx.x % LPR;
    const int t = beg + (int)((blockIdx.x * TB + threadIdx.x) / LPR);
    if (t >= end) return;
    const int i = perm[t];
    const int rb = rp[i], re = rp[i + 1];
    T s = T(0), c = T(0);
    if (lane == 0) s = alpha == T(1) ? b[i] : alpha * b[i];
    for (int k = rb + lane; k < re; k += LPR) {
        const int j = ci[k];
        if (UPPER ? (j > i) : (j < i)) {
            if (COMP) dev::dot2_step<T>(s, c, -val[k], x[j]);
            else s = dev::fma_rn<T>(-val[k], x[j], s);
        }
    }
#pragma unroll
    for (int m = LPR / 2; m >= 1; m >>= 1) {
        const T s2 = __shfl_xor(s, m, 64);
        if (COMP) {
            const T c2 = __shfl_xor(c, m, 64);
            dev::merge<T>(s, c, s2, c2);
        } else {
            s = s + s2;
        }
    }
    if (lane == 0) {
        T v = COMP ? s + c : s;
        if (!UNIT) v = v / val[diag[i]];
        x[i] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_ilu0_rows(int beg, const int* __restrict__ perm, const int* __restrict__ rp,
                                                  const int* __restrict__ ci, const int* __restrict__ diag,
                                                  T* __restrict__ val)
{
    const int i = perm[beg + blockIdx.x];
    const int rb = rp[i], re = rp[i + 1], di = diag[i];
    for (int kk = rb; kk < di; ++kk) {
        const int k = ci[kk];
        const T lik = val[kk] / val[diag[k]];
        __syncthreads();   // every lane has read val[kk] before it changes
        if (threadIdx.x == 0) val[kk] = lik;
        // a_ij -= l_ik u_kj for j > k in row i that also appear in row k (upper part)
        const int kb = diag[k] + 1, ke = rp[k + 1];
        for (int jj = kk + 1 + (int)threadIdx.x; jj < re; jj += 64) {
            const int j = ci[jj];
            int lo = kb, hi = ke;   // binary search of column j in row k's upper part
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (lo < ke && ci[lo] == j) val[jj] = dev::fma_rn<T>(-lik, val[lo], val[jj]);
        }
        __syncthreads();   // row i updated before the next k reads it
    }
}

template <typename T, bool COMP, bool UPPER, bool UNIT, int LPR>
void trsv_level(hipStream_t st, const hpb_trsv_s* t, int beg, int end, const T* val, T alpha, const T* b, T* x)
{
    const int rows = end - beg;
    const int per_block = TB / LPR;
    const unsigned grid = (unsigned)((rows + per_block - 1) / per_block);
    hipLaunchKernelGGL((k_trsv<T, LPR, COMP, UPPER, UNIT>), dim3(grid), dim3(TB), 0, st, beg, end, t->perm, t->rp,
                       t->ci, val, t->diag, alpha, b, x);
}

template <typename T, bool COMP, bool UPPER, bool UNIT>
void trsv_all(hipStream_t st, const hpb_trsv_s* t, const T* val, T alpha, const T* b, T* x)
{
    const int nl = (int)t->level_ptr.size() - 1;
    for (int l = 0; l < nl; ++l) {
        const int beg = t->level_ptr[l], end = t->level_ptr[l + 1];
        switch (t->lpr) {
        case 1: trsv_level<T, COMP, UPPER, UNIT, 1>(st, t, beg, end, val, alpha, b, x); break;
        case 2: trsv_level<T, COMP, UPPER, UNIT, 2>(st, t, beg, end, val, alpha, b, x); break;
        case 4: trsv_level<T, COMP, UPPER, UNIT, 4>(st, t, beg, end, val, alpha, b, x); break;
        case 8: trsv_level<T, COMP, UPPER, UNIT, 8>(st, t, beg, end, val, alpha, b, x); break;
        default: trsv_level<T, COMP, UPPER, UNIT, 16>(st, t, beg, end, val, alpha, b, x); break;
        }
    }
}

template <typename T>
int trsv_entry(hpb_handle_t h, hpb_trsv_t t, const T* val, T alpha, const T* b, T* x, int comp)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && t && val && b && x, HPB_ERR_INVALID, "hpb_csr_trsv: null argument");
    hipStream_t st = h->stream;
    switch (t->uplo) {
    case 0:
        if (comp) trsv_all<T, true, false, true>(st, t, val, alpha, b, x);
        else trsv_all<T, false, false, true>(st, t, val, alpha, b, x);
        break;
    case 1:
        if (comp) trsv_all<T, true, true, false>(st, t, val, alpha, b, x);
        else trsv_all<T, false, true, false>(st, t, val, alpha, b, x);
        break;
    case 2:
        if (comp) trsv_all<T, true, false, false>(st, t, val, alpha, b, x);
        else trsv_all<T, false, false, false>(st, t, val, alpha, b, x);
        break;
    default:
        if (comp) trsv_all<T, true, true, true>(st, t, val, alpha, b, x);
        else trsv_all<T, false, true, true>(st, t, val, alpha, b, x);
        break;
    }
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename T>
int ilu0_entry(hpb_handle_t h, hpb_trsv_t t, T* val)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && t && val, HPB_ERR_INVALID, "hpb_csr_ilu0: null argument");
    HPB_REQUIRE(t->uplo == 0, HPB_ERR_INVALID, "hpb_csr_ilu0: needs the analysis of the lower triangle");
    HPB_REQUIRE(t->all_diag, HPB_ERR_INVALID, "hpb_csr_ilu0: every row needs a stored diagonal entry");
    const int nl = (int)t->level_ptr.size() - 1;
    for (int l = 0; l < nl; ++l) {
        const int beg = t->level_ptr[l], end = t->level_ptr[l + 1];
        hipLaunchKernelGGL((k_ilu0_rows<T>), dim3((unsigned)(end - beg)), dim3(64), 0, h->stream, beg, t->perm, t->rp,
                           t->ci, t->diag, val);
    }
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

}  // namespace
}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_csr_trsv_analysis(hpb_handle_t h, int uplo, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                          hpb_trsv_t* out)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && row_ptr && col_ind && out, HPB_ERR_INVALID, "hpb_csr_trsv_analysis: null argument");
    HPB_REQUIRE(uplo >= 0 && uplo <= 3, HPB_ERR_INVALID, "hpb_csr_trsv_analysis: uplo must be 0, 1, 2 or 3");
    const bool lower = (uplo == 0 || uplo == 2), unit = (uplo == 0 || uplo == 3);
    HPB_REQUIRE(rows > 0 && nnz > 0 && nnz < (1ll << 31), HPB_ERR_INVALID, "hpb_csr_trsv_analysis: bad size");
    *out = nullptr;
    // the pattern on the host, through the handle's pinned staging buffer (a pageable
    // copy of the 128^3 pattern took 15-18 ms)
    const size_t need = sizeof(int) * ((size_t)rows + 1 + (size_t)nnz);
    if (h->hstage_bytes < need) {
        if (h->hstage) (void)hipHostFree(h->hstage);
        h->hstage = nullptr;
        h->hstage_bytes = 0;
        HPB_HIP(hipHostMalloc(&h->hstage, need, hipHostMallocDefault));
        h->hstage_bytes = need;
    }
    int* const rp = static_cast<int*>(h->hstage);
    int* const ci = rp + (size_t)rows + 1;
    HPB_HIP(hipMemcpyAsync(rp, row_ptr, sizeof(int) * ((size_t)rows + 1), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipMemcpyAsync(ci, col_ind, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    HPB_REQUIRE(rp[0] == 0 && rp[rows] == nnz, HPB_ERR_INVALID, "hpb_csr_trsv_analysis: row_ptr inconsistent with nnz");
    // checks, diagonal positions and the end of each row's lower part (the columns
    // are sorted, so a row's j < i come first), in parallel over row ranges; the
    // first failing row in dependency order names the error, as a sequential scan
    std::vector<int> level((size_t)rows, 0), diag((size_t)rows, -1), lend((size_t)rows);
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned T = (unsigned)std::max<int64_t>(1, std::min<int64_t>({(int64_t)hw, 16, (int64_t)rows / 65536 + 1}));
        std::vector<int> first_code(T, 0);
        auto scan = [&](unsigned t) {
            const int s0 = (int)((int64_t)rows * t / T), s1 = (int)((int64_t)rows * (t + 1) / T);
            for (int s = s0; s < s1; ++s) {
                const int i = lower ? s : rows - 1 - s;   // dependency order
                int code = 0, le = rp[i + 1];
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    const int j = ci[k];
                    if (!(j >= 0 && j < rows)) { code = 1; break; }
                    if (!(k == rp[i] || ci[k - 1] < j)) { code = 2; break; }
                    if (j >= i && le == rp[i + 1]) le = k;
                    if (j == i) diag[i] = k;
                }
                lend[i] = le;
                if (!code && !unit && diag[i] < 0) code = 3;
                if (code) {   // later rows of this range cannot come first
                    first_code[t] = code;
                    return;
                }
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < T; ++t) th.emplace_back(scan, t);
        scan(0);
        for (auto& x : th) x.join();
        for (unsigned t = 0; t < T; ++t) {   // ranges in dependency order: the first error wins
            HPB_REQUIRE(first_code[t] != 1, HPB_ERR_INVALID, "hpb_csr_trsv_analysis: column index out of range");
            HPB_REQUIRE(first_code[t] != 2, HPB_ERR_INVALID,
                        "hpb_csr_trsv_analysis: columns must be sorted and unique within a row");
            HPB_REQUIRE(first_code[t] != 3, HPB_ERR_INVALID,
                        "hpb_csr_trsv_analysis: a row without a stored diagonal entry (structurally singular)");
        }
    }
    // levels (longest dependency chain): one pass in dependency order over the
    // dependencies only
    int nlev = 0;
    for (int s = 0; s < rows; ++s) {
        const int i = lower ? s : rows - 1 - s;
        int lv = 0;
        if (lower) {
            for (int k = rp[i]; k < lend[i]; ++k) lv = std::max(lv, level[ci[k]] + 1);
        } else {
            for (int k = lend[i] + (diag[i] >= 0 ? 1 : 0); k < rp[i + 1]; ++k) lv = std::max(lv, level[ci[k]] + 1);
        }
        level[i] = lv;
        nlev = std::max(nlev, lv + 1);
    }
    hpb_trsv_s* t = new hpb_trsv_s();
    t->uplo = uplo;
    t->upper = !lower;
    t->unit = unit;
    t->rows = rows;
    t->nnz = nnz;
    t->rp = row_ptr;
    t->ci = col_ind;
    t->level_ptr.assign((size_t)nlev + 1, 0);
    for (int i = 0; i < rows; ++i) t->level_ptr[level[i] + 1]++;
    for (int l = 0; l < nlev; ++l) t->level_ptr[l + 1] += t->level_ptr[l];
    std::vector<int> perm((size_t)rows), fill(t->level_ptr.begin(), t->level_ptr.end() - 1);
    for (int i = 0; i < rows; ++i) perm[fill[level[i]]++] = i;   // ascending row order within a level
    t->all_diag = std::all_of(diag.begin(), diag.end(), [](int d) { return d >= 0; });
    const double mean = (double)nnz / rows;
    t->lpr = mean <= 1.5 ? 1 : mean <= 3.0 ? 2 : mean <= 6.0 ? 4 : mean <= 12.0 ? 8 : 16;
    hipError_t e = hipMalloc((void**)&t->perm, sizeof(int) * rows);
    if (e == hipSuccess) e = hipMalloc((void**)&t->diag, sizeof(int) * rows);
    if (e == hipSuccess) e = hipMemcpyAsync(t->perm, perm.data(), sizeof(int) * rows, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t->diag, diag.data(), sizeof(int) * rows, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        if (t->perm) (void)hipFree(t->perm);
        if (t->diag) (void)hipFree(t->diag);
        delete t;
        return hip_status(e, __FILE__, __LINE__, "hpb_csr_trsv_analysis");
    }
    *out = t;
    return HPB_OK;
}

int hpb_csr_trsv_levels(hpb_trsv_t t, int* nlevels)
{
    HPB_REQUIRE(t && nlevels, HPB_ERR_INVALID, "hpb_csr_trsv_levels: null argument");
    *nlevels = (int)t->level_ptr.size() - 1;
    return HPB_OK;
}

int hpb_trsv_destroy(hpb_trsv_t t)
{
    if (t) {
        if (t->perm) (void)hipFree(t->perm);
        if (t->diag) (void)hipFree(t->diag);
        delete t;
    }
    return HPB_OK;
}

int hpb_csr_trsv_f64(hpb_handle_t h, hpb_trsv_t t, const double* val, double alpha, const double* b, double* x,
                     int compensated)
{
    return trsv_entry<double>(h, t, val, alpha, b, x, compensated);
}
int hpb_csr_trsv_f32(hpb_handle_t h, hpb_trsv_t t, const float* val, float alpha, const float* b, float* x,
                     int compensated)
{
    return trsv_entry<float>(h, t, val, alpha, b, x, compensated);
}

int hpb_csr_ilu0_f64(hpb_handle_t h, hpb_trsv_t lower, double* val)
{
    return ilu0_entry<double>(h, lower, val);
}
int hpb_csr_ilu0_f32(hpb_handle_t h, hpb_trsv_t lower, float* val)
{
    return ilu0_entry<float>(h, lower, val);
}

}  // extern "C"
