// hpb_transpose.hip -- CSR transpose on the device, for the transposed products.
//
// The reference runs its transposed products through cuSPARSE with
// CUSPARSE_OPERATION_TRANSPOSE: csr::gpu_matrix::axpy(..., transpose = true)
// (source/common/csr/gpu_matrix.h:322-333) and the transposed triangular solves
// (gpu_matrix.h:142-154, 232-244).  Here the transpose B = A^T is formed once as a
// CSR matrix of its own and the ordinary row kernels run on it, so the transposed
// product has the same summation order rules, cost and compensated-row option as
// the plain one, and no floating-point atomics.
//
// A setup call: it allocates and frees a temporary of about 4 (cols + 1) bytes and
// synchronises the handle's stream.
// Steps (all on the device):
//  1. count the entries of every column (integer atomics: the counts do not depend
//     on the order of the increments);
//  2. inclusive scan of the counts: row_ptr of B (blocks of 2048 counts scanned in
//     LDS, the block totals scanned the same way, then added back);
//  3. scatter every entry a_ij to row j of B at a cursor taken with an integer
//     atomic -- the order inside a row of B is not yet fixed;
//  4. sort every row of B by (column, value bits): insertion sort for short rows,
//     heap sort for long ones.  The result is one fixed matrix whatever the order
//     of step 3 (for duplicate (i, j) entries the value bits decide).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hpb_internal.hpp"

namespace hpb {
namespace {

constexpr int TT = 256;          // threads per block
constexpr int SCAN_ITEMS = 8;    // counts per thread in the scan
constexpr int SCAN_TILE = TT * SCAN_ITEMS;

__global__ __launch_bounds__(TT) void k_fill_int(size_t n, int v, int* __restrict__ p)
{
    for (size_t i = (size_t)blockIdx.x * TT + threadIdx.x; i < n; i += (size_t)gridDim.x * TT) p[i] = v;
}

// counts of column j land in tp[j + 1]; a column index outside [0, cols) sets *bad
__global__ __launch_bounds__(TT) void k_count_cols(int64_t nnz, int cols, const int* __restrict__ ci,
                                                   int* __restrict__ tp, int* __restrict__ bad)
{
    for (int64_t k = (int64_t)blockIdx.x * TT + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * TT) {
        const int j = ci[k];
        if (j < 0 || j >= cols) {
            *bad = 1;
            continue;
        }
        atomicAdd(&tp[j + 1], 1);
    }
}

// inclusive scan of SCAN_TILE consecutive ints per block; block totals to aux
__global__ __launch_bounds__(TT) void k_scan_tile(size_t n, int* __restrict__ a, int* __restrict__ aux)
{
    __shared__ int wsum[TT / 64];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = base + k < n ? a[base + k] : 0;
        run += v[k];
        v[k] = run;
    }
    // exclusive prefix of the thread totals: wave scan, then the waves in order
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wid; ++w) before += wsum[w];
    const int excl = before + incl - run;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) a[base + k] = v[k] + excl;
    if (threadIdx.x == TT - 1) aux[blockIdx.x] = before + incl;
}

// adds the inclusive prefix of the earlier blocks' totals (aux scanned) to block b
__global__ __launch_bounds__(TT) void k_scan_add(size_t n, int* __restrict__ a, const int* __restrict__ aux)
{
    if (blockIdx.x == 0) return;
    const int add = aux[blockIdx.x - 1];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    for (int k = threadIdx.x; k < SCAN_TILE; k += TT)
        if (base + k < n) a[base + k] += add;
}

__device__ __forceinline__ int row_of(const int* __restrict__ rp, int rows, int64_t k)
{
    int lo = 0, hi = rows;   // last row with rp[row] <= k
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)rp[mid] <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T>
__global__ __launch_bounds__(TT) void k_scatter(int rows, int64_t nnz, const int* __restrict__ rp,
                                                const int* __restrict__ ci, const T* __restrict__ val,
                                                int* __restrict__ cursor, int* __restrict__ tci, T* __restrict__ tval)
{
    for (int64_t k = (int64_t)blockIdx.x * TT + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * TT) {
        const int j = ci[k];
        const int p = atomicAdd(&cursor[j], 1);
        tci[p] = row_of(rp, rows, k);
        tval[p] = val[k];
    }
}

template <typename T> struct key_bits;
template <> struct key_bits<double> { using U = unsigned long long; };
template <> struct key_bits<float> { using U = unsigned; };

template <typename T>
__device__ __forceinline__ bool less_entry(int c1, T v1, int c2, T v2)
{
    using U = typename key_bits<T>::U;
    if (c1 != c2) return c1 < c2;
    return __builtin_bit_cast(U, v1) < __builtin_bit_cast(U, v2);
}

template <typename T>
__device__ void sift_down(int* c, T* v, int root, int len)
{
    for (;;) {
        int child = 2 * root + 1;
        if (child >= len) return;
        if (child + 1 < len && less_entry<T>(c[child], v[child], c[child + 1], v[child + 1])) ++child;
        if (!less_entry<T>(c[root], v[root], c[child], v[child])) return;
        const int tc = c[root];
        c[root] = c[child];
        c[child] = tc;
        const T tv = v[root];
        v[root] = v[child];
        v[child] = tv;
        root = child;
    }
}

template <typename T>
__global__ __launch_bounds__(TT) void k_sort_rows(int trows, const int* __restrict__ tp, int* __restrict__ tci,
                                                  T* __restrict__ tval)
{
    for (int j = blockIdx.x * TT + threadIdx.x; j < trows; j += gridDim.x * TT) {
        const int b = tp[j], len = tp[j + 1] - tp[j];
        int* c = tci + b;
        T* v = tval + b;
        if (len <= 32) {
            for (int a = 1; a < len; ++a) {
                const int kc = c[a];
                const T kv = v[a];
                int p = a - 1;
                while (p >= 0 && less_entry<T>(kc, kv, c[p], v[p])) {
                    c[p + 1] = c[p];
                    v[p + 1] = v[p];
                    --p;
                }
                c[p + 1] = kc;
                v[p + 1] = kv;
            }
        } else {
            for (int r = len / 2 - 1; r >= 0; --r) sift_down<T>(c, v, r, len);
            for (int e = len - 1; e > 0; --e) {
                const int tc = c[0];
                c[0] = c[e];
                c[e] = tc;
                const T tv = v[0];
                v[0] = v[e];
                v[e] = tv;
                sift_down<T>(c, v, 0, e);
            }
        }
    }
}

unsigned grid_for(hpb_handle_t h, int64_t work)
{
    const int64_t cap = (int64_t)(h->num_cu > 0 ? h->num_cu : 256) * 8;
    int64_t g = (work + TT - 1) / TT;
    if (g < 1) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

// inclusive scan of a[0..n) in place; tmp holds at least scan_tmp_ints(n) ints
size_t scan_tmp_ints(size_t n)
{
    size_t total = 0;
    while (n > 1) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        total += n;
    }
    return total > 0 ? total : 1;
}

void scan_inclusive(hipStream_t st, size_t n, int* a, int* tmp)
{
    if (n == 0) return;
    const size_t blocks = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)blocks), dim3(TT), 0, st, n, a, tmp);
    if (blocks > 1) {
        scan_inclusive(st, blocks, tmp, tmp + blocks);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)blocks), dim3(TT), 0, st, n, a, (const int*)tmp);
    }
}

template <typename T>
int transpose_entry(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* rp, const int* ci, const T* val,
                    int* trp, int* tci, T* tval)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && trp, HPB_ERR_INVALID, "hpb_csr_transpose: null argument");
    HPB_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && nnz < (1ll << 31), HPB_ERR_INVALID,
                "hpb_csr_transpose: bad size (int32 row pointers need nnz < 2^31)");
    HPB_REQUIRE(nnz == 0 || (rp && ci && val && tci && tval), HPB_ERR_INVALID, "hpb_csr_transpose: null argument");
    hipStream_t st = h->stream;
    // temporaries: the cursors (cols ints), the scan's block totals, the error flag
    const size_t ntmp = (size_t)cols + scan_tmp_ints((size_t)cols + 1) + 1;
    int* tmp = nullptr;
    HPB_HIP(hipMalloc((void**)&tmp, ntmp * sizeof(int)));
    int* cursor = tmp;
    int* bad = tmp + cols;
    int* scan_tmp = bad + 1;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(h, cols + 1)), dim3(TT), 0, st, (size_t)cols + 1, 0, trp);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(TT), 0, st, (size_t)1, 0, bad);
    if (nnz > 0)
        hipLaunchKernelGGL(k_count_cols, dim3(grid_for(h, nnz)), dim3(TT), 0, st, nnz, cols, ci, trp, bad);
    scan_inclusive(st, (size_t)cols + 1, trp, scan_tmp);
    int hbad = 0;
    hipError_t e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess || hbad) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_status(e, __FILE__, __LINE__, "hpb_csr_transpose");
        HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_csr_transpose: column index out of range");
    }
    if (nnz > 0) {
        HPB_HIP(hipMemcpyAsync(cursor, trp, sizeof(int) * (size_t)cols, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL((k_scatter<T>), dim3(grid_for(h, nnz)), dim3(TT), 0, st, rows, nnz, rp, ci, val, cursor,
                           tci, tval);
        hipLaunchKernelGGL((k_sort_rows<T>), dim3(grid_for(h, cols)), dim3(TT), 0, st, cols, (const int*)trp, tci,
                           tval);
    }
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    if (le != hipSuccess) return hip_status(le, __FILE__, __LINE__, "hpb_csr_transpose: kernel launch");
    HPB_HIP(se);
    return HPB_OK;
}

}  // namespace
}  // namespace hpb

extern "C" {

int hpb_csr_transpose_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const double* val, int* t_row_ptr, int* t_col_ind, double* t_val)
{
    return hpb::transpose_entry<double>(h, rows, cols, nnz, row_ptr, col_ind, val, t_row_ptr, t_col_ind, t_val);
}
int hpb_csr_transpose_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const float* val, int* t_row_ptr, int* t_col_ind, float* t_val)
{
    return hpb::transpose_entry<float>(h, rows, cols, nnz, row_ptr, col_ind, val, t_row_ptr, t_col_ind, t_val);
}

}  // extern "C"
