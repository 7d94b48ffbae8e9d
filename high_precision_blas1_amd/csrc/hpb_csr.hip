// hpb_csr.hip -- CSR sparse matrix-vector products and structured test matrices.
//
// Replaces csr::gpu_matrix::axpy (source/common/csr/gpu_matrix.h:322-375), which
// calls cusparseSpMV with a cudaMalloc/cudaFree and two descriptor create/destroy
// per call.  Storage as in source/common/csr/matrix.h:91-158: zero-based int32
// row_ptr[rows+1], col_ind[nnz], values[nnz].
//
// Kernel: LPR lanes per row (LPR in {1, 2, 4, 8, 16, 32, 64}, chosen from the
// mean row length), lane l of a row takes entries beg+l, beg+l+LPR, ...; the row
// total is merged across the LPR lanes with an xor butterfly.  The plain kernel
// accumulates s = fma(a_ij, x_j, s); the compensated-row kernel runs an
// Ogita-Rump-Oishi Dot2 over the row (per-lane TwoProd/TwoSum, TwoSum merges) so
// that every y_i is accurate to ~u|row.x| + gamma^2 |row|.|x|.
// The residual kernel forms r = b - A x in the same way with b_i entered as the
// first term of the compensated sum (used by iterative refinement).
// Rows of up to 4 lanes are formed from staged tiles (csr_stage_tile) whose columns
// and values are loaded nontemporally; y is then stored nontemporally too (NT).
#include <hip/hip_runtime.h>

#include <vector>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"
#include "hpb_philox.hpp"

namespace hpb {

constexpr int SPMV_BLOCK = 256;

// Staged tiles (csr_stage_tile) for up to 4 lanes per row: RT rows, CAP entries
// (mean row length up to 2 for one lane, 8 for two or four); longer rows read their
// entries directly.
// FP32 tiles hold twice the rows and entries (the same 16 KB of LDS): a tile's fixed
// latency then moves as many bytes as an FP64 tile.  FP32 at 256^3 / 4096^2:
// 4850 / 5090 GB/s against 4440 / 4210 with FP64-sized tiles; doubling the FP64
// tiles instead halves the blocks per CU (4870 against 5420 GB/s at 256^3).
template <int LPR, typename T = double> struct spmv_tile {
    static constexpr int F = sizeof(T) == 4 ? 2 : 1;
    static constexpr int RT = (LPR == 1 ? 256 : 128) * F, CAP = (LPR == 1 ? 512 : 1024) * F;
};
template <int LPR> constexpr bool spmv_staged() { return LPR <= 4; }

template <typename T, int LPR, bool COMP, int MODE, bool NT>
__global__ __launch_bounds__(SPMV_BLOCK) void k_spmv(int rows, const int* __restrict__ rp,
                                                     const int* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ x, T alpha, T beta,
                                                     const T* __restrict__ b, T* __restrict__ y,
                                                     const int* __restrict__ skip, const T* __restrict__ dalpha)
{
    // MODE 0: y = alpha*A*x + beta*y (beta == 0: y not read); MODE 1: y = b - A*x
    if (skip != nullptr && *skip) return;   // solver graphs: flag set, launch does nothing
    if (dalpha != nullptr) alpha = *dalpha;  // solver graphs: scale factor from device memory
    const int lane = threadIdx.x % LPR;
    auto finish = [&](long r, T s, T c) {
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1) {
            T s2 = __shfl_xor(s, m, 64);
            if (COMP) {
                T c2 = __shfl_xor(c, m, 64);
                dev::merge<T>(s, c, s2, c2);
            } else {
                s = s + s2;
            }
        }
        if (lane == 0) {
            const T v = COMP ? s + c : s;
            T out;
            if (MODE == 1) out = v;
            else if (beta == T(0)) out = (alpha == T(1)) ? v : alpha * v;
            else out = dev::fma_rn<T>(alpha, v, beta * y[r]);
            if constexpr (NT) __builtin_nontemporal_store(out, y + r);
            else y[r] = out;
        }
    };
    if constexpr (spmv_staged<LPR>()) {
        constexpr int RT = spmv_tile<LPR, T>::RT, CAP = spmv_tile<LPR, T>::CAP;
        __shared__ dev::csr_stage<T, RT, CAP> S;
        const dev::xcd_rows R = dev::xcd_rows::of(rows, RT);
        for (long r0 = R.first; r0 < R.end; r0 += R.stride) {
            const int nr = (int)((R.end - r0) < RT ? (R.end - r0) : RT);
            const bool staged = dev::csr_stage_tile<T, RT, CAP, SPMV_BLOCK, NT>(S, r0, nr, rp, col, val, x);
            const int e0 = S.rp[0];
            for (int rr = threadIdx.x / LPR; rr < nr; rr += SPMV_BLOCK / LPR) {
                T s = T(0), c = T(0);
                if (MODE == 1 && lane == 0) s = b[r0 + rr];
                if (staged)
                    dev::csr_row_acc_staged<T, LPR, COMP, MODE == 1>(S.rp[rr] - e0, S.rp[rr + 1] - e0, lane, S.v, S.x,
                                                                    s, c);
                else
                    dev::csr_row_acc<T, LPR, COMP, MODE == 1>(S.rp[rr], S.rp[rr + 1], lane, col, val, x, s, c);
                finish(r0 + rr, s, c);
            }
            __syncthreads();   // the next tile overwrites S
        }
    } else {
        const dev::xcd_rows R = dev::xcd_rows::of(rows, SPMV_BLOCK / LPR);
        for (long r = R.first + (long)threadIdx.x / LPR; r < R.end; r += R.stride) {
            T s = T(0), c = T(0);
            if (MODE == 1 && lane == 0) s = b[r];
            dev::csr_row_acc<T, LPR, COMP, MODE == 1>(rp[r], rp[r + 1], lane, col, val, x, s, c);
            finish(r, s, c);
        }
    }
}

// Staged tiles stream the columns and values in and y out nontemporally (NT).
template <typename T, bool COMP, int MODE, int LPR>
static void spmv_launch_l(hpb_handle_t h, int grid, int rows, const int* rp, const int* col, const T* val,
                          const T* x, T alpha, T beta, const T* b, T* y, const int* skip, const T* dalpha)
{
    hipLaunchKernelGGL((k_spmv<T, LPR, COMP, MODE, spmv_staged<LPR>()>), dim3(grid), dim3(SPMV_BLOCK), 0, h->stream,
                       rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha);
}

// Rows of up to 8 entries on average fill the staged tiles (128 rows, 1024 entries)
// and take 2 lanes: 256^3 7-point FP64, 5271 GB/s with 2 lanes, 5156 with 4, 3397
// with 8 (direct rows); 4096^2 5-point, 5246 with 2, 5045 with 4, 3689 with 1
// (128 blocks per CU).
// Longer rows read their entries directly, about two entries per lane.
static int pick_lpr(long rows, long nnz)
{
    const double mean = rows > 0 ? (double)nnz / (double)rows : 1.0;
    if (mean <= 1.5) return 1;
    if (mean <= 8.0) return 2;
    if (mean <= 12.0) return 4;
    if (mean <= 24.0) return 8;
    if (mean <= 48.0) return 16;
    if (mean <= 96.0) return 32;
    return 64;
}

template <typename T, bool COMP, int MODE>
static int spmv_entry(hpb_handle_t h, int rows, int cols, long nnz, const int* rp, const int* col, const T* val,
                      const T* x, T alpha, T beta, const T* b, T* y, int lpr, const int* skip = nullptr,
                      const T* dalpha = nullptr)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h, HPB_ERR_INVALID, "csr spmv: null handle");
    HPB_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, HPB_ERR_INVALID, "csr spmv: negative size");
    if (rows == 0) return HPB_OK;
    HPB_REQUIRE(rp && y && (nnz == 0 || (col && val && x)), HPB_ERR_INVALID, "csr spmv: null pointer");
    HPB_REQUIRE(MODE == 0 || b, HPB_ERR_INVALID, "csr residual: null b");
    if (lpr <= 0) lpr = pick_lpr(rows, nnz);
    HPB_REQUIRE(lpr == 1 || lpr == 2 || lpr == 4 || lpr == 8 || lpr == 16 || lpr == 32 || lpr == 64,
                HPB_ERR_INVALID, "csr spmv: lanes per row must be a power of two <= 64");
    const long rows_per_block = lpr == 1 ? spmv_tile<1, T>::RT : lpr <= 4 ? spmv_tile<4, T>::RT : SPMV_BLOCK / lpr;
    long grid = (rows + rows_per_block - 1) / rows_per_block;
    // staged tiles: up to 256 blocks per CU, so a block forms a few tiles, strided
    // over its XCD's slice; 256^3 7-point: 5455 GB/s, against 5334 with one tile
    // per block, 5265 with 128 blocks per CU and 4918 with 16 (HBM reads 1.08, 1.01,
    // 1.18 and 1.02 x the algorithmic bytes).  Direct rows: 16 blocks per CU.
    const long cap = (long)(h->num_cu > 0 ? h->num_cu : 256) * (lpr <= 4 ? 256 : 16);
    if (grid > cap) grid = cap;
    if (grid < 1) grid = 1;
    switch (lpr) {
    case 1: spmv_launch_l<T, COMP, MODE, 1>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    case 2: spmv_launch_l<T, COMP, MODE, 2>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    case 4: spmv_launch_l<T, COMP, MODE, 4>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    case 8: spmv_launch_l<T, COMP, MODE, 8>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    case 16: spmv_launch_l<T, COMP, MODE, 16>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    case 32: spmv_launch_l<T, COMP, MODE, 32>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    default: spmv_launch_l<T, COMP, MODE, 64>(h, (int)grid, rows, rp, col, val, x, alpha, beta, b, y, skip, dalpha); break;
    }
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

// Internal launcher for the solver graphs: MODE 0 y = A x, MODE 1 y = b - A x.
template <typename T>
int csr_apply(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* col, const T* val, const T* x,
              const T* b, T* y, bool comp, const int* skip, const T* dev_alpha)
{
    if (b) {
        return comp ? spmv_entry<T, true, 1>(h, rows, rows, nnz, rp, col, val, x, T(1), T(0), b, y, 0, skip)
                    : spmv_entry<T, false, 1>(h, rows, rows, nnz, rp, col, val, x, T(1), T(0), b, y, 0, skip);
    }
    return comp ? spmv_entry<T, true, 0>(h, rows, rows, nnz, rp, col, val, x, T(1), T(0), nullptr, y, 0, skip,
                                         dev_alpha)
                : spmv_entry<T, false, 0>(h, rows, rows, nnz, rp, col, val, x, T(1), T(0), nullptr, y, 0, skip,
                                          dev_alpha);
}
template int csr_apply<double>(hpb_handle_t, int, int64_t, const int*, const int*, const double*, const double*,
                               const double*, double*, bool, const int*, const double*);
template int csr_apply<float>(hpb_handle_t, int, int64_t, const int*, const int*, const float*, const float*,
                              const float*, float*, bool, const int*, const float*);

// ---------------------------------------------------------------------------
// structured test matrices
// ---------------------------------------------------------------------------

// Fill col/val of a 7-point (nz > 1) or 5-point (nz == 1) stencil, lexicographic
// order idx = i + nx*(j + ny*k), neighbours outside the box dropped (Dirichlet),
// columns ascending: k-1, j-1, i-1, diag, i+1, j+1, k+1.
// coef = {zm, ym, xm, diag, xp, yp, zp}.
__global__ __launch_bounds__(256) void k_stencil7(int nx, int ny, int nz, const int* __restrict__ rp, double c_zm,
                                                  double c_ym, double c_xm, double c_d, double c_xp, double c_yp,
                                                  double c_zp, int* __restrict__ col, double* __restrict__ val)
{
    const long n = (long)nx * ny * nz;
    for (long r = (long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long)gridDim.x * 256) {
        const int i = (int)(r % nx);
        const int j = (int)((r / nx) % ny);
        const int k = (int)(r / ((long)nx * ny));
        int p = rp[r];
        if (k > 0) { col[p] = (int)(r - (long)nx * ny); val[p++] = c_zm; }
        if (j > 0) { col[p] = (int)(r - nx); val[p++] = c_ym; }
        if (i > 0) { col[p] = (int)(r - 1); val[p++] = c_xm; }
        col[p] = (int)r; val[p++] = c_d;
        if (i < nx - 1) { col[p] = (int)(r + 1); val[p++] = c_xp; }
        if (j < ny - 1) { col[p] = (int)(r + nx); val[p++] = c_yp; }
        if (k < nz - 1) { col[p] = (int)(r + (long)nx * ny); val[p++] = c_zp; }
    }
}

static void stencil_row_ptr(int nx, int ny, int nz, std::vector<int>& rp)
{
    const long n = (long)nx * ny * nz;
    rp.resize(n + 1);
    long p = 0;
    long r = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i, ++r) {
                rp[r] = (int)p;
                p += 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
            }
    rp[n] = (int)p;
}

// Log-normal cell field kappa = exp(sigma * g), g ~ N(0,1) by Box-Muller from
// Philox (key = seed, counter = (cell, 0, 0, 0)).
__global__ __launch_bounds__(256) void k_lognormal_field(long ncell, double sigma, uint64_t seed,
                                                         double* __restrict__ kappa)
{
    for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (long)gridDim.x * 256) {
        philox_out o = philox4x32_10((uint32_t)c, (uint32_t)(c >> 32), 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = unif53_open0(k53(o.v[0], o.v[1]));
        const double u2 = unif53(k53(o.v[2], o.v[3]));
        const double g = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
        kappa[c] = exp(sigma * g);
    }
}

// 5-point diffusion on an m x m grid with cell coefficients kappa: face
// coefficient = harmonic mean 2 a b / (a + b) between cells, kappa of the cell on
// boundary faces; row: -k_face for each neighbour, diagonal = sum of the four faces.
__global__ __launch_bounds__(256) void k_diffusion2d(int m, const int* __restrict__ rp,
                                                     const double* __restrict__ kappa, int* __restrict__ col,
                                                     double* __restrict__ val)
{
    const long n = (long)m * m;
    for (long r = (long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long)gridDim.x * 256) {
        const int i = (int)(r % m), j = (int)(r / m);
        const double kc = kappa[r];
        auto face = [&](long nb) { const double kn = kappa[nb]; return 2.0 * kc * kn / (kc + kn); };
        const double fs = j > 0 ? face(r - m) : kc;
        const double fw = i > 0 ? face(r - 1) : kc;
        const double fe = i < m - 1 ? face(r + 1) : kc;
        const double fn = j < m - 1 ? face(r + m) : kc;
        int p = rp[r];
        if (j > 0) { col[p] = (int)(r - m); val[p++] = -fs; }
        if (i > 0) { col[p] = (int)(r - 1); val[p++] = -fw; }
        col[p] = (int)r; val[p++] = ((fs + fw) + fe) + fn;
        if (i < m - 1) { col[p] = (int)(r + 1); val[p++] = -fe; }
        if (j < m - 1) { col[p] = (int)(r + m); val[p++] = -fn; }
    }
}

static int fill_grid(hpb_handle_t h, long n)
{
    long g = (n + 255) / 256;
    const long cap = (long)(h->num_cu > 0 ? h->num_cu : 256) * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_csr_spmv_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const double* val, double alpha, const double* x, double beta, double* y)
{ return spmv_entry<double, false, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, 0); }
int hpb_csr_spmv_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                     const float* val, float alpha, const float* x, float beta, float* y)
{ return spmv_entry<float, false, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, 0); }
int hpb_csr_spmvc_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const double* val, double alpha, const double* x, double beta, double* y)
{ return spmv_entry<double, true, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, 0); }
int hpb_csr_spmvc_f32(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const float* val, float alpha, const float* x, float beta, float* y)
{ return spmv_entry<float, true, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, 0); }
int hpb_csr_residual_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                         const double* val, const double* b, const double* x, double* r)
{ return spmv_entry<double, false, 1>(h, rows, cols, nnz, row_ptr, col_ind, val, x, 1.0, 0.0, b, r, 0); }
int hpb_csr_residualc_f64(hpb_handle_t h, int rows, int cols, int64_t nnz, const int* row_ptr, const int* col_ind,
                          const double* val, const double* b, const double* x, double* r)
{ return spmv_entry<double, true, 1>(h, rows, cols, nnz, row_ptr, col_ind, val, x, 1.0, 0.0, b, r, 0); }
int hpb_csr_spmv_lpr_f64(hpb_handle_t h, int lpr, int comp, int rows, int cols, int64_t nnz, const int* row_ptr,
                         const int* col_ind, const double* val, double alpha, const double* x, double beta, double* y)
{
    if (comp) return spmv_entry<double, true, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, lpr);
    return spmv_entry<double, false, 0>(h, rows, cols, nnz, row_ptr, col_ind, val, x, alpha, beta, nullptr, y, lpr);
}

int hpb_stencil_nnz(int nx, int ny, int nz, int64_t* out_rows, int64_t* out_nnz)
{
    HPB_REQUIRE(nx > 0 && ny > 0 && nz > 0 && out_rows && out_nnz, HPB_ERR_INVALID, "hpb_stencil_nnz: bad argument");
    const int64_t n = (int64_t)nx * ny * nz;
    HPB_REQUIRE(n < (1l << 31), HPB_ERR_INVALID, "hpb_stencil_nnz: more than 2^31-1 rows");
    const int64_t nnz = n + 2l * ((long)(nx - 1) * ny * nz + (long)nx * (ny - 1) * nz + (long)nx * ny * (nz - 1));
    HPB_REQUIRE(nnz < (1l << 31), HPB_ERR_INVALID, "hpb_stencil_nnz: more than 2^31-1 non-zeros (int32 indices)");
    *out_rows = n;
    *out_nnz = nnz;
    return HPB_OK;
}

int hpb_csr_stencil7_f64(hpb_handle_t h, int nx, int ny, int nz, const double* coef7, int* row_ptr, int* col_ind,
                         double* val)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && coef7 && row_ptr && col_ind && val, HPB_ERR_INVALID, "hpb_csr_stencil7_f64: null argument");
    int64_t n, nnz;
    int st = hpb_stencil_nnz(nx, ny, nz, &n, &nnz);
    if (st) return st;
    std::vector<int> rp;
    stencil_row_ptr(nx, ny, nz, rp);
    HPB_HIP(hipMemcpyAsync(row_ptr, rp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_stencil7, dim3(fill_grid(h, n)), dim3(256), 0, h->stream, nx, ny, nz, row_ptr, coef7[0],
                       coef7[1], coef7[2], coef7[3], coef7[4], coef7[5], coef7[6], col_ind, val);
    HPB_LAUNCH_CHECK();
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

int hpb_lognormal_field_f64(hpb_handle_t h, int64_t ncell, double sigma, uint64_t seed, double* kappa)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && kappa && ncell > 0, HPB_ERR_INVALID, "hpb_lognormal_field_f64: bad argument");
    hipLaunchKernelGGL(k_lognormal_field, dim3(fill_grid(h, ncell)), dim3(256), 0, h->stream, ncell, sigma, seed, kappa);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

int hpb_csr_diffusion2d_f64(hpb_handle_t h, int m, const double* kappa, int* row_ptr, int* col_ind, double* val)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && kappa && row_ptr && col_ind && val && m > 0, HPB_ERR_INVALID, "hpb_csr_diffusion2d_f64: bad argument");
    int64_t n, nnz;
    int st = hpb_stencil_nnz(m, m, 1, &n, &nnz);
    if (st) return st;
    std::vector<int> rp;
    stencil_row_ptr(m, m, 1, rp);
    HPB_HIP(hipMemcpyAsync(row_ptr, rp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_diffusion2d, dim3(fill_grid(h, n)), dim3(256), 0, h->stream, m, row_ptr, kappa, col_ind, val);
    HPB_LAUNCH_CHECK();
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

}  // extern "C"
