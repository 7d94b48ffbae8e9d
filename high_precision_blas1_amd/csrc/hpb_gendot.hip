// hpb_gendot.hip -- device-side GenDot generator of ill-conditioned FP64 pairs
// with an exactly known dot product.
//
// Replaces generate_vector_pair<...>::generate (source/generate_vector_pair.hpp:119-182),
// which is host-driven, seeded from std::random_device and runs one full-length
// GPU Dot2 per fixed tail element (O(N_fix * N)).  Here the construction is
// Ogita-Rump-Oishi's Alg. 6.1: a random head whose exponents are uniform in
// [0, b/2], b = log2(cond), and a tail with exponents descending b/2 -> 0 that
// cancels the running dot.  Entries are multiples of 2^-53 so that the running dot
// is exact in a 256-bit fixed-point accumulator; the tail is built in chunks whose
// last element cancels the EXACT running sum (a sequential pass over chunks),
// every other tail element cancels its predecessor's target.  Chunks hold 256
// elements, doubled until there are at most 4096 of them, so the sequential pass
// stays short at any length.  The arithmetic is specified so that the host
// restatement in the oracle gives the same bits (see hpb_oracle.c, "GenDot-style
// ... generator").
//
// Passes: k_head (parallel, per-block exact partials) -> k_reduce_parts (one block)
// -> k_tail_chunks (one block per chunk: every element but the chunk's last only
// needs its own and its predecessor's draws, so the lanes take them in parallel
// and sum their exact products in LDS) -> k_tail_fix (one wave; the chain over
// chunks is sequential) -> results.
#include <hip/hip_runtime.h>

#include <cmath>

#include "hpb_acc256.hpp"
#include "hpb_device.hpp"
#include "hpb_internal.hpp"
#include "hpb_philox.hpp"

namespace hpb {

constexpr int GD_CHUNK_MIN = 256;
constexpr uint64_t GD_MAX_CHUNKS = 4096;
constexpr int GD_BLOCK = 256;

// tail chunk length: 256, doubled until the tail has at most GD_MAX_CHUNKS chunks
static uint64_t gd_chunk_len(uint64_t n_tail)
{
    uint64_t c = GD_CHUNK_MIN;
    while ((n_tail + c - 1) / c > GD_MAX_CHUNKS) c *= 2;
    return c;
}

struct gd_draws {
    uint64_t ka, kb, kc;
};

__device__ __forceinline__ gd_draws gd_draw(uint64_t seed, uint64_t i)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    philox_out w = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 0u, 0u, k0, k1);
    philox_out v = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 1u, 0u, k0, k1);
    gd_draws d;
    d.ka = k53(w.v[0], w.v[1]);
    d.kb = k53(w.v[2], w.v[3]);
    d.kc = k53(v.v[0], v.v[1]);
    return d;
}

__device__ __forceinline__ double gd_mant(uint64_t k)
{
    return (double)((int64_t)(2 * k + 1) - ((int64_t)1 << 53)) * 0x1p-53;
}
__device__ __forceinline__ double gd_grid53(double v)
{
    return fabs(v) >= 0.5 ? v : rint(v * 0x1p53) * 0x1p-53;
}
__device__ __forceinline__ double gd_exp_head(double hb, uint64_t i, uint64_t n_head, uint64_t ka)
{
    if (i == 0) return rint(hb) + 1.0;
    if (i == n_head - 1) return 0.0;
    return rint(unif53(ka) * hb);
}
__device__ __forceinline__ double gd_exp_tail(double hb, uint64_t j, uint64_t n_tail)
{
    if (n_tail <= 1) return 0.0;
    return rint(hb * (1.0 - (double)j / (double)(n_tail - 1)));
}

// exact |x y| = |p| + sign(p) e  for p = fl(xy), e = xy - p
__device__ __forceinline__ void gd_add_prod(acc256& S, acc256& A, double x, double y)
{
    double e;
    double p = dev::two_prod<double>(x, y, e);
    acc_add_double(S, p);
    acc_add_double(S, e);
    acc_add_double(A, fabs(p));
    acc_add_double(A, p < 0.0 ? -e : e);
}

// Sum the block's accumulators through LDS (integer adds: order-free).
__device__ void block_sum_acc(acc256& S, acc256& A)
{
    __shared__ uint64_t lds[GD_BLOCK][8];
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) {
        lds[t][k] = S.w[k];
        lds[t][4 + k] = A.w[k];
    }
    __syncthreads();
    for (int st = GD_BLOCK / 2; st >= 1; st >>= 1) {
        if (t < st) {
            acc256 a, b;
            for (int k = 0; k < 4; ++k) {
                a.w[k] = lds[t][k];
                b.w[k] = lds[t + st][k];
            }
            acc_add(a, b);
            for (int k = 0; k < 4; ++k) {
                lds[t][k] = a.w[k];
                a.w[k] = lds[t][4 + k];
                b.w[k] = lds[t + st][4 + k];
            }
            acc_add(a, b);
            for (int k = 0; k < 4; ++k) lds[t][4 + k] = a.w[k];
        }
        __syncthreads();
    }
    for (int k = 0; k < 4; ++k) {
        S.w[k] = lds[0][k];
        A.w[k] = lds[0][4 + k];
    }
}

__global__ __launch_bounds__(GD_BLOCK) void k_head(uint64_t n_head, double hb, uint64_t seed, double* __restrict__ x,
                                                   double* __restrict__ y, uint64_t* __restrict__ part)
{
    acc256 S, A;
    acc_zero(S);
    acc_zero(A);
    for (uint64_t i = (uint64_t)blockIdx.x * GD_BLOCK + threadIdx.x; i < n_head; i += (uint64_t)gridDim.x * GD_BLOCK) {
        gd_draws d = gd_draw(seed, i);
        const int e = (int)gd_exp_head(hb, i, n_head, d.ka);
        const double xi = ldexp(gd_mant(d.kb), e);
        const double yi = ldexp(gd_mant(d.kc), e);
        x[i] = xi;
        y[i] = yi;
        gd_add_prod(S, A, xi, yi);
    }
    block_sum_acc(S, A);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            part[blockIdx.x * 8 + k] = S.w[k];
            part[blockIdx.x * 8 + 4 + k] = A.w[k];
        }
}

// One block: sum nparts partials -> out[0..7]
__global__ __launch_bounds__(GD_BLOCK) void k_reduce_parts(int nparts, const uint64_t* __restrict__ part,
                                                           uint64_t* __restrict__ out)
{
    acc256 S, A;
    acc_zero(S);
    acc_zero(A);
    for (int p = threadIdx.x; p < nparts; p += GD_BLOCK) {
        acc256 s, a;
        for (int k = 0; k < 4; ++k) {
            s.w[k] = part[p * 8 + k];
            a.w[k] = part[p * 8 + 4 + k];
        }
        acc_add(S, s);
        acc_add(A, a);
    }
    block_sum_acc(S, A);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            out[k] = S.w[k];
            out[4 + k] = A.w[k];
        }
}

// One block per tail chunk.  head_acc = exact head sum (S limbs 0..3).  Element j
// of the chunk (not its last) takes y = grid53((r_j - r_{j-1}) / x_j) from its own
// draws and its predecessor's (the head sum, rounded, for the first tail element);
// the chunk's exact partial sums are taken over those elements.
__global__ __launch_bounds__(GD_BLOCK) void k_tail_chunks(uint64_t n_head, uint64_t n_tail, uint64_t chunk, double hb,
                                                          uint64_t seed, const uint64_t* __restrict__ head_acc,
                                                          double* __restrict__ x, double* __restrict__ y,
                                                          uint64_t* __restrict__ cpart, double* __restrict__ r_last)
{
    const uint64_t ch = blockIdx.x;
    const uint64_t j0 = ch * chunk;
    uint64_t j1 = j0 + chunk;
    if (j1 > n_tail) j1 = n_tail;
    acc256 S, A;
    acc_zero(S);
    acc_zero(A);
    for (uint64_t j = j0 + threadIdx.x; j < j1; j += GD_BLOCK) {
        const uint64_t i = n_head + j;
        gd_draws d = gd_draw(seed, i);
        const int e = (int)gd_exp_tail(hb, j, n_tail);
        const double xi = ldexp(gd_mant(d.kb), e);
        const double r = ldexp(gd_mant(d.kc), e);
        x[i] = xi;
        if (j + 1 < j1) {
            double r_prev;
            if (j == 0) {
                acc256 Sh;
                for (int k = 0; k < 4; ++k) Sh.w[k] = head_acc[k];
                r_prev = acc_round(Sh);
            } else {
                gd_draws dp = gd_draw(seed, i - 1);
                r_prev = ldexp(gd_mant(dp.kc), (int)gd_exp_tail(hb, j - 1, n_tail));
            }
            const double yi = gd_grid53((r - r_prev) / xi);
            y[i] = yi;
            gd_add_prod(S, A, xi, yi);
        } else {
            r_last[ch] = r;
        }
    }
    block_sum_acc(S, A);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            cpart[ch * 8 + k] = S.w[k];
            cpart[ch * 8 + 4 + k] = A.w[k];
        }
}

// Sequential pass over the chunks: the last element of every chunk cancels the
// exact running sum.  acc[0..7] holds the head sums on entry and the totals on exit.
// The 64 lanes stage the inputs of FIX_BATCH chunks in LDS (independent loads in
// parallel), lane 0 runs the dependent chain from LDS, then the lanes store the
// new y values; the arithmetic is the sequential one.
constexpr int FIX_BATCH = 64;

__global__ __launch_bounds__(64) void k_tail_fix(uint64_t n_head, uint64_t n_tail, uint64_t chunk,
                                                 const uint64_t* __restrict__ cpart,
                                                 const double* __restrict__ r_last, double* __restrict__ x,
                                                 double* __restrict__ y, uint64_t* __restrict__ acc,
                                                 double* __restrict__ out3)
{
    __shared__ uint64_t sp[FIX_BATCH * 8];
    __shared__ double sx[FIX_BATCH], sr[FIX_BATCH], sy[FIX_BATCH];
    acc256 S, A;
    for (int k = 0; k < 4; ++k) {
        S.w[k] = acc[k];
        A.w[k] = acc[4 + k];
    }
    const uint64_t nchunks = (n_tail + chunk - 1) / chunk;
    const int lane = threadIdx.x;
    for (uint64_t c0 = 0; c0 < nchunks; c0 += FIX_BATCH) {
        const int nb = (int)((nchunks - c0) < (uint64_t)FIX_BATCH ? (nchunks - c0) : (uint64_t)FIX_BATCH);
        for (int q = lane; q < nb * 8; q += 64) sp[q] = cpart[c0 * 8 + q];
        if (lane < nb) {
            const uint64_t ch = c0 + lane;
            uint64_t j1 = (ch + 1) * chunk;
            if (j1 > n_tail) j1 = n_tail;
            sx[lane] = x[n_head + j1 - 1];
            sr[lane] = r_last[ch];
        }
        __syncthreads();
        if (lane == 0) {
            for (int b = 0; b < nb; ++b) {
                acc256 c, a;
                for (int k = 0; k < 4; ++k) {
                    c.w[k] = sp[b * 8 + k];
                    a.w[k] = sp[b * 8 + 4 + k];
                }
                acc_add(S, c);
                acc_add(A, a);
                const double xi = sx[b];
                const double yi = gd_grid53((sr[b] - acc_round(S)) / xi);
                sy[b] = yi;
                gd_add_prod(S, A, xi, yi);
            }
        }
        __syncthreads();
        if (lane < nb) {
            const uint64_t ch = c0 + lane;
            uint64_t j1 = (ch + 1) * chunk;
            if (j1 > n_tail) j1 = n_tail;
            y[n_head + j1 - 1] = sy[lane];
        }
        __syncthreads();
    }
    if (lane == 0) {
        for (int k = 0; k < 4; ++k) {
            acc[k] = S.w[k];
            acc[4 + k] = A.w[k];
        }
        const double d = acc_round(S), ab = acc_round(A);
        out3[0] = d;
        out3[1] = ab;
        out3[2] = d != 0.0 ? ab / fabs(d) : INFINITY;
    }
}

}  // namespace hpb

using namespace hpb;

extern "C" int hpb_gendot_f64(hpb_handle_t h, size_t n, double log2cond, uint64_t seed, size_t n_tail, double* x,
                              double* y, uint64_t* out_exact4, double* out3)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && x && y, HPB_ERR_INVALID, "hpb_gendot_f64: null argument");
    HPB_REQUIRE(n >= 2 && n_tail >= 1 && n_tail < n, HPB_ERR_INVALID, "hpb_gendot_f64: need n >= 2, 1 <= n_tail < n");
    const double hb = 0.5 * log2cond;
    HPB_REQUIRE(hb >= 0.0 && log2cond + 4.0 + std::log2((double)n) <= 147.0, HPB_ERR_INVALID,
                "hpb_gendot_f64: condition number out of range (need 1 <= cond, log2(cond) + log2(n) <= 143)");
    const size_t n_head = n - n_tail;
    const uint64_t chunk = gd_chunk_len(n_tail);
    const size_t nchunks = (n_tail + chunk - 1) / chunk;
    size_t gh = (n_head + GD_BLOCK - 1) / GD_BLOCK;
    const size_t gcap = (size_t)(h->num_cu > 0 ? h->num_cu : 256) * 4;
    if (gh > gcap) gh = gcap;
    if (gh < 1) gh = 1;
    // workspace: head partials | head/total acc (8) | chunk partials | r_last | out3
    const size_t bytes = gh * 64 + 64 + nchunks * 64 + nchunks * 8 + 32;
    void* ws = nullptr;
    HPB_HIP(hipMalloc(&ws, bytes));
    uint64_t* hpart = (uint64_t*)ws;
    uint64_t* acc = hpart + gh * 8;
    uint64_t* cpart = acc + 8;
    double* rl = (double*)(cpart + nchunks * 8);
    double* o3 = rl + nchunks;
    hipLaunchKernelGGL(k_head, dim3((unsigned)gh), dim3(GD_BLOCK), 0, h->stream, (uint64_t)n_head, hb, seed, x, y, hpart);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(GD_BLOCK), 0, h->stream, (int)gh, hpart, acc);
    hipLaunchKernelGGL(k_tail_chunks, dim3((unsigned)nchunks), dim3(GD_BLOCK), 0, h->stream, (uint64_t)n_head,
                       (uint64_t)n_tail, chunk, hb, seed, acc, x, y, cpart, rl);
    hipLaunchKernelGGL(k_tail_fix, dim3(1), dim3(64), 0, h->stream, (uint64_t)n_head, (uint64_t)n_tail, chunk, cpart, rl,
                       x, y, acc, o3);
    hipError_t le = hipGetLastError();
    uint64_t hacc[8];
    double ho[3];
    hipError_t e1 = hipMemcpyAsync(hacc, acc, sizeof hacc, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipMemcpyAsync(ho, o3, sizeof ho, hipMemcpyDeviceToHost, h->stream);
    hipError_t e3 = hipStreamSynchronize(h->stream);
    (void)hipFree(ws);
    if (le != hipSuccess) return hpb::hip_status(le, __FILE__, __LINE__, "gendot kernels");
    if (e1 != hipSuccess) return hpb::hip_status(e1, __FILE__, __LINE__, "hipMemcpyAsync");
    if (e2 != hipSuccess) return hpb::hip_status(e2, __FILE__, __LINE__, "hipMemcpyAsync");
    if (e3 != hipSuccess) return hpb::hip_status(e3, __FILE__, __LINE__, "hipStreamSynchronize");
    if (out_exact4)
        for (int k = 0; k < 4; ++k) out_exact4[k] = hacc[k];
    if (out3)
        for (int k = 0; k < 3; ++k) out3[k] = ho[k];
    return HPB_OK;
}
