// hpb_batch.hip -- many independent CSR systems solved together by BiCGStab(2),
// one captured hipGraph per BiCGStab(2) cycle.
//
// No reference counterpart: the reference solves one system at a time with a host
// round trip per dot (numerical_algos/lin_solvers/bicgstabl.h:151-274).  Here all
// systems live in one block-diagonal CSR matrix; rows are cut into tiles that never
// cross a system, so every kernel finds its system's scalars from the tile.
//   * segmented compensated dots: per-tile (s, c) partials, merged per system in
//     tile order by one wave (reproducible);
//   * per-system scalar recurrences on the device (one lane per system);
//   * updates and SpMV over all systems in one launch each; tiles of converged
//     systems return at once;
//   * every dot product of the cycle is an epilogue of the SpMV that makes its
//     new vector, so no pass reads a fresh vector back only to take its dots;
//   * the whole cycle (5 SpMV incl. the true-residual recomputation, 5 update
//     kernels, 6 scalar kernels) is captured once as a hipGraph and replayed; the
//     host reads the count of active systems every few cycles;
//   * by default (resid_recalc_freq 0, the reference's default) the cycles test
//     the recursive residual and compute the true one only for the systems that
//     claim convergence; with resid_recalc_freq k every k-th cycle replaces the
//     recursive residual by the true one for all systems (a second captured
//     graph, chosen per cycle by the host).
// The recurrences are the reference's BiCGStab(L) for L = 2 (bicgstabl.h:199-258)
// with the MR part evaluated from one fused Gram matrix pass: tau12 = (r2,r1)/s1,
// s2 = (r2,r2) - 2 tau12 (r2,r1) + tau12^2 s1, which equals the modified
// Gram-Schmidt values in exact arithmetic (the dots are compensated).
//
// Test matrices: 2-D 5-point diffusion, cell coefficients exp(sigma g), g ~ N(0,1)
// from Philox (key seed + s * 0x9E3779B97F4A7C15, counter = cell), harmonic face
// means as in hpb_csr_diffusion2d_f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <vector>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"
#include "hpb_philox.hpp"

namespace hpb {

constexpr int BT_ROWS = 2048;  // rows per tile
constexpr int BT_THREADS = 256;
constexpr int BQ = 5;          // dot products per group, at most

struct btile {
    int sys;
    int pad;
    int64_t r0, r1;  // global rows [r0, r1)
};

// state per system
enum { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_BNORM, S_RNORM, S_G1, S_GPP1, S_GP1, S_GP2, S_TAU, S_G2, S_NSTATE };

struct hpb_batch_s;
struct batch_t {
    hpb_handle_t h = nullptr;
    int nsys = 0;
    int64_t nrows = 0, nnz = 0;
    int ntiles = 0;
    std::vector<int64_t> row_off;   // nsys + 1
    std::vector<int> tile_first;    // nsys + 1
    btile* d_tiles = nullptr;
    int* d_tile_first = nullptr;
    int64_t* d_row_off = nullptr;
    int* rp = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    double* v[10] = {};             // b, x, r0, r1, r2, u0, u1, u2, rt, xs
    double* st = nullptr;           // nsys * S_NSTATE
    int* active = nullptr;          // nsys
    int* iters = nullptr;           // nsys
    int* nactive = nullptr;         // 1
    double* part = nullptr;         // ntiles * BQ * 2
    double* dots = nullptr;         // nsys * BQ
    int* hcount = nullptr;          // pinned
    int* tile_list = nullptr;       // ntiles: indices of the tiles of active systems
    int* tile_count = nullptr;      // 1: entries of tile_list in use
    double* dinv = nullptr;         // nrows: 1 / a_ii (Jacobi right preconditioning)
    double* vald = nullptr;         // nnz: the values of A D (columns scaled by dinv), made on first use
    int jacobi = 0;
    int* confirm = nullptr;         // nsys: the recursive residual met the tolerance this cycle
    int recalc = 0;                 // true residual b - A x every recalc cycles (0: only to confirm)
    double graph_tol2 = -1.0;       // settings the captured graphs were recorded with
    int graph_jacobi = -1;
    int grid = 1;                   // blocks of the SpMV and dot kernels (over tile_list)
    int ugrid = 1;                  // blocks of the update kernels
    hipGraph_t graph = nullptr;     // cycle that recomputes the true residual
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph2 = nullptr;    // cycle on the recursive residual
    hipGraphExec_t exec2 = nullptr;
    int verbose = 0;                // print the active count at every check
    struct pbar* bar = nullptr;     // grid barrier state of the persistent launch
};
enum { VB = 0, VX, VR0, VR1, VR2, VU0, VU1, VU2, VRT, VXS };

// ---------------------------------------------------------------- builders

__device__ __forceinline__ double cell_kappa(int m, int i, int j, double sigma, uint64_t seed)
{
    const int64_t c = (int64_t)i + (int64_t)m * j;
    philox_out o = philox4x32_10((uint32_t)c, (uint32_t)(c >> 32), 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = unif53_open0(k53(o.v[0], o.v[1]));
    const double u2 = unif53(k53(o.v[2], o.v[3]));
    return exp(sigma * (sqrt(-2.0 * log(u1)) * cospi(2.0 * u2)));
}

// entries of the 5-point stencil rows before row (i, j) of an m x m grid
__device__ __forceinline__ int64_t rows_before(int m, int i, int j)
{
    const int64_t full = (int64_t)j * (3 * m - 2) + (int64_t)m * ((j > 0 ? j - 1 : 0) + (j < m - 1 ? j : m - 1));
    const int ny = (j > 0) + (j < m - 1);
    const int64_t part = (int64_t)i + (i > 0 ? i - 1 : 0) + (i < m - 1 ? i : m - 1) + (int64_t)i * ny;
    return full + part;
}

__device__ __forceinline__ int find_sys(const int64_t* row_off, int nsys, int64_t r)
{
    int lo = 0, hi = nsys;  // row_off[lo] <= r < row_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row_off[mid] <= r) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_batch_build(int nsys, int64_t nrows, const int64_t* __restrict__ row_off,
                                                     const int64_t* __restrict__ nnz_off, const int* __restrict__ side,
                                                     double sigma, uint64_t seed, int* __restrict__ rp,
                                                     int* __restrict__ col, double* __restrict__ val)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * 256) {
        const int s = find_sys(row_off, nsys, r);
        const int m = side[s];
        const int64_t lr = r - row_off[s];
        const int i = (int)(lr % m), j = (int)(lr / m);
        const uint64_t sd = seed + (uint64_t)s * 0x9E3779B97F4A7C15ull;
        const double kc = cell_kappa(m, i, j, sigma, sd);
        auto face = [&](int ii, int jj) {
            const double kn = cell_kappa(m, ii, jj, sigma, sd);
            return 2.0 * kc * kn / (kc + kn);
        };
        const double fs = j > 0 ? face(i, j - 1) : kc;
        const double fw = i > 0 ? face(i - 1, j) : kc;
        const double fe = i < m - 1 ? face(i + 1, j) : kc;
        const double fn = j < m - 1 ? face(i, j + 1) : kc;
        const int64_t p0 = nnz_off[s] + rows_before(m, i, j);
        int64_t p = p0;
        const int64_t g = row_off[s];
        if (j > 0) { col[p] = (int)(g + lr - m); val[p++] = -fs; }
        if (i > 0) { col[p] = (int)(g + lr - 1); val[p++] = -fw; }
        col[p] = (int)r; val[p++] = ((fs + fw) + fe) + fn;
        if (i < m - 1) { col[p] = (int)(g + lr + 1); val[p++] = -fe; }
        if (j < m - 1) { col[p] = (int)(g + lr + m); val[p++] = -fn; }
        rp[r] = (int)p0;
        if (r == nrows - 1) rp[nrows] = (int)p;
    }
}

// ---------------------------------------------------------------- kernels over tiles
//
// The tile kernels walk the list of tiles of active systems.  The list is cut into
// one contiguous run per XCD (dev::xcd_rows over list positions), so that the
// neighbouring tiles of a system, whose SpMV rows gather each other's x entries,
// are read through the same L2.

// y = A x (MODE 0) or y = b - A x (MODE 1, compensated rows) for active systems;
// with Jacobi right preconditioning val holds the values of A D, the columns
// scaled once by dinv, so the row products gather x alone.
// EPI adds the dot products the cycle takes of the new vector, as per-tile
// compensated partials in the layout of k_bdots (slot q of tile ti at
// part[(ti * BQ + q) * 2]), so that no second pass reads y back:
//   EPI 0: none;
//   EPI 1: (y, rt) in slot 0;
//   EPI 2: (y, y) in slot 0 and (y, rt) in slot 1 (the true residual: its norm for
//          the convergence check, and rho for the next cycle's first BiCG step);
//   EPI 5: the Gram entries of the MR step with y = r2, x = r1 and z = r0:
//          (r1, r1), (r2, r1), (r2, r2), (r0, r1), (r0, r2) in slots 0..4.
// With only != null the tiles of systems with only[sys] == 0 are skipped.
constexpr int BSPMV_LPR = 2;                         // lanes per row of the batch SpMV
constexpr int BSPMV_SRT = BT_THREADS / BSPMV_LPR;    // staged sub-tile: rows (one per lane group)
constexpr int BSPMV_SCAP = 1024;                     // staged sub-tile: entries
using bstage = dev::csr_stage<double, BSPMV_SRT, BSPMV_SCAP>;

template <int MODE, int EPI>
__device__ __forceinline__ void bspmv_body(bstage& S, unsigned bid, unsigned nblk, const btile* __restrict__ tiles,
                                           const int* __restrict__ list, const int* __restrict__ count,
                                           const int* __restrict__ active, const int* __restrict__ rp,
                                           const int* __restrict__ col, const double* __restrict__ val,
                                           const double* __restrict__ x, const double* __restrict__ b,
                                           double* __restrict__ y, const double* __restrict__ rt,
                                           const double* __restrict__ z, double* __restrict__ part,
                                           const int* __restrict__ only)
{
    static_assert(EPI == 0 || EPI == 1 || EPI == 2 || EPI == 5, "k_bspmv: epilogue");
    constexpr int NQ = EPI == 0 ? 1 : EPI;
    constexpr int NW = BT_THREADS / 64;
    // wave totals of the tile partials; two buffers, alternating by tile, so that one
    // barrier per tile suffices: buffer k is read after the barrier of its tile and
    // written again only two tiles later, after every thread passed the next barrier
    __shared__ double wtot[2][NQ][NW][2];
    // lanes per row: on these 5-entry rows 2 lanes beat 4 by 8 % over a whole
    // cycle (1 lane: 8 % slower than 4, 8 lanes: 18 % slower)
    constexpr int LPR = BSPMV_LPR;
    constexpr int SRT = BSPMV_SRT, SCAP = BSPMV_SCAP;
    int buf = 0;
    const int nt = *count;
    const dev::xcd_rows Q = dev::xcd_rows::of(nt, 1, nblk, bid);
    for (long q = Q.first; q < Q.end; q += Q.stride) {
        const int ti = list[q];
        const btile t = tiles[ti];
        if (!active[t.sys] || (only != nullptr && !only[t.sys])) continue;
        const int lane = threadIdx.x % LPR;
        double ds[NQ], dc[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) ds[k] = dc[k] = 0.0;
        // sub-tiles of SRT rows, one per lane group, staged in LDS (csr_stage_tile):
        // each thread keeps the rows t.r0 + threadIdx.x / LPR + k SRT of the direct walk
        for (int64_t r0 = t.r0; r0 < t.r1; r0 += SRT) {
            const int nr = (int)((t.r1 - r0) < SRT ? (t.r1 - r0) : SRT);
            // columns and values streamed in, y out, nontemporally (the x lines stay in
            // L2: 19.81 s against 20.05 s for the 512-system solve)
            const bool staged = dev::csr_stage_tile<double, SRT, SCAP, BT_THREADS, true>(S, r0, nr, rp, col, val, x);
            const int eb = staged ? S.rp[0] : 0;
            const int rr = threadIdx.x / LPR;
            // (the barrier at the end of the sub-tile stays outside this branch: a wave
            // whose lanes split between two barriers would execute both)
            if (rr < nr) {
                const int64_t r = r0 + rr;
                // the epilogue's own operands, loaded ahead of the row's gathers
                double e0 = 0.0, e1 = 0.0;
                if (EPI != 0 && lane == 0) {
                    if (EPI == 5) {
                        e0 = x[r];
                        e1 = z[r];
                    } else {
                        e0 = rt[r];
                    }
                }
                double s = (MODE == 1 && lane == 0) ? b[r] : 0.0, c = 0.0;
                if (staged)
                    dev::csr_row_acc_staged<double, LPR, true, MODE == 1>(S.rp[rr] - eb, S.rp[rr + 1] - eb, lane, S.v,
                                                                         S.x, s, c);
                else
                    dev::csr_row_acc<double, LPR, true, MODE == 1>(S.rp[rr], S.rp[rr + 1], lane, col, val, x, s, c);
#pragma unroll
                for (int m = LPR / 2; m >= 1; m >>= 1) {
                    double s2 = __shfl_xor(s, m, 64), c2 = __shfl_xor(c, m, 64);
                    dev::merge<double>(s, c, s2, c2);
                }
                if (lane == 0) {
                    const double yv = s + c;
                    __builtin_nontemporal_store(yv, y + r);
                    if (EPI == 1) {
                        dev::dot2_step<double>(ds[0], dc[0], yv, e0);
                    } else if (EPI == 2) {
                        dev::dot2_step<double>(ds[0], dc[0], yv, yv);
                        dev::dot2_step<double>(ds[NQ - 1], dc[NQ - 1], yv, e0);
                    } else if (EPI == 5) {   // e0 = r1, e1 = r0, yv = r2
                        dev::dot2_step<double>(ds[0], dc[0], e0, e0);
                        dev::dot2_step<double>(ds[1 % NQ], dc[1 % NQ], yv, e0);
                        dev::dot2_step<double>(ds[2 % NQ], dc[2 % NQ], yv, yv);
                        dev::dot2_step<double>(ds[3 % NQ], dc[3 % NQ], e1, e0);
                        dev::dot2_step<double>(ds[4 % NQ], dc[4 % NQ], e1, yv);
                    }
                }
            }
            __syncthreads();   // the next sub-tile overwrites S
        }
        if (EPI != 0) {
            const int w = threadIdx.x >> 6;
#pragma unroll
            for (int k = 0; k < NQ; ++k) {
                dev::wave_reduce<double, true>(ds[k], dc[k]);
                if ((threadIdx.x & 63) == 0) {
                    wtot[buf][k][w][0] = ds[k];
                    wtot[buf][k][w][1] = dc[k];
                }
            }
            __syncthreads();
            if (threadIdx.x < NQ) {   // waves merged in wave order, as block_reduce does
                const int k = threadIdx.x;
                double fs = wtot[buf][k][0][0], fc = wtot[buf][k][0][1];
#pragma unroll
                for (int v = 1; v < NW; ++v) dev::merge<double>(fs, fc, wtot[buf][k][v][0], wtot[buf][k][v][1]);
                part[((size_t)ti * BQ + k) * 2] = fs;
                part[((size_t)ti * BQ + k) * 2 + 1] = fc;
            }
            buf ^= 1;
        }
    }
}

template <int MODE, int EPI>
__global__ __launch_bounds__(BT_THREADS) void k_bspmv(const btile* __restrict__ tiles, const int* __restrict__ list,
                                                      const int* __restrict__ count, const int* __restrict__ active,
                                                      const int* __restrict__ rp, const int* __restrict__ col,
                                                      const double* __restrict__ val, const double* __restrict__ x,
                                                      const double* __restrict__ b, double* __restrict__ y,
                                                      const double* __restrict__ rt,
                                                      const double* __restrict__ z, double* __restrict__ part,
                                                      const int* __restrict__ only)
{
    __shared__ bstage S;
    bspmv_body<MODE, EPI>(S, blockIdx.x, gridDim.x, tiles, list, count, active, rp, col, val, x, b, y, rt, z, part,
                          only);
}

struct dot_pairs {
    const double* a[BQ];
    const double* b[BQ];
};

// per-tile compensated partials of NQ dot products (tiles of active systems)
template <int NQ>
__global__ __launch_bounds__(BT_THREADS) void k_bdots(const btile* __restrict__ tiles, const int* __restrict__ list,
                                                      const int* __restrict__ count, const int* __restrict__ active,
                                                      dot_pairs P, double* __restrict__ part)
{
    const int nt = *count;
    const dev::xcd_rows Q = dev::xcd_rows::of(nt, 1);
    for (long q = Q.first; q < Q.end; q += Q.stride) {
        const int ti = list[q];
        const btile t = tiles[ti];
        if (!active[t.sys]) continue;
        double s[NQ], c[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) s[k] = c[k] = 0.0;
        for (int64_t r = t.r0 + threadIdx.x; r < t.r1; r += BT_THREADS) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) dev::dot2_step<double>(s[k], c[k], P.a[k][r], P.b[k][r]);
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            dev::block_reduce<double, true, BT_THREADS>(s[k], c[k]);
            if (threadIdx.x == 0) {
                part[((size_t)ti * BQ + k) * 2] = s[k];
                part[((size_t)ti * BQ + k) * 2 + 1] = c[k];
            }
        }
    }
}

// one wave per system: merge the system's tile partials of slots Q0 .. Q0 + NQ - 1
// in tile order
template <int NQ, int Q0>
__device__ __forceinline__ void merge_sys(int s, const int* tile_first, const double* part, double* d)
{
    const int lane = threadIdx.x & 63;
    for (int q = 0; q < NQ; ++q) {
        double fs = 0.0, fc = 0.0;
        for (int t = tile_first[s] + lane; t < tile_first[s + 1]; t += 64)
            dev::merge<double>(fs, fc, part[((size_t)t * BQ + Q0 + q) * 2], part[((size_t)t * BQ + Q0 + q) * 2 + 1]);
        dev::wave_reduce<double, true>(fs, fc);
        d[q] = fs + fc;
    }
}

// phase: 0 rho1 -> beta; 1 alpha; 2 MR; 3 true-residual check; 4 init (rnorm,
// bnorm); 5 recursive-residual check: a system that meets the tolerance is marked
// for confirmation; 6 true-residual check of the marked systems.  The NQ dot
// products come from slots Q0 ..
// One system per wave: wave w of block bid takes systems (bid * waves + w) +
// k * nblk * waves.  The 256-thread blocks of the standalone kernel take one
// system per wave.
template <int NQ, int PHASE, int Q0>
__device__ __forceinline__ void bsys_one(int s, const int* __restrict__ tile_first, const double* __restrict__ part,
                                         double* __restrict__ st, int* __restrict__ active, int* __restrict__ iters,
                                         int* __restrict__ nactive, int* __restrict__ confirm, double tol2, int first);

template <int NQ, int PHASE, int Q0 = 0>
__device__ __forceinline__ void bsys_body(unsigned bid, unsigned nblk, int nsys, const int* __restrict__ tile_first,
                                          const double* __restrict__ part, double* __restrict__ st,
                                          int* __restrict__ active, int* __restrict__ iters,
                                          int* __restrict__ nactive, int* __restrict__ confirm, double tol2,
                                          int first)
{
    const int waves = (int)(blockDim.x >> 6);
    for (int s = (int)bid * waves + (int)(threadIdx.x >> 6); s < nsys; s += (int)nblk * waves)
        bsys_one<NQ, PHASE, Q0>(s, tile_first, part, st, active, iters, nactive, confirm, tol2, first);
}

template <int NQ, int PHASE, int Q0 = 0>
__global__ __launch_bounds__(256) void k_bsys(int nsys, const int* __restrict__ tile_first,
                                              const double* __restrict__ part, double* __restrict__ st,
                                              int* __restrict__ active, int* __restrict__ iters, int* __restrict__ nactive,
                                              int* __restrict__ confirm, double tol2, int first)
{
    bsys_body<NQ, PHASE, Q0>(blockIdx.x, gridDim.x, nsys, tile_first, part, st, active, iters, nactive, confirm, tol2,
                             first);
}

// the work of one wave on system s (every lane of the wave calls it)
template <int NQ, int PHASE, int Q0>
__device__ __forceinline__ void bsys_one(int s, const int* __restrict__ tile_first, const double* __restrict__ part,
                                         double* __restrict__ st, int* __restrict__ active, int* __restrict__ iters,
                                         int* __restrict__ nactive, int* __restrict__ confirm, double tol2, int first)
{
    if (!active[s]) return;
    if (PHASE == 6 && !confirm[s]) return;
    static_assert(Q0 + NQ <= BQ, "k_bsys: slots");
    double d[NQ];
    merge_sys<NQ, Q0>(s, tile_first, part, d);
    if ((threadIdx.x & 63) != 0) return;
    double* S = st + (size_t)s * S_NSTATE;
    if (PHASE == 0) {  // d0 = (r_j, rt)
        double rho = S[S_RHO];
        if (first) rho = -S[S_OMEGA] * rho;
        if (rho == 0.0) {
            active[s] = 0;
            S[S_BETA] = 0.0;
            return;
        }
        const double rho1 = d[0];
        S[S_BETA] = S[S_ALPHA] * rho1 / rho;
        S[S_RHO] = rho1;
    } else if (PHASE == 1) {  // d0 = (u_{j+1}, rt)
        S[S_ALPHA] = S[S_RHO] / d[0];
    } else if (PHASE == 2) {  // d = (r1,r1), (r2,r1), (r2,r2), (r0,r1), (r0,r2)
        const double s1 = d[0];
        const double tau = d[1] / s1;
        const double s2 = d[2] - 2.0 * tau * d[1] + tau * tau * s1;
        const double gp1 = d[3] / s1;
        const double gp2 = (d[4] - tau * d[3]) / s2;
        const double g2 = gp2;
        const double g1 = gp1 - tau * g2;
        S[S_TAU] = tau;
        S[S_GP1] = gp1;
        S[S_GP2] = gp2;
        S[S_G2] = g2;
        S[S_G1] = g1;
        S[S_GPP1] = g2;
        S[S_OMEGA] = g2;
        iters[s] += 2;
    } else if (PHASE == 5) {  // d0 = (r0, r0) of the recursive residual
        S[S_RNORM] = d[0];
        if (d[0] <= tol2 * S[S_BNORM]) confirm[s] = 1;   // decided by the true residual, this cycle
        else if (d[0] == d[0]) atomicAdd(nactive, 1);
        else active[s] = 0;
    } else if (PHASE == 3 || PHASE == 6) {  // d0 = (r0, r0) of the true residual
        if (PHASE == 6) confirm[s] = 0;
        S[S_RNORM] = d[0];
        if (!(d[0] <= tol2 * S[S_BNORM])) {
            if (d[0] == d[0]) atomicAdd(nactive, 1);
            else active[s] = 0;
        } else {
            active[s] = 0;
        }
    } else {  // init: d0 = (b, b), d1 = (r0, r0)
        S[S_BNORM] = d[0];
        S[S_RNORM] = d[1];
        S[S_RHO] = 1.0;
        S[S_ALPHA] = 0.0;
        S[S_OMEGA] = 1.0;
        if (d[1] <= tol2 * d[0]) active[s] = 0;
    }
}

// the scalars of one system's update
struct upd_coef {
    double beta, alpha;
    double g1, gpp1, c1, gp2, g2;   // MR update: c1 = gp2 tau - gp1
};

template <int KIND>
__device__ __forceinline__ upd_coef load_coef(const double* __restrict__ S)
{
    upd_coef k;
    k.beta = S[S_BETA];
    k.alpha = S[S_ALPHA];
    if (KIND == 4) {
        k.g1 = S[S_G1];
        k.gpp1 = S[S_GPP1];
        k.c1 = S[S_GP2] * S[S_TAU] - S[S_GP1];
        k.gp2 = S[S_GP2];
        k.g2 = S[S_G2];
    }
    return k;
}

// vectors an update kind reads
enum { RD_X = 1, RD_R0 = 2, RD_R1 = 4, RD_R2 = 8, RD_U0 = 16, RD_U1 = 32, RD_U2 = 64 };
constexpr int upd_reads(int kind)
{
    return kind == 0 ? RD_U0 | RD_R0
         : kind == 1 ? RD_U1 | RD_R0 | RD_U0 | RD_X
         : kind == 2 ? RD_U0 | RD_R0 | RD_U1 | RD_R1
         : kind == 3 ? RD_U1 | RD_R0 | RD_U2 | RD_R1 | RD_U0 | RD_X
                     : RD_X | RD_R0 | RD_R1 | RD_R2 | RD_U0 | RD_U1 | RD_U2;
}
struct upd_vals {
    double x, r0, r1, r2, u0, u1, u2, rt;
};

constexpr int UPD_ROWS = 4;   // rows per thread per step, all loaded before any is written

template <int KIND, bool DOTS>
__device__ __forceinline__ void upd_load(int64_t i, upd_vals& v, const double* __restrict__ x,
                                         const double* __restrict__ r0, const double* __restrict__ r1,
                                         const double* __restrict__ r2, const double* __restrict__ u0,
                                         const double* __restrict__ u1, const double* __restrict__ u2,
                                         const double* __restrict__ rt)
{
    constexpr int M = upd_reads(KIND);
    if (M & RD_X) v.x = x[i];
    if (M & RD_R0) v.r0 = r0[i];
    if (M & RD_R1) v.r1 = r1[i];
    if (M & RD_R2) v.r2 = r2[i];
    if (M & RD_U0) v.u0 = u0[i];
    if (M & RD_U1) v.u1 = u1[i];
    if (M & RD_U2) v.u2 = u2[i];
    if (DOTS) v.rt = rt[i];
}

// stores row i's update; returns the new r0 (KIND 4)
template <int KIND>
__device__ __forceinline__ double upd_store(int64_t i, const upd_vals& a, const upd_coef& k, double* __restrict__ x,
                                            double* __restrict__ r0, double* __restrict__ r1,
                                            double* __restrict__ u0, double* __restrict__ u1)
{
    double rn = 0.0;
    if (KIND == 0) {
        u0[i] = dev::fma_rn<double>(-k.beta, a.u0, a.r0);
    } else if (KIND == 1) {
        r0[i] = dev::fma_rn<double>(-k.alpha, a.u1, a.r0);
        x[i] = dev::fma_rn<double>(k.alpha, a.u0, a.x);
    } else if (KIND == 2) {
        u0[i] = dev::fma_rn<double>(-k.beta, a.u0, a.r0);
        u1[i] = dev::fma_rn<double>(-k.beta, a.u1, a.r1);
    } else if (KIND == 3) {
        r0[i] = dev::fma_rn<double>(-k.alpha, a.u1, a.r0);
        r1[i] = dev::fma_rn<double>(-k.alpha, a.u2, a.r1);
        x[i] = dev::fma_rn<double>(k.alpha, a.u0, a.x);
    } else {
        x[i] = dev::fma_rn<double>(k.gpp1, a.r1, dev::fma_rn<double>(k.g1, a.r0, a.x));
        rn = dev::fma_rn<double>(k.c1, a.r1, dev::fma_rn<double>(-k.gp2, a.r2, a.r0));
        r0[i] = rn;
        u0[i] = dev::fma_rn<double>(-k.g1, a.u1, dev::fma_rn<double>(-k.g2, a.u2, a.u0));
    }
    return rn;
}

// KIND 0: u0 = r0 - beta u0; 1: r0 -= alpha u1, x += alpha u0; 2: u0 = r0 - beta u0,
// u1 = r1 - beta u1; 3: r0 -= alpha u1, r1 -= alpha u2, x += alpha u0; 4: the MR
// update x += g1 r0 + gpp1 r1, r0 -= gp2 r2 - c1 r1, u0 -= g2 u2 + g1 u1; with DOTS
// (KIND 4) also (r0, r0) and (r0, r~) of the new r0 into (ds, dc)[0] and [1].
// Rows r + u * BT_THREADS, u < UPD_ROWS: when all lie below end their loads are
// issued together before any store, so a wave keeps UPD_ROWS times as many loads
// in flight; the rows of a partial step go one at a time.
template <int KIND, bool DOTS>
__device__ __forceinline__ void upd_step(int64_t r, int64_t end, const upd_coef& k, double* __restrict__ x,
                                         double* __restrict__ r0, double* __restrict__ r1,
                                         const double* __restrict__ r2, double* __restrict__ u0,
                                         double* __restrict__ u1, const double* __restrict__ u2,
                                         const double* __restrict__ rt, double* ds, double* dc)
{
    auto dots = [&](double rn, double rtv) {
        if (DOTS) {
            dev::dot2_step<double>(ds[0], dc[0], rn, rn);
            dev::dot2_step<double>(ds[1], dc[1], rn, rtv);
        }
    };
    if (r + (int64_t)(UPD_ROWS - 1) * BT_THREADS < end) {
        upd_vals v[UPD_ROWS];
#pragma unroll
        for (int u = 0; u < UPD_ROWS; ++u)
            upd_load<KIND, DOTS>(r + (int64_t)u * BT_THREADS, v[u], x, r0, r1, r2, u0, u1, u2, rt);
#pragma unroll
        for (int u = 0; u < UPD_ROWS; ++u)
            dots(upd_store<KIND>(r + (int64_t)u * BT_THREADS, v[u], k, x, r0, r1, u0, u1), v[u].rt);
    } else {
        for (int64_t i = r; i < end; i += BT_THREADS) {
            upd_vals v;
            upd_load<KIND, DOTS>(i, v, x, r0, r1, r2, u0, u1, u2, rt);
            dots(upd_store<KIND>(i, v, k, x, r0, r1, u0, u1), v.rt);
        }
    }
}

// updates; every tile reads its system's scalars.  DOTS (the MR update of a cycle
// on the recursive residual): per-tile partials of (r0, r0) and (r0, r~) of the new
// r0 in slots 0 and 1, as the true-residual SpMV leaves them.
template <int KIND, bool DOTS = false>
__device__ __forceinline__ void bupd_body(unsigned bid, unsigned nblk, const btile* __restrict__ tiles,
                                          const int* __restrict__ list, const int* __restrict__ count,
                                          const int* __restrict__ active, const double* __restrict__ st,
                                          double* __restrict__ x, double* __restrict__ r0, double* __restrict__ r1,
                                          double* __restrict__ r2, double* __restrict__ u0, double* __restrict__ u1,
                                          double* __restrict__ u2, const double* __restrict__ rt,
                                          double* __restrict__ part)
{
    static_assert(!DOTS || KIND == 4, "k_bupd: dot products of the MR update only");
    constexpr int NW = BT_THREADS / 64;
    // wave totals, double-buffered by tile as in k_bspmv: one barrier per tile
    __shared__ double wtot[2][2][NW][2];
    int buf = 0;
    const int nt = *count;
    const dev::xcd_rows Q = dev::xcd_rows::of(nt, 1, nblk, bid);
    for (long q = Q.first; q < Q.end; q += Q.stride) {
        const int ti = list[q];
        const btile t = tiles[ti];
        if (!active[t.sys]) continue;
        const upd_coef k = load_coef<KIND>(st + (size_t)t.sys * S_NSTATE);
        double ds[2] = {0.0, 0.0}, dc[2] = {0.0, 0.0};
        for (int64_t r = t.r0 + threadIdx.x; r < t.r1; r += (int64_t)UPD_ROWS * BT_THREADS)
            upd_step<KIND, DOTS>(r, t.r1, k, x, r0, r1, r2, u0, u1, u2, rt, ds, dc);
        if (DOTS) {
            const int w = threadIdx.x >> 6;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                dev::wave_reduce<double, true>(ds[j], dc[j]);
                if ((threadIdx.x & 63) == 0) {
                    wtot[buf][j][w][0] = ds[j];
                    wtot[buf][j][w][1] = dc[j];
                }
            }
            __syncthreads();
            if (threadIdx.x < 2) {   // waves merged in wave order
                const int j = threadIdx.x;
                double fs = wtot[buf][j][0][0], fc = wtot[buf][j][0][1];
#pragma unroll
                for (int v = 1; v < NW; ++v) dev::merge<double>(fs, fc, wtot[buf][j][v][0], wtot[buf][j][v][1]);
                part[((size_t)ti * BQ + j) * 2] = fs;
                part[((size_t)ti * BQ + j) * 2 + 1] = fc;
            }
            buf ^= 1;
        }
    }
}

template <int KIND, bool DOTS = false>
__global__ __launch_bounds__(BT_THREADS) void k_bupd(const btile* __restrict__ tiles, const int* __restrict__ list,
                                                     const int* __restrict__ count, const int* __restrict__ active,
                                                     const double* __restrict__ st, double* __restrict__ x,
                                                     double* __restrict__ r0, double* __restrict__ r1,
                                                     double* __restrict__ r2, double* __restrict__ u0,
                                                     double* __restrict__ u1, double* __restrict__ u2,
                                                     const double* __restrict__ rt, double* __restrict__ part)
{
    bupd_body<KIND, DOTS>(blockIdx.x, gridDim.x, tiles, list, count, active, st, x, r0, r1, r2, u0, u1, u2, rt,
                          part);
}

// list of the tiles of active systems (one block; systems in order, so the list
// is ordered and the per-system merges see their tiles in tile order)
__device__ __forceinline__ void bcompact_body(int nsys, const int* __restrict__ tile_first,
                                              const int* __restrict__ active, int* __restrict__ list,
                                              int* __restrict__ count)
{
    __shared__ int base;
    __shared__ int scan[256];
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int s0 = 0; s0 < nsys; s0 += 256) {
        const int s = s0 + threadIdx.x;
        const int nt = (s < nsys && active[s]) ? tile_first[s + 1] - tile_first[s] : 0;
        scan[threadIdx.x] = nt;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {   // inclusive Hillis-Steele scan
            const int v = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        const int start = base + scan[threadIdx.x] - nt;
        for (int k = 0; k < nt; ++k) list[start + k] = tile_first[s] + k;
        __syncthreads();
        if (threadIdx.x == 255) base += scan[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = base;
}
__global__ __launch_bounds__(256) void k_bcompact(int nsys, const int* __restrict__ tile_first,
                                                  const int* __restrict__ active, int* __restrict__ list,
                                                  int* __restrict__ count)
{
    bcompact_body(nsys, tile_first, active, list, count);
}

// dinv_i = 1 / a_ii
__global__ __launch_bounds__(256) void k_bdiag(int64_t nrows, const int* __restrict__ rp, const int* __restrict__ col,
                                               const double* __restrict__ val, double* __restrict__ dinv)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * 256) {
        double d = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (col[k] == (int)r) d = val[k];
        dinv[r] = 1.0 / d;
    }
}

// the values of A D: vald_k = val_k * dinv_col(k)
__global__ __launch_bounds__(256) void k_bcolscale(int64_t nrows, const int* __restrict__ rp,
                                                   const int* __restrict__ col, const double* __restrict__ val,
                                                   const double* __restrict__ dinv, double* __restrict__ vald)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * 256)
        for (int k = rp[r]; k < rp[r + 1]; ++k) vald[k] = val[k] * dinv[col[k]];
}

// x := dinv * x (the solution of the preconditioned system back to x)
__global__ __launch_bounds__(256) void k_bscale(int64_t nrows, const double* __restrict__ dinv, double* __restrict__ x)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * 256)
        x[r] = x[r] * dinv[r];
}

__global__ void k_zero_int(int* p) { *p = 0; }

// ---------------------------------------------------------------- one persistent launch
// All cycles in one cooperative launch (hipLaunchCooperativeKernel: every block is
// resident, so blocks can wait for each other): the phases of record_cycle, each
// followed by a grid barrier, until no system is active or the cycle cap.  Same
// phase bodies as the standalone kernels, so the iterates are the same bit for bit.
//
// Grid barrier.  The blocks of one XCD share its L2, and a release at agent scope
// writes that whole L2 back, whichever block issues it; issuing one per block (a
// thousand or more per barrier) costs far more than the phases between barriers.  So
// the blocks are grouped by the XCD they run on (the XCC_ID hardware register, read
// once; the group sizes are counted at the start of the launch): a block waits for
// its own stores to reach L2 and arrives on its XCD's counter with a relaxed
// increment; the last block of an XCD resets that counter and arrives on the top
// counter with an acq_rel increment (one L2 write-back per XCD); the last of those
// advances the generation word.  Waiting polls with relaxed loads and takes one
// acquire fence once the generation has moved (an acquire in the poll loop would
// invalidate the caches on every poll and slow the blocks still at work).  The
// wait is bounded: after 2^24 polls (a second or more) without progress it sets the
// error word, and a set error word releases every wait and ends the cycle loop, so
// the launch ends even if a barrier were missed.
constexpr int PB_XCC = 16;   // XCC_ID is a 4-bit field
struct pbar {
    unsigned group[PB_XCC * 16];   // arrivals per XCD, 64 bytes apart
    unsigned reg[PB_XCC * 16];     // blocks per XCD, counted at the start
    unsigned flat;                 // arrivals of the start-up barrier
    unsigned pad0[15];
    unsigned top;
    unsigned pad1[15];
    unsigned gen;
    unsigned pad2[15];
    int err;
    int cycles;
    int pad3[14];
};
constexpr unsigned PB_SPIN_LIMIT = 1u << 24;

struct pgroup {
    unsigned xcc, members, ngroups;
};

// polls until *w != v (relaxed) or the error word is set; then one acquire fence
__device__ __forceinline__ void pb_wait_ne(pbar* P, const unsigned* w, unsigned v)
{
    unsigned spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == v) {
        if (__hip_atomic_load(&P->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        __builtin_amdgcn_s_sleep(2);
        if (++spins == PB_SPIN_LIMIT) {
            __hip_atomic_store(&P->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Start of the launch: count the blocks of each XCD.  Every block registers with an
// acq_rel increment and waits at a one-off barrier on all blocks (per-block
// releases, once per launch).
__device__ __forceinline__ pgroup pb_register(pbar* P)
{
    __shared__ pgroup g;
    if (threadIdx.x == 0) {
        const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (PB_XCC - 1);   // HW_REG_XCC_ID
        __hip_atomic_fetch_add(&P->reg[xcc * 16], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(&P->flat, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u)
            __hip_atomic_store(&P->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        pb_wait_ne(P, &P->gen, 0u);
        unsigned ng = 0;
        for (int x = 0; x < PB_XCC; ++x)
            ng += __hip_atomic_load(&P->reg[x * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ? 1u : 0u;
        g.xcc = xcc;
        g.members = __hip_atomic_load(&P->reg[xcc * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        g.ngroups = ng;
    }
    __syncthreads();
    return g;
}

__device__ __forceinline__ void grid_sync(pbar* P, const pgroup& G)
{
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have reached the XCD's L2
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned gen = __hip_atomic_load(&P->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(&P->group[G.xcc * 16], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            G.members - 1u) {
            // every block of this XCD has arrived, its stores in this L2: one release
            // writes them back for all
            __hip_atomic_store(&P->group[G.xcc * 16], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(&P->top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == G.ngroups - 1u) {
                __hip_atomic_store(&P->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&P->gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        pb_wait_ne(P, &P->gen, gen);
    }
    __syncthreads();
}

struct pargs {
    const btile* tiles;
    const int* tile_first;
    int* list;
    int* count;
    int* active;
    int* iters;
    int* nactive;
    int* confirm;
    const int* rp;
    const int* col;
    const double* mval;
    double* st;
    double* part;
    double* v[10];
    int nsys;
    int grid;    // blocks of the tile-walking SpMV phases (at most the launch's)
    int ugrid;   // blocks of the update phases
    int max_cycles;
    int recalc;
    double tol2;
    pbar* bar;
};

// Occupancy: left alone the compiler keeps every pointer of every phase live and
// gives the kernel 136 VGPRs (3 blocks per CU); capped at 6 waves per SIMD (80
// VGPRs, a few cold spills) the 512-system benchmark solve takes 21.95 s against
// 26.97 s at 3 and 22.94 s at 8.
__global__ __launch_bounds__(BT_THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_bpersistent(pargs A)
{
    const unsigned P = gridDim.x, b = blockIdx.x;
    const pgroup PG = pb_register(A.bar);
    __shared__ bstage S;   // one staging buffer for the three SpMV phases
    const unsigned Gs = (unsigned)A.grid < P ? (unsigned)A.grid : P;
    const unsigned Gu = (unsigned)A.ugrid < P ? (unsigned)A.ugrid : P;
    double* const* v = A.v;
    const double* none = nullptr;
    auto sys = [&](auto nq, auto phase, auto q0, int first) {
        bsys_body<decltype(nq)::value, decltype(phase)::value, decltype(q0)::value>(
            b, P, A.nsys, A.tile_first, A.part, A.st, A.active, A.iters, A.nactive, A.confirm, A.tol2, first);
        grid_sync(A.bar, PG);
    };
    auto upd = [&](auto kind) {
        if (b < Gu)
            bupd_body<decltype(kind)::value>(b, Gu, A.tiles, A.list, A.count, A.active, A.st, v[VX], v[VR0], v[VR1],
                                             v[VR2], v[VU0], v[VU1], v[VU2], none, (double*)nullptr);
        grid_sync(A.bar, PG);
    };
    auto spmv_rt = [&](const double* x, double* y) {   // y = A D x and (y, r~)
        if (b < Gs)
            bspmv_body<0, 1>(S, b, Gs, A.tiles, A.list, A.count, A.active, A.rp, A.col, A.mval, x, none, y,
                             (const double*)v[VRT], none, A.part, (const int*)nullptr);
        grid_sync(A.bar, PG);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    using I6 = std::integral_constant<int, 6>;
    int c = 1;
    for (; c <= A.max_cycles; ++c) {
        const bool recalc = A.recalc > 0 && c % A.recalc == 0;
        if (b == 0 && threadIdx.x == 0) __hip_atomic_store(A.nactive, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sys(I1{}, I0{}, I1{}, 1);        // rho1 = (r0, r~), beta
        upd(I0{});                       // u0 = r0 - beta u0
        spmv_rt(v[VU0], v[VU1]);         // u1 = A D u0, (u1, r~)
        sys(I1{}, I1{}, I0{}, 0);        // alpha
        upd(I1{});                       // r0 -= alpha u1, x += alpha u0
        spmv_rt(v[VR0], v[VR1]);         // r1 = A D r0, (r1, r~)
        sys(I1{}, I0{}, I0{}, 0);
        upd(I2{});                       // u0, u1
        spmv_rt(v[VU1], v[VU2]);         // u2 = A D u1, (u2, r~)
        sys(I1{}, I1{}, I0{}, 0);
        upd(I3{});                       // r0, r1, x
        if (b < Gs)                      // r2 = A D r1 and the Gram entries of r0, r1, r2
            bspmv_body<0, 5>(S, b, Gs, A.tiles, A.list, A.count, A.active, A.rp, A.col, A.mval, (const double*)v[VR1],
                             none, v[VR2], none, (const double*)v[VR0], A.part, (const int*)nullptr);
        grid_sync(A.bar, PG);
        sys(I5{}, I2{}, I0{}, 0);
        auto true_resid = [&](const int* only) {
            if (b < Gs)
                bspmv_body<1, 2>(S, b, Gs, A.tiles, A.list, A.count, A.active, A.rp, A.col, A.mval,
                                 (const double*)v[VX], (const double*)v[VB], v[VR0], (const double*)v[VRT], none,
                                 A.part, only);
            grid_sync(A.bar, PG);
        };
        if (recalc) {
            upd(I4{});
            true_resid(nullptr);
            sys(I1{}, I3{}, I0{}, 0);
        } else {
            if (b < Gu)
                bupd_body<4, true>(b, Gu, A.tiles, A.list, A.count, A.active, A.st, v[VX], v[VR0], v[VR1], v[VR2],
                                   v[VU0], v[VU1], v[VU2], (const double*)v[VRT], A.part);
            grid_sync(A.bar, PG);
            sys(I1{}, I5{}, I0{}, 0);
            true_resid(A.confirm);
            sys(I1{}, I6{}, I0{}, 0);
        }
        if (b == 0) bcompact_body(A.nsys, A.tile_first, A.active, A.list, A.count);
        grid_sync(A.bar, PG);
        const int na = __hip_atomic_load(A.nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int err = __hip_atomic_load(&A.bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        grid_sync(A.bar, PG);   // every block has read the count before the next cycle clears it
        if (na == 0 || err) break;
    }
    if (b == 0 && threadIdx.x == 0) A.bar->cycles = c <= A.max_cycles ? c : A.max_cycles;
}

// ---------------------------------------------------------------- host side

static int free_batch(batch_t* B)
{
    if (!B) return HPB_OK;
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->exec2) (void)hipGraphExecDestroy(B->exec2);
    if (B->graph2) (void)hipGraphDestroy(B->graph2);
    void* ptrs[] = {B->d_tiles, B->d_tile_first, B->d_row_off, B->rp, B->col, B->val, B->st, B->active,
                    B->iters, B->nactive, B->part, B->dots, B->tile_list, B->tile_count, B->dinv, B->vald,
                    B->confirm, B->bar};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (double* p : B->v)
        if (p) (void)hipFree(p);
    if (B->hcount) (void)hipHostFree(B->hcount);
    delete B;
    return HPB_OK;
}

static unsigned sys_blocks(int nsys) { return (unsigned)((nsys * 64 + 255) / 256); }

// record one BiCGStab(2) cycle on the batch's stream.  Every dot product of the
// cycle is an epilogue of the kernel that makes its new vector; the cycle's first
// (r0, r~) comes from slot 1 of the previous cycle's last pass (of the set-up pass,
// where r~ = r0, on the first cycle).  recalc: the cycle ends with the true
// residual b - A x replacing r0 and the convergence check on it; otherwise the
// check runs on the recursive r0 of the MR update, and only the systems that meet
// the tolerance get the true residual (which replaces r0) and the final check.
static void record_cycle(batch_t* B, double tol2, bool recalc)
{
    hipStream_t st = B->h->stream;
    const unsigned G = (unsigned)B->grid, GU = (unsigned)B->ugrid, SB = sys_blocks(B->nsys);
    double** v = B->v;
    const double* mval = B->jacobi ? B->vald : B->val;   // A D or A
    const double* none = nullptr;
    auto spmv_rt = [&](const double* x, double* y) {   // y = A D x and (y, r~)
        hipLaunchKernelGGL((k_bspmv<0, 1>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list,
                           B->tile_count, B->active, B->rp, B->col, mval, x, none, y, (const double*)v[VRT],
                           none, B->part, (const int*)nullptr);
    };
    auto upd = [&](auto kind) {
        hipLaunchKernelGGL((k_bupd<decltype(kind)::value>), dim3(GU), dim3(BT_THREADS), 0, st, B->d_tiles,
                           B->tile_list, B->tile_count, B->active, B->st, v[VX], v[VR0], v[VR1], v[VR2], v[VU0],
                           v[VU1], v[VU2], none, (double*)nullptr);
    };
    auto sys = [&](auto nq, auto phase, auto q0, int first) {
        hipLaunchKernelGGL((k_bsys<decltype(nq)::value, decltype(phase)::value, decltype(q0)::value>), dim3(SB),
                           dim3(256), 0, st, B->nsys, B->d_tile_first, B->part, B->st, B->active, B->iters,
                           B->nactive, B->confirm, tol2, first);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    using I5 = std::integral_constant<int, 5>;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, B->nactive);
    // j = 0
    sys(I1{}, I0{}, I1{}, 1);        // rho1 = (r0, r~), beta
    upd(I0{});                       // u0 = r0 - beta u0
    spmv_rt(v[VU0], v[VU1]);         // u1 = A D u0, (u1, r~)
    sys(I1{}, I1{}, I0{}, 0);        // alpha
    upd(I1{});                       // r0 -= alpha u1, x += alpha u0
    spmv_rt(v[VR0], v[VR1]);         // r1 = A D r0, (r1, r~)
    // j = 1
    sys(I1{}, I0{}, I0{}, 0);
    upd(I2{});                       // u0, u1
    spmv_rt(v[VU1], v[VU2]);         // u2 = A D u1, (u2, r~)
    sys(I1{}, I1{}, I0{}, 0);
    upd(I3{});                       // r0, r1, x
    // MR part: r2 = A D r1 and the Gram entries of r0, r1, r2
    hipLaunchKernelGGL((k_bspmv<0, 5>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list, B->tile_count,
                       B->active, B->rp, B->col, mval, (const double*)v[VR1], none, v[VR2], none,
                       (const double*)v[VR0], B->part, (const int*)nullptr);
    sys(I5{}, I2{}, I0{}, 0);
    // the true residual r0 = b - A D x with (r0, r0) and (r0, r~), for all systems
    // or for those marked in confirm
    auto true_resid = [&](const int* only) {
        hipLaunchKernelGGL((k_bspmv<1, 2>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list,
                           B->tile_count, B->active, B->rp, B->col, mval, (const double*)v[VX], (const double*)v[VB],
                           v[VR0], (const double*)v[VRT], none, B->part, only);
    };
    if (recalc) {
        upd(I4{});
        true_resid(nullptr);
        sys(I1{}, I3{}, I0{}, 0);
    } else {
        hipLaunchKernelGGL((k_bupd<4, true>), dim3(GU), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list,
                           B->tile_count, B->active, B->st, v[VX], v[VR0], v[VR1], v[VR2], v[VU0], v[VU1], v[VU2],
                           (const double*)v[VRT], B->part);
        sys(I1{}, std::integral_constant<int, 5>{}, I0{}, 0);
        true_resid(B->confirm);
        sys(I1{}, std::integral_constant<int, 6>{}, I0{}, 0);
    }
    // next cycle only visits the tiles of systems still active
    hipLaunchKernelGGL(k_bcompact, dim3(1), dim3(256), 0, st, B->nsys, B->d_tile_first, B->active, B->tile_list,
                       B->tile_count);
}

}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_batch_create_diffusion2d(hpb_handle_t h, int nsys, const int* sides, double sigma, uint64_t seed,
                                 uint64_t rhs_seed, hpb_batch_t* out)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && sides && out && nsys >= 1, HPB_ERR_INVALID, "hpb_batch_create_diffusion2d: bad argument");
    *out = nullptr;
    batch_t* B = new batch_t();
    B->h = h;
    B->nsys = nsys;
    B->row_off.assign(nsys + 1, 0);
    std::vector<int64_t> nnz_off(nsys + 1, 0);
    std::vector<btile> tiles;
    B->tile_first.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        const int64_t m = sides[s];
        if (m < 2 || m > 46340) {
            free_batch(B);
            HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_batch_create_diffusion2d: side length must be in [2, 46340]");
        }
        B->row_off[s + 1] = B->row_off[s] + m * m;
        nnz_off[s + 1] = nnz_off[s] + m * m + 4 * m * (m - 1);
        B->tile_first[s] = (int)tiles.size();
        for (int64_t r = B->row_off[s]; r < B->row_off[s + 1]; r += BT_ROWS)
            tiles.push_back(btile{s, 0, r, std::min<int64_t>(r + BT_ROWS, B->row_off[s + 1])});
    }
    B->tile_first[nsys] = (int)tiles.size();
    B->nrows = B->row_off[nsys];
    B->nnz = nnz_off[nsys];
    B->ntiles = (int)tiles.size();
    if (B->nnz >= (1ll << 31)) {
        free_batch(B);
        HPB_REQUIRE(false, HPB_ERR_INVALID, "hpb_batch_create_diffusion2d: more than 2^31-1 non-zeros");
    }
    hipError_t e = hipSuccess;
    auto M = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    M((void**)&B->d_tiles, sizeof(btile) * B->ntiles);
    M((void**)&B->d_tile_first, sizeof(int) * (nsys + 1));
    M((void**)&B->d_row_off, sizeof(int64_t) * (nsys + 1));
    M((void**)&B->rp, sizeof(int) * (B->nrows + 1));
    M((void**)&B->col, sizeof(int) * B->nnz);
    M((void**)&B->val, sizeof(double) * B->nnz);
    for (int k = 0; k < 10; ++k) M((void**)&B->v[k], sizeof(double) * B->nrows);
    M((void**)&B->st, sizeof(double) * S_NSTATE * nsys);
    M((void**)&B->active, sizeof(int) * nsys);
    M((void**)&B->iters, sizeof(int) * nsys);
    M((void**)&B->nactive, sizeof(int));
    M((void**)&B->part, sizeof(double) * 2 * BQ * B->ntiles);
    M((void**)&B->dots, sizeof(double) * BQ * nsys);
    M((void**)&B->tile_list, sizeof(int) * B->ntiles);
    M((void**)&B->tile_count, sizeof(int));
    M((void**)&B->dinv, sizeof(double) * B->nrows);
    M((void**)&B->confirm, sizeof(int) * nsys);
    M((void**)&B->bar, sizeof(pbar));
    if (e == hipSuccess) e = hipHostMalloc((void**)&B->hcount, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_batch(B);
        return hpb::hip_status(e, __FILE__, __LINE__, "hpb_batch_create_diffusion2d: allocation");
    }
    int64_t* d_nnz_off = nullptr;
    int* d_side = nullptr;
    HPB_HIP(hipMalloc((void**)&d_nnz_off, sizeof(int64_t) * (nsys + 1)));
    HPB_HIP(hipMalloc((void**)&d_side, sizeof(int) * nsys));
    hipStream_t st = h->stream;
    HPB_HIP(hipMemcpyAsync(B->d_tiles, tiles.data(), sizeof(btile) * B->ntiles, hipMemcpyHostToDevice, st));
    HPB_HIP(hipMemcpyAsync(B->d_tile_first, B->tile_first.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, st));
    HPB_HIP(hipMemcpyAsync(B->d_row_off, B->row_off.data(), sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice, st));
    HPB_HIP(hipMemcpyAsync(d_nnz_off, nnz_off.data(), sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice, st));
    HPB_HIP(hipMemcpyAsync(d_side, sides, sizeof(int) * nsys, hipMemcpyHostToDevice, st));
    int64_t g = (B->nrows + 255) / 256;
    const int64_t cap = (int64_t)(h->num_cu > 0 ? h->num_cu : 256) * 16;
    if (g > cap) g = cap;
    hipLaunchKernelGGL(k_batch_build, dim3((unsigned)g), dim3(256), 0, st, nsys, B->nrows, B->d_row_off, d_nnz_off,
                       d_side, sigma, seed, B->rp, B->col, B->val);
    HPB_LAUNCH_CHECK();
    // x* uniform in [-1, 1] (Philox, rhs_seed), b = A x*
    HPB_HIP(hipStreamSynchronize(st));
    (void)hipFree(d_nnz_off);
    (void)hipFree(d_side);
    int rc = hpb_random_uniform_f64(h, (size_t)B->nrows, rhs_seed, 0, -1.0, 1.0, B->v[VXS]);
    if (rc) {
        free_batch(B);
        return rc;
    }
    rc = hpb_csr_spmvc_f64(h, (int)B->nrows, (int)B->nrows, B->nnz, B->rp, B->col, B->val, 1.0, B->v[VXS], 0.0,
                           B->v[VB]);
    if (rc) {
        free_batch(B);
        return rc;
    }
    {
        int64_t g2 = (B->nrows + 255) / 256;
        if (g2 > cap) g2 = cap;
        hipLaunchKernelGGL(k_bdiag, dim3((unsigned)g2), dim3(256), 0, st, B->nrows, B->rp, B->col, B->val, B->dinv);
        HPB_LAUNCH_CHECK();
    }
    {
        const int cu = h->num_cu > 0 ? h->num_cu : 256;
        // measured on 512 systems (186.6M rows, 91 368 tiles): the row-product kernels
        // run best with one tile per block (whole solve 18.59 s, against 18.71 s at 128
        // blocks per CU, 18.92 s at 32 and 19.86 s at 8; with the nontemporal staging);
        // the streaming updates with 1 block per CU (18.61 s, against 19.11 s at 4,
        // 19.15 s at 32 and 18.72 s with one tile per block)
        B->grid = (int)std::min<int64_t>(B->ntiles, (int64_t)512 * cu);
        B->ugrid = std::min(B->ntiles, cu);
    }
    HPB_HIP(hipStreamSynchronize(st));
    *out = (hpb_batch_t)B;
    return HPB_OK;
}

int hpb_batch_info(hpb_batch_t b, int64_t* rows, int64_t* nnz, int* ntiles)
{
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_info: null batch");
    batch_t* B = (batch_t*)b;
    if (rows) *rows = B->nrows;
    if (nnz) *nnz = B->nnz;
    if (ntiles) *ntiles = B->ntiles;
    return HPB_OK;
}

int hpb_batch_solve_bicgstab2(hpb_batch_t b, double rel_tol, int max_iters, int use_graph, int check_every,
                              int* out_iters, double* out_true_rel_resid, int* out_converged, double* out_seconds,
                              int* out_cycles)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(b && rel_tol > 0 && max_iters >= 2 && check_every >= 1 && use_graph >= 0 && use_graph <= 2,
                HPB_ERR_INVALID, "hpb_batch_solve_bicgstab2: bad argument");
    batch_t* B = (batch_t*)b;
    hpb_handle_t h = B->h;
    hipStream_t st = h->stream;
    const double tol2 = rel_tol * rel_tol;
    double** v = B->v;
    const size_t nb = sizeof(double) * B->nrows;
    const unsigned G = (unsigned)B->grid;
    auto all_active = [&]() -> int {
        std::vector<int> ones(B->nsys, 1);
        HPB_HIP(hipMemcpyAsync(B->active, ones.data(), sizeof(int) * B->nsys, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bcompact, dim3(1), dim3(256), 0, st, B->nsys, B->d_tile_first, B->active, B->tile_list,
                           B->tile_count);
        HPB_LAUNCH_CHECK();
        HPB_HIP(hipStreamSynchronize(st));
        return HPB_OK;
    };
    // x = 0 (x^ = D^-1 x with Jacobi), r0 = b, rt = r0 (unnormalised: scaling r~
    // rescales rho and (u, r~) alike), u0 = 0
    HPB_HIP(hipMemsetAsync(v[VX], 0, nb, st));
    HPB_HIP(hipMemsetAsync(v[VU0], 0, nb, st));
    HPB_HIP(hipMemsetAsync(v[VU1], 0, nb, st));
    HPB_HIP(hipMemsetAsync(v[VU2], 0, nb, st));
    HPB_HIP(hipMemcpyAsync(v[VR0], v[VB], nb, hipMemcpyDeviceToDevice, st));
    HPB_HIP(hipMemcpyAsync(v[VRT], v[VB], nb, hipMemcpyDeviceToDevice, st));
    HPB_HIP(hipMemsetAsync(B->iters, 0, sizeof(int) * B->nsys, st));
    if (int rc = all_active()) return rc;
    {
        // slot 0: (b, b); slot 1: (r0, r0), which is also the (r0, r~) the first
        // cycle starts from
        dot_pairs P{};
        P.a[0] = v[VB];
        P.b[0] = v[VB];
        P.a[1] = v[VR0];
        P.b[1] = v[VR0];
        hipLaunchKernelGGL((k_bdots<2>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list, B->tile_count,
                           B->active, P, B->part);
        hipLaunchKernelGGL((k_bsys<2, 4>), dim3(sys_blocks(B->nsys)), dim3(256), 0, st, B->nsys, B->d_tile_first,
                           B->part, B->st, B->active, B->iters, B->nactive, B->confirm, tol2, 0);
        hipLaunchKernelGGL(k_bcompact, dim3(1), dim3(256), 0, st, B->nsys, B->d_tile_first, B->active, B->tile_list,
                           B->tile_count);
        HPB_LAUNCH_CHECK();
    }
    HPB_HIP(hipMemsetAsync(B->confirm, 0, sizeof(int) * B->nsys, st));
    // the captured cycles bake in the tolerance and the preconditioner choice
    if (B->graph_tol2 != tol2 || B->graph_jacobi != B->jacobi) {
        for (int k = 0; k < 2; ++k) {
            hipGraphExec_t& e = k ? B->exec2 : B->exec;
            hipGraph_t& g = k ? B->graph2 : B->graph;
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
            e = nullptr;
            g = nullptr;
        }
    }
    // cycle c (from 1) recomputes the true residual when recalc divides c
    const int recalc = B->recalc;
    auto is_recalc = [recalc](int c) { return recalc > 0 && c % recalc == 0; };
    if (use_graph == 1) {
        for (int k = 0; k < 2; ++k) {
            const bool rc = k == 0;
            if (rc && recalc == 0) continue;        // never used
            if (!rc && recalc == 1) continue;
            hipGraphExec_t& e = rc ? B->exec : B->exec2;
            hipGraph_t& g = rc ? B->graph : B->graph2;
            if (e) continue;
            HPB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            record_cycle(B, tol2, rc);
            hipError_t ce = hipStreamEndCapture(st, &g);
            if (ce != hipSuccess) return hpb::hip_status(ce, __FILE__, __LINE__, "hipStreamEndCapture");
            HPB_HIP(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
        }
        B->graph_tol2 = tol2;
        B->graph_jacobi = B->jacobi;
    }
    HPB_HIP(hipStreamSynchronize(st));
    auto t0 = std::chrono::steady_clock::now();
    int cycles = 0;
    const int max_cycles = max_iters / 2;
    if (use_graph == 2) {
        // every cycle in one cooperative launch of as many blocks as can be resident
        pargs A{};
        A.tiles = B->d_tiles;
        A.tile_first = B->d_tile_first;
        A.list = B->tile_list;
        A.count = B->tile_count;
        A.active = B->active;
        A.iters = B->iters;
        A.nactive = B->nactive;
        A.confirm = B->confirm;
        A.rp = B->rp;
        A.col = B->col;
        A.mval = B->jacobi ? B->vald : B->val;
        A.st = B->st;
        A.part = B->part;
        for (int k = 0; k < 10; ++k) A.v[k] = v[k];
        A.nsys = B->nsys;
        A.grid = B->grid;
        A.ugrid = B->ugrid;
        A.max_cycles = max_cycles;
        A.recalc = recalc;
        A.tol2 = tol2;
        A.bar = B->bar;
        int per_cu = 0;
        HPB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bpersistent, BT_THREADS, 0));
        HPB_REQUIRE(per_cu >= 1, HPB_ERR_UNSUPPORTED, "hpb_batch_solve_bicgstab2: persistent kernel does not fit a CU");
        const int cu = h->num_cu > 0 ? h->num_cu : 256;
        const int P = std::min(per_cu, 8) * cu;
        HPB_HIP(hipMemsetAsync(B->bar, 0, sizeof(pbar), st));
        void* args[] = {&A};
        HPB_HIP(hipLaunchCooperativeKernel((const void*)k_bpersistent, dim3((unsigned)P), dim3(BT_THREADS), args, 0,
                                           st));
        pbar hb;
        HPB_HIP(hipMemcpyAsync(&hb, B->bar, sizeof(pbar), hipMemcpyDeviceToHost, st));
        HPB_HIP(hipStreamSynchronize(st));
        HPB_REQUIRE(hb.err == 0, HPB_ERR_UNSUPPORTED,
                    "hpb_batch_solve_bicgstab2: the persistent launch's grid barrier made no progress for seconds");
        cycles = hb.cycles;
        if (B->verbose)
            std::fprintf(stderr, "hpb_batch: one persistent launch of %d blocks, %d cycles, %.3f s\n", P, cycles,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    while (use_graph != 2 && cycles < max_cycles) {
        const int k = std::min(check_every, max_cycles - cycles);
        for (int i = 0; i < k; ++i) {
            const bool rc = is_recalc(cycles + i + 1);
            if (use_graph == 1) HPB_HIP(hipGraphLaunch(rc ? B->exec : B->exec2, st));
            else record_cycle(B, tol2, rc);
        }
        HPB_LAUNCH_CHECK();
        cycles += k;
        HPB_HIP(hipMemcpyAsync(B->hcount, B->nactive, sizeof(int), hipMemcpyDeviceToHost, st));
        HPB_HIP(hipStreamSynchronize(st));
        if (B->verbose)
            std::fprintf(stderr, "hpb_batch: cycle %d, active systems %d, %.3f s\n", cycles, *B->hcount,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        if (*B->hcount == 0) break;
    }
    // back from the preconditioned unknowns: x = D^-1 x^
    if (B->jacobi) {
        const int cu = h->num_cu > 0 ? h->num_cu : 256;
        int64_t g = (B->nrows + 255) / 256;
        if (g > (int64_t)cu * 16) g = (int64_t)cu * 16;
        hipLaunchKernelGGL(k_bscale, dim3((unsigned)g), dim3(256), 0, st, B->nrows, B->dinv, v[VX]);
        HPB_LAUNCH_CHECK();
    }
    HPB_HIP(hipStreamSynchronize(st));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // per-system true relative residuals ||b - A x|| / ||b|| of the original systems, compensated
    if (int rc = all_active()) return rc;
    hipLaunchKernelGGL((k_bspmv<1, 0>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list, B->tile_count,
                       B->active, B->rp, B->col, B->val, (const double*)v[VX], (const double*)v[VB], v[VR1],
                       (const double*)nullptr, (const double*)nullptr, B->part, (const int*)nullptr);
    {
        dot_pairs P{};
        P.a[0] = v[VR1];
        P.b[0] = v[VR1];
        P.a[1] = v[VB];
        P.b[1] = v[VB];
        hipLaunchKernelGGL((k_bdots<2>), dim3(G), dim3(BT_THREADS), 0, st, B->d_tiles, B->tile_list, B->tile_count,
                           B->active, P, B->part);
    }
    HPB_LAUNCH_CHECK();
    std::vector<double> part((size_t)B->ntiles * BQ * 2);
    std::vector<int> it(B->nsys);
    HPB_HIP(hipMemcpyAsync(part.data(), B->part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    HPB_HIP(hipMemcpyAsync(it.data(), B->iters, sizeof(int) * B->nsys, hipMemcpyDeviceToHost, st));
    HPB_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < B->nsys; ++s) {
        double rr = 0.0, bb = 0.0;
        for (int t = B->tile_first[s]; t < B->tile_first[s + 1]; ++t) {
            rr += part[((size_t)t * BQ + 0) * 2] + part[((size_t)t * BQ + 0) * 2 + 1];
            bb += part[((size_t)t * BQ + 1) * 2] + part[((size_t)t * BQ + 1) * 2 + 1];
        }
        const double rel = std::sqrt(rr) / std::sqrt(bb);
        if (out_true_rel_resid) out_true_rel_resid[s] = rel;
        if (out_iters) out_iters[s] = it[s];
        if (out_converged) out_converged[s] = rel <= rel_tol ? 1 : 0;
    }
    if (out_seconds) *out_seconds = secs;
    if (out_cycles) *out_cycles = cycles;
    return HPB_OK;
}

int hpb_batch_set_jacobi(hpb_batch_t b, int on)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_set_jacobi: null batch");
    batch_t* B = (batch_t*)b;
    if (on && !B->vald) {   // the values of A D, once
        HPB_HIP(hipMalloc((void**)&B->vald, sizeof(double) * B->nnz));
        hipStream_t st = B->h->stream;
        const int64_t cap = (int64_t)(B->h->num_cu > 0 ? B->h->num_cu : 256) * 16;
        const int64_t g = std::min<int64_t>((B->nrows + 255) / 256, cap);
        hipLaunchKernelGGL(k_bcolscale, dim3((unsigned)g), dim3(256), 0, st, B->nrows, B->rp, B->col, B->val,
                           B->dinv, B->vald);
        HPB_LAUNCH_CHECK();
        HPB_HIP(hipStreamSynchronize(st));
    }
    B->jacobi = on ? 1 : 0;
    return HPB_OK;
}

int hpb_batch_set_resid_recalc_freq(hpb_batch_t b, int freq)
{
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_set_resid_recalc_freq: null batch");
    HPB_REQUIRE(freq >= 0, HPB_ERR_INVALID, "hpb_batch_set_resid_recalc_freq: freq must be >= 0");
    ((batch_t*)b)->recalc = freq;
    return HPB_OK;
}

int hpb_batch_set_verbose(hpb_batch_t b, int verbose)
{
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_set_verbose: null batch");
    ((batch_t*)b)->verbose = verbose;
    return HPB_OK;
}

int hpb_batch_solution(hpb_batch_t b, double* host_x, double* host_xstar)
{
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_solution: null batch");
    batch_t* B = (batch_t*)b;
    hipStream_t st = B->h->stream;
    if (host_x) HPB_HIP(hipMemcpyAsync(host_x, B->v[VX], sizeof(double) * B->nrows, hipMemcpyDeviceToHost, st));
    if (host_xstar) HPB_HIP(hipMemcpyAsync(host_xstar, B->v[VXS], sizeof(double) * B->nrows, hipMemcpyDeviceToHost, st));
    HPB_HIP(hipStreamSynchronize(st));
    return HPB_OK;
}

int hpb_batch_csr(hpb_batch_t b, const int** row_ptr, const int** col_ind, const double** val, const double** rhs)
{
    HPB_REQUIRE(b, HPB_ERR_INVALID, "hpb_batch_csr: null batch");
    batch_t* B = (batch_t*)b;
    if (row_ptr) *row_ptr = B->rp;
    if (col_ind) *col_ind = B->col;
    if (val) *val = B->val;
    if (rhs) *rhs = B->v[VB];
    return HPB_OK;
}

int hpb_batch_destroy(hpb_batch_t b)
{
    if (b) {
        batch_t* B = (batch_t*)b;
        if (B->h) (void)hipStreamSynchronize(B->h->stream);
        free_batch(B);
    }
    return HPB_OK;
}

}  // extern "C"
