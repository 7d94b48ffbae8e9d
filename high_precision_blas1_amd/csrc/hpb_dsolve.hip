// hpb_dsolve.hip -- device-resident GMRES(m) and BiCGStab(L).
//
// Same algorithms and monitor rules as the C++ solver templates (which follow
// source/numerical_algos/lin_solvers/gmres.h:376-494 and bicgstabl.h:151-274), with
// every scalar of the iteration kept in device memory, so that no step waits for
// the host:
//   * dots, norms and multi-dots write (s, c) pairs to device memory; small
//     single-thread kernels turn them into the recurrence scalars in FP64 (for
//     FP32 vectors too: the Hessenberg matrix, the Givens rotations, the back
//     substitution and BiCGStab's L x L minimal-residual solve are FP64);
//   * GMRES: one restart cycle (restart residual, m Arnoldi steps with CGS2 --
//     two fused multi-dot / multi-update passes --, back substitution, x += V y) is
//     captured once as a hipGraph and replayed; the host reads one flag per cycle;
//   * BiCGStab(L): one cycle (L BiCG steps, the MR part, the optional true
//     residual, the norm check) per graph; the host reads the flag every few
//     cycles;
//   * convergence, breakdown and the iteration cap raise device flags; every
//     kernel of a graph tests its flag on entry, so replays after the end do no
//     work and leave x untouched.
// MR part of BiCGStab(L): the normal equations of min ||r_0 - sum_j g_j r_j||,
// G g = c with G_ij = (r_i, r_j), c_i = (r_i, r_0) (1 <= i, j <= L), are formed by
// fused compensated multi-dots and solved by Cholesky in FP64; in exact arithmetic
// this is the reference's modified Gram-Schmidt solution (bicgstabl.h:224-245),
// and x, r_0, u_0 are then updated in one pass.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

namespace hpb {
namespace {

constexpr int DB = 256;        // threads per block of the vector kernels
constexpr int MAXB = 256;      // GMRES basis vectors (m + 1 <= 256)
constexpr int MAXL = 8;        // BiCGStab(L) basis in the device path

struct dstate {
    double tol, bnorm, rnorm;
    double rlast;   // reliable updates: FP64 true residual norm at the last update
    int restart;    // reliable updates: the FP32 iteration restarts from the new residual
    double rho, alpha, omega;
    int iters, max_iters, min_iters, cycles;
    int done, stop, converged, breakdown, invalid;
    int cycle_len, hist_len, hist_cap, recalc_freq, skip_recalc;
    int early_exit, basis;
    // reliable updates: the FP32 residual is the true one / rscale, and x64 gains
    // xscale * x32 (both powers of two: 1 in every other solve)
    double rscale, xscale;
    double wn2;   // GMRES selective reorthogonalisation: ||w||^2 before the projection
    int skip1;    // ... 1 when the second Gram-Schmidt pass is not needed (or the cycle stopped)
};

template <typename T>
__device__ __forceinline__ double dval(const T* o4)
{
    return (double)o4[0] + (double)o4[1];
}

__device__ __forceinline__ bool finite(double v) { return v == v && v - v == 0.0; }

__device__ void record(dstate* st, double* hist, double rn)
{
    if (hist && st->hist_len < st->hist_cap) {
        hist[2 * st->hist_len] = (double)st->iters;
        hist[2 * st->hist_len + 1] = st->bnorm > 0.0 ? rn / st->bnorm : rn;
        st->hist_len++;
    }
}

// monitor check (default_monitor.h:200-227): converged when the residual is a
// valid number, <= tol and at least min_iters steps ran; finished also at the cap
__device__ void monitor_check(dstate* st, double* hist, double rn)
{
    st->rnorm = rn;
    record(st, hist, rn);
    if (!finite(rn)) {
        st->invalid = 1;
        st->done = 1;
        return;
    }
    if (rn <= st->tol && st->iters >= st->min_iters) {
        st->converged = 1;
        st->done = 1;
        return;
    }
    if (st->iters >= st->max_iters) st->done = 1;
}

// ---------------------------------------------------------------- common kernels

template <typename T>
__global__ void k_init(dstate* st, const T* nb4, double rel_tol, double abs_tol)
{
    const double bn = sqrt(dval(nb4));
    st->bnorm = bn;
    st->tol = abs_tol + rel_tol * bn;
}

// v := s v, s read from device memory
template <typename T>
__global__ __launch_bounds__(DB) void k_scale(size_t n, T* __restrict__ v, const T* __restrict__ s,
                                              const int* __restrict__ skip)
{
    if (*skip) return;
    const T a = *s;
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) v[i] = a * v[i];
}

// y := s x
template <typename T>
__global__ __launch_bounds__(DB) void k_scale_copy(size_t n, const T* __restrict__ x, T* __restrict__ y,
                                                   const T* __restrict__ s, const int* __restrict__ skip)
{
    if (*skip) return;
    const T a = *s;
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) y[i] = a * x[i];
}

template <typename T>
__global__ __launch_bounds__(DB) void k_fill(size_t n, T* __restrict__ v, T a)
{
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) v[i] = a;
}

// max |x_i|: non-negative doubles order like their bit patterns, so blocks fold
// their maxima with one integer atomic max (order-independent)
__global__ __launch_bounds__(DB) void k_absmax(size_t n, const double* __restrict__ x,
                                               unsigned long long* __restrict__ out)
{
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) m = fmax(m, fabs(x[i]));
    for (int k = 32; k >= 1; k >>= 1) m = fmax(m, __shfl_xor(m, k, 64));
    __shared__ double wm[DB / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < DB / 64; ++w) m = fmax(m, wm[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// ---------------------------------------------------------------- GMRES kernels

// restart: beta = ||r||, monitor check on the true residual, g = beta e_1
template <typename T>
__global__ void k_gm_start(dstate* st, const T* n4, double* g, int m, T* sig, double* hist)
{
    if (st->done) {
        st->stop = 1;
        return;
    }
    const double beta = sqrt(dval(n4));
    monitor_check(st, hist, beta);
    if (st->done) {
        st->stop = 1;
        return;
    }
    st->stop = 0;
    st->cycle_len = 0;
    st->cycles++;
    g[0] = beta;
    for (int i = 1; i <= m; ++i) g[i] = 0.0;
    sig[0] = (T)(1.0 / beta);
}

__device__ __forceinline__ void apply_rot(double& dx, double& dy, double cs, double sn)
{
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

// New Hessenberg column j with h_{j+1,j} = hn: previous rotations, new rotation,
// residual estimate (gmres.h:217-274); sig[j+1] = 1 / hn scales the stored basis
// vector (V holds unnormalised vectors; the basis is sig_i V_i).
template <typename T>
__device__ void givens_step(int j, int m, double hn, double* H, int ldh, double* cs, double* sn, double* g, T* sig,
                            dstate* st)
{
    double* Hj = H + (size_t)j * ldh;
    Hj[j + 1] = hn;
    for (int k = 0; k < j; ++k) apply_rot(Hj[k], Hj[k + 1], cs[k], sn[k]);
    double c, s;
    const double dx = Hj[j], dy = Hj[j + 1];
    if (dy == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (fabs(dy) > fabs(dx)) {
        const double t = dx / dy;
        s = 1.0 / sqrt(1.0 + t * t);
        c = t * s;
    } else {
        const double t = dy / dx;
        c = 1.0 / sqrt(1.0 + t * t);
        s = t * c;
    }
    cs[j] = c;
    sn[j] = s;
    apply_rot(Hj[j], Hj[j + 1], c, s);
    Hj[j + 1] = 0.0;
    apply_rot(g[j], g[j + 1], c, s);
    st->iters++;
    st->cycle_len = j + 1;
    const double est = fabs(g[j + 1]);
    st->rnorm = est;
    sig[j + 1] = hn != 0.0 ? (T)(1.0 / hn) : T(0);
    if (hn == 0.0 || !finite(est) || (st->early_exit && est <= st->tol) || j + 1 >= m || st->iters >= st->max_iters)
        st->stop = 1;
}

template <typename T>
struct basis_pack {
    const T* p[MAXB];
};

// Reductions over many blocks end in per-block totals stored plainly in `part`;
// a one-block kernel launched next merges them in a fixed order (block by block
// per thread, then threads in order) and runs the scalar step.  A last-block
// ticket would need an agent-scope release in every block -- an L2 write-back on
// this multi-XCD part, about 20 ns serialised per block -- which costs more than
// the extra launch once a grid has more than a few hundred blocks.
template <typename T, bool COMP>
__device__ __forceinline__ void part_out(T s, T c, T* __restrict__ part, size_t slot)
{
    dev::block_reduce<T, COMP, DB>(s, c);
    if (threadIdx.x == 0) {
        part[2 * slot] = s;
        part[2 * slot + 1] = c;
    }
}

// One block of DB threads: columns 0..NQ-1 of G partials each merged (thread t
// takes partials t, t + DB, ... of every column -- independent loads in flight --,
// then the waves' butterflies and the wave totals in order); out valid in thread 0.
template <typename T, bool COMP, int NQ>
__device__ __forceinline__ void merge_columns(const T* __restrict__ part, unsigned G, double (&out)[NQ])
{
    T fs[NQ], fc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) fs[q] = fc[q] = T(0);
    for (unsigned b = threadIdx.x; b < G; b += DB) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const T ps = part[2 * ((size_t)q * G + b)];
            const T pc = part[2 * ((size_t)q * G + b) + 1];
            if (COMP) dev::merge<T>(fs[q], fc[q], ps, pc);
            else fs[q] = fs[q] + ps;
        }
    }
    constexpr int NW = DB / 64;
    __shared__ T ls[NW][NQ], lc[NW][NQ];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        dev::wave_reduce<T, COMP>(fs[q], fc[q]);
        if (lane == 0) {
            ls[wid][q] = fs[q];
            lc[wid][q] = fc[q];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            T s = ls[0][q], c = lc[0][q];
            for (int w = 1; w < NW; ++w) {
                if (COMP) dev::merge<T>(s, c, ls[w][q], lc[w][q]);
                else s = s + ls[w][q];
            }
            out[q] = (double)s + (double)c;
        }
    }
}

// One Gram-Schmidt projection pass for column j: per-block partials of
// (V_i, w), i <= j, compensated or plain, KC columns per blockIdx.y chunk;
// k_gm_hfin merges them.  NT: the basis columns are loaded nontemporally (chosen
// when the basis is larger than the 256 MB MALL, so it cannot stay there between
// passes anyway).
template <typename T, int KC, bool COMP, bool NT>
__global__ __launch_bounds__(DB) void k_gm_dot(size_t n, int k, basis_pack<T> V, const T* __restrict__ w,
                                               T* __restrict__ part, const int* __restrict__ stop)
{
    if (*stop) return;
    using Vt = typename dev::vec16<T>::type;   // the vectors come from hipMalloc: 16-byte aligned
    constexpr int W = dev::vec16<T>::W;
    const unsigned G = gridDim.x;
    const int j0 = blockIdx.y * KC;
    const int kc = (k - j0) < KC ? (k - j0) : KC;
    T s[KC], c[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) s[q] = c[q] = T(0);
    const size_t nv = n / W;
    const Vt* wv = reinterpret_cast<const Vt*>(w);
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < nv; i += (size_t)G * DB) {
        const Vt a = wv[i];
#pragma unroll
        for (int q = 0; q < KC; ++q)
            if (q < kc) {
                const Vt* col = reinterpret_cast<const Vt*>(V.p[j0 + q]);
                const Vt v = NT ? dev::load_nt<Vt>(col + i) : col[i];
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    if (COMP) dev::dot2_step<T>(s[q], c[q], dev::vget<T>(v, e), dev::vget<T>(a, e));
                    else s[q] = dev::fma_rn<T>(dev::vget<T>(v, e), dev::vget<T>(a, e), s[q]);
                }
            }
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {   // the < W trailing elements
        const size_t i = nv * W + threadIdx.x;
#pragma unroll
        for (int q = 0; q < KC; ++q)
            if (q < kc) {
                if (COMP) dev::dot2_step<T>(s[q], c[q], V.p[j0 + q][i], w[i]);
                else s[q] = dev::fma_rn<T>(V.p[j0 + q][i], w[i], s[q]);
            }
    }
#pragma unroll
    for (int q = 0; q < KC; ++q)
        if (q < kc) part_out<T, COMP>(s[q], c[q], part, (size_t)(j0 + q) * G + blockIdx.x);
}

// h_i = sig_i (V_i, w) from G partials per column: H(i, j) (pass 0) or += (pass 1),
// and coef_i = -h_i sig_i for the update.  One block, a wave per column.
// With kk = k + 1 the last column is (w, w) (pass 0 of the selective
// reorthogonalisation), kept unscaled in st->wn2.
template <typename T, bool COMP>
__global__ __launch_bounds__(1024) void k_gm_hfin(const T* __restrict__ part, unsigned G, int k, int kk,
                                                  const T* __restrict__ sig, double* __restrict__ Hj,
                                                  T* __restrict__ coef, int pass, const int* __restrict__ stop,
                                                  dstate* __restrict__ st)
{
    if (*stop) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int col = wid; col < kk; col += 1024 / 64) {
        T fs = T(0), fc = T(0);
        for (unsigned b = lane; b < G; b += 64) {
            const T ps = part[2 * ((size_t)col * G + b)];
            const T pc = part[2 * ((size_t)col * G + b) + 1];
            if (COMP) dev::merge<T>(fs, fc, ps, pc);
            else fs = fs + ps;
        }
        dev::wave_reduce<T, COMP>(fs, fc);
        if (lane == 0) {
            if (col == k) {
                st->wn2 = (double)fs + (double)fc;
                continue;
            }
            const double sg = (double)sig[col];
            const double h = ((double)fs + (double)fc) * sg;
            if (pass == 0) Hj[col] = h;
            else Hj[col] += h;
            coef[col] = (T)(-h * sg);
        }
    }
}

// w += sum_{i<k} coef_i V_i (Dot2 over the terms of each element when COMP);
// NORM: also per-block partials of ||w||^2 (compensated) for k_gm_nfin.  NT as in
// k_gm_dot.
template <typename T, bool COMP, bool NORM, bool NT>
__global__ __launch_bounds__(DB) void k_gm_upd(size_t n, int k, basis_pack<T> V, const T* __restrict__ coef,
                                               T* __restrict__ w, T* __restrict__ part,
                                               const int* __restrict__ skip)
{
    if (*skip) return;
    using Vt = typename dev::vec16<T>::type;
    constexpr int W = dev::vec16<T>::W;
    __shared__ T sc[MAXB];
    for (int q = threadIdx.x; q < k; q += DB) sc[q] = coef[q];
    __syncthreads();
    T ns = T(0), nc = T(0);
    const size_t nv = n / W;
    Vt* wv = reinterpret_cast<Vt*>(w);
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < nv; i += (size_t)gridDim.x * DB) {
        Vt a = wv[i];
        T s[W], c[W];
#pragma unroll
        for (int e = 0; e < W; ++e) {
            s[e] = dev::vget<T>(a, e);
            c[e] = T(0);
        }
        for (int q = 0; q < k; ++q) {
            const Vt* col = reinterpret_cast<const Vt*>(V.p[q]);
            const Vt v = NT ? dev::load_nt<Vt>(col + i) : col[i];
            const T cq = sc[q];
#pragma unroll
            for (int e = 0; e < W; ++e) {
                if (COMP) dev::dot2_step<T>(s[e], c[e], cq, dev::vget<T>(v, e));
                else s[e] = dev::fma_rn<T>(cq, dev::vget<T>(v, e), s[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < W; ++e) {
            const T r = COMP ? s[e] + c[e] : s[e];
            reinterpret_cast<T*>(&a)[e] = r;
            if (NORM) {
                if (COMP) dev::dot2_step<T>(ns, nc, r, r);
                else ns = dev::fma_rn<T>(r, r, ns);
            }
        }
        wv[i] = a;
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {   // the < W trailing elements
        const size_t i = nv * W + threadIdx.x;
        T s = w[i], c = T(0);
        for (int q = 0; q < k; ++q) {
            if (COMP) dev::dot2_step<T>(s, c, sc[q], V.p[q][i]);
            else s = dev::fma_rn<T>(sc[q], V.p[q][i], s);
        }
        const T r = COMP ? s + c : s;
        w[i] = r;
        if (NORM) {
            if (COMP) dev::dot2_step<T>(ns, nc, r, r);
            else ns = dev::fma_rn<T>(r, r, ns);
        }
    }
    if (NORM) part_out<T, COMP>(ns, nc, part, blockIdx.x);
}

// ||w|| from G partials, then the Givens step for column j with h_{j+1,j} = ||w||
template <typename T, bool COMP>
__global__ __launch_bounds__(DB) void k_gm_nfin(const T* __restrict__ part, unsigned G, int j, int m,
                                                double* __restrict__ H, int ldh, double* __restrict__ cs,
                                                double* __restrict__ sn, double* __restrict__ g, T* __restrict__ sig,
                                                dstate* __restrict__ st, const int* __restrict__ skip)
{
    if (*skip) return;
    double nn[1];
    merge_columns<T, COMP, 1>(part, G, nn);
    if (threadIdx.x == 0) givens_step<T>(j, m, sqrt(nn[0]), H, ldh, cs, sn, g, sig, st);
}

// Selective reorthogonalisation after the first classical pass (Daniel, Gragg,
// Kaufman and Stewart 1976): w' = w - V h is accepted when ||w'|| >= ||w|| / sqrt 2,
// and the Givens step runs with h_{j+1,j} = ||w'||; otherwise the second pass runs
// (skip1 = 0) and k_gm_nfin takes the step after it.
template <typename T, bool COMP>
__global__ __launch_bounds__(DB) void k_gm_dgks(const T* __restrict__ part, unsigned G, int j, int m,
                                                double* __restrict__ H, int ldh, double* __restrict__ cs,
                                                double* __restrict__ sn, double* __restrict__ g, T* __restrict__ sig,
                                                dstate* __restrict__ st)
{
    if (st->stop) {
        if (threadIdx.x == 0) st->skip1 = 1;
        return;
    }
    double nn[1];
    merge_columns<T, COMP, 1>(part, G, nn);
    if (threadIdx.x != 0) return;
    if (nn[0] >= 0.5 * st->wn2) {
        st->skip1 = 1;
        givens_step<T>(j, m, sqrt(nn[0]), H, ldh, cs, sn, g, sig, st);
    } else {
        st->skip1 = 0;
    }
}

// y = R^{-1} g over the steps of this cycle (gmres.h:204-215); y[k..m) = 0
template <typename T>
__global__ void k_gm_backsub(int m, const double* H, int ldh, double* g, const T* sig, T* y, const dstate* st)
{
    if (st->done) return;
    const int k = st->cycle_len;
    for (int j = k - 1; j >= 0; --j) {
        g[j] /= H[(size_t)j * ldh + j];
        for (int i = j - 1; i >= 0; --i) g[i] -= H[(size_t)j * ldh + i] * g[j];
    }
    // x += sum_j g_j (sig_j V_j)
    for (int j = 0; j < m; ++j) y[j] = j < k ? (T)(g[j] * (double)sig[j]) : T(0);
}

// x += sum_{j < cycle_len} y_j V_j (compensated: Dot2 over the terms of each element)
template <typename T, bool COMP>
__global__ __launch_bounds__(DB) void k_gm_xupd(size_t n, basis_pack<T> V, const T* __restrict__ y,
                                                T* __restrict__ x, const dstate* __restrict__ st)
{
    if (st->done) return;
    const int k = st->cycle_len;
    __shared__ T sy[MAXB];
    for (int j = threadIdx.x; j < k; j += DB) sy[j] = y[j];
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        T s = x[i], c = T(0);
        for (int j = 0; j < k; ++j) {
            if (COMP) dev::dot2_step<T>(s, c, sy[j], V.p[j][i]);
            else s = dev::fma_rn<T>(sy[j], V.p[j][i], s);
        }
        x[i] = COMP ? s + c : s;
    }
}

// ---------------------------------------------------------------- BiCGStab(L) kernels

template <typename T>
struct bs_pack {
    T* r[MAXL + 1];
    T* u[MAXL + 1];
};

template <typename T>
__global__ void k_bs_init(dstate* st, const T* n4, T* scal, double* hist)
{
    const double rn = sqrt(dval(n4));
    st->rho = 1.0;
    st->alpha = 0.0;
    st->omega = 1.0;
    monitor_check(st, hist, rn);
    *scal = rn != 0.0 ? (T)(1.0 / rn) : T(1);
}

// rho1 = (r_j, rt); beta = alpha rho1 / rho (bicgstabl.h:199-206)
__device__ __forceinline__ void breakdown_stop(dstate* st)
{
    st->breakdown = 1;
    st->done = 1;
    st->stop = 1;
}

template <typename T>
__global__ void k_bs_beta(dstate* st, const T* d4, T* coef, int first, const int* skip)
{
    if (*skip) return;
    double rho = st->rho;
    if (first) rho = -st->omega * rho;
    if (rho == 0.0 || !finite(rho)) {
        breakdown_stop(st);
        return;
    }
    const double rho1 = dval(d4);
    const double beta = st->alpha * rho1 / rho;
    st->rho = rho1;
    coef[0] = (T)beta;
}

// alpha = rho / (u_{j+1}, rt)
template <typename T>
__global__ void k_bs_alpha(dstate* st, const T* d4, T* coef, const int* skip)
{
    if (*skip) return;
    const double sg = dval(d4);
    if (sg == 0.0 || !finite(sg)) {
        breakdown_stop(st);
        return;
    }
    st->alpha = st->rho / sg;
    coef[1] = (T)st->alpha;
}

// u_i = r_i - beta u_i, i = 0..j
template <typename T>
__global__ __launch_bounds__(DB) void k_bs_upd_u(size_t n, int j, bs_pack<T> P, const T* __restrict__ coef,
                                                 const int* __restrict__ skip)
{
    if (*skip) return;
    const T beta = coef[0];
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB)
        for (int q = 0; q <= j; ++q) P.u[q][i] = dev::fma_rn<T>(-beta, P.u[q][i], P.r[q][i]);
}

// r_i -= alpha u_{i+1}, i = 0..j;  x += alpha u_0
template <typename T>
__global__ __launch_bounds__(DB) void k_bs_upd_r(size_t n, int j, bs_pack<T> P, T* __restrict__ x,
                                                 const T* __restrict__ coef, const int* __restrict__ skip)
{
    if (*skip) return;
    const T alpha = coef[1];
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        for (int q = 0; q <= j; ++q) P.r[q][i] = dev::fma_rn<T>(-alpha, P.u[q + 1][i], P.r[q][i]);
        x[i] = dev::fma_rn<T>(alpha, P.u[0][i], x[i]);
    }
}

// MR part: gram[k] holds (r_i, r_k), i = 0..k, as 4-word (s, c, ...) entries.
// Solves G g = c by Cholesky in FP64; coef[2 + j] = g_{j+1}.
template <typename T>
__global__ void k_bs_mr(dstate* st, int L, const T* gram, int gstride, T* coef, const int* skip)
{
    if (*skip) return;
    double G[MAXL][MAXL], c[MAXL], y[MAXL];
    for (int k = 1; k <= L; ++k) {
        const T* gk = gram + (size_t)k * gstride;   // entries (r_i, r_k), i = 0..k
        c[k - 1] = dval(gk);
        for (int i = 1; i <= k; ++i) {
            const double v = dval(gk + 4 * i);
            G[i - 1][k - 1] = v;
            G[k - 1][i - 1] = v;
        }
    }
    // Cholesky G = R^T R (upper R stored in G)
    for (int j = 0; j < L; ++j) {
        double d = G[j][j];
        for (int k = 0; k < j; ++k) d -= G[k][j] * G[k][j];
        if (!(d > 0.0) || !finite(d)) {
            breakdown_stop(st);
            return;
        }
        d = sqrt(d);
        G[j][j] = d;
        for (int i = j + 1; i < L; ++i) {
            double v = G[j][i];
            for (int k = 0; k < j; ++k) v -= G[k][j] * G[k][i];
            G[j][i] = v / d;
        }
    }
    for (int i = 0; i < L; ++i) {   // R^T y = c
        double v = c[i];
        for (int k = 0; k < i; ++k) v -= G[k][i] * y[k];
        y[i] = v / G[i][i];
    }
    for (int i = L - 1; i >= 0; --i) {   // R g = y
        double v = y[i];
        for (int k = i + 1; k < L; ++k) v -= G[i][k] * y[k];
        y[i] = v / G[i][i];
    }
    st->omega = y[L - 1];
    for (int j = 0; j < L; ++j) coef[2 + j] = (T)y[j];
}

// x += sum_j g_j r_{j-1};  r_0 -= sum_j g_j r_j;  u_0 -= sum_j g_j u_j  (j = 1..L)
template <typename T, bool COMP>
__global__ __launch_bounds__(DB) void k_bs_upd_mr(size_t n, int L, bs_pack<T> P, T* __restrict__ x,
                                                  const T* __restrict__ coef, const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ T gm[MAXL];
    if (threadIdx.x < (unsigned)L) gm[threadIdx.x] = coef[2 + threadIdx.x];
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        T sx = x[i], cx = T(0), sr = P.r[0][i], cr = T(0), su = P.u[0][i], cu = T(0);
        T rprev = sr;
        for (int j = 1; j <= L; ++j) {
            const T g = gm[j - 1];
            const T rj = P.r[j][i];
            const T uj = P.u[j][i];
            if (COMP) {
                dev::dot2_step<T>(sx, cx, g, rprev);
                dev::dot2_step<T>(sr, cr, -g, rj);
                dev::dot2_step<T>(su, cu, -g, uj);
            } else {
                sx = dev::fma_rn<T>(g, rprev, sx);
                sr = dev::fma_rn<T>(-g, rj, sr);
                su = dev::fma_rn<T>(-g, uj, su);
            }
            rprev = rj;
        }
        x[i] = COMP ? sx + cx : sx;
        P.r[0][i] = COMP ? sr + cr : sr;
        P.u[0][i] = COMP ? su + cu : su;
    }
}

// true residual every recalc_freq cycles (bicgstabl.h:247-263)
__global__ void k_bs_recalc_flag(dstate* st)
{
    st->skip_recalc = 1;
    if (st->done) return;
    const int f = st->recalc_freq;
    if (f > 0 && ((st->iters / st->basis + 1) % f) == 0) st->skip_recalc = 0;
}

template <typename T>
__global__ void k_bs_check(dstate* st, const T* n4, double* hist)
{
    if (st->done) return;
    st->iters += st->basis;
    st->cycles++;
    monitor_check(st, hist, sqrt(dval(n4)));
}

// ---- fused BiCGStab(L) cycle: dots ride on the kernels that produce their operands,
// a one-block kernel merges them and runs the scalar step

// scalar steps of bicgstabl.h:199-210 on merged dot products
template <typename T>
__device__ void epi_alpha(dstate* st, double sg, T* coef)
{
    if (sg == 0.0 || !finite(sg)) {
        breakdown_stop(st);
        return;
    }
    st->alpha = st->rho / sg;
    coef[1] = (T)st->alpha;
}
template <typename T>
__device__ void epi_beta(dstate* st, double rho1, int first, T* coef)
{
    double rho = st->rho;
    if (first) rho = -st->omega * rho;
    if (rho == 0.0 || !finite(rho)) {
        breakdown_stop(st);
        return;
    }
    coef[0] = (T)(st->alpha * rho1 / rho);
    st->rho = rho1;
}

enum { EPI_NONE = 0, EPI_ALPHA, EPI_BETA, EPI_CHECK, EPI_REL };

// end of a cycle: d0 = (r_0, r~), d1 = (r_0, r_0)
template <typename T>
__device__ void epi_cycle(int kind, dstate* st, double d0, double d1, T* coef, double* hist, double delta)
{
    st->iters += st->basis;
    st->cycles++;
    const double rn = sqrt(d1) * st->rscale;
    if (kind == EPI_CHECK) {
        monitor_check(st, hist, rn);
        if (st->done) {
            st->stop = 1;
            return;
        }
    } else {   // reliable updates: pause for an FP64 update, or continue
        st->rnorm = rn;
        if (!finite(rn)) {
            st->invalid = 1;
            st->done = 1;
            st->stop = 1;
            return;
        }
        if (rn <= delta * st->rlast || rn <= st->tol || st->iters >= st->max_iters) {
            st->stop = 1;
            return;
        }
    }
    epi_beta<T>(st, d0, 1, coef);
}

// number of dot products a fused kernel hands to k_ds_fin for step EPI
template <int EPI> struct epi_terms {
    static constexpr int value = EPI == EPI_NONE ? 0 : (EPI == EPI_ALPHA || EPI == EPI_BETA ? 1 : 2);
};

// y = A x (MODE 0) or y = b - A x (MODE 1), LPR lanes per row (Dot2 rows when
// CROWS), with partials of the dot products of y that the cycle needs next
// (k_ds_fin then runs the step):
//   EPI_ALPHA  (y, r~) -> alpha;  EPI_BETA (y, r~) -> beta of the next BiCG step;
//   EPI_CHECK / EPI_REL (y, r~), (y, y) -> monitor check and beta of the next cycle.

template <typename T, int LPR, bool CROWS, bool COMP, int MODE, int EPI>
__global__ __launch_bounds__(DB) void k_ds_spmv(int rows, const int* __restrict__ rp, const int* __restrict__ col,
                                                const T* __restrict__ val, const T* __restrict__ x,
                                                const T* __restrict__ b, T* __restrict__ y, const T* __restrict__ rt,
                                                T* __restrict__ part, const int* __restrict__ skip)
{
    if (*skip) return;
    constexpr int NQ = epi_terms<EPI>::value;
    T ds[NQ > 0 ? NQ : 1], dc[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : 1); ++q) ds[q] = dc[q] = T(0);
    const int lane = threadIdx.x % LPR;
    // one row: its lanes accumulate (staged: from the tile in LDS), merge, and lane 0
    // stores y_r and takes the dot products
    auto row = [&](long r, int beg, int end, const T* sv, const T* sx) {
        const T rtr = (NQ >= 1 && lane == 0) ? rt[r] : T(0);   // issued with the row bounds
        T s = T(0), c = T(0);
        if (MODE == 1 && lane == 0) s = b[r];
        if (sv != nullptr) dev::csr_row_acc_staged<T, LPR, CROWS, MODE == 1>(beg, end, lane, sv, sx, s, c);
        else dev::csr_row_acc<T, LPR, CROWS, MODE == 1>(beg, end, lane, col, val, x, s, c);
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1) {
            const T s2 = __shfl_xor(s, m, 64);
            if (CROWS) {
                const T c2 = __shfl_xor(c, m, 64);
                dev::merge<T>(s, c, s2, c2);
            } else {
                s = s + s2;
            }
        }
        if (lane == 0) {
            const T v = CROWS ? s + c : s;
            y[r] = v;
            if constexpr (NQ >= 1) {
                if (COMP) dev::dot2_step<T>(ds[0], dc[0], v, rtr);
                else ds[0] = dev::fma_rn<T>(v, rtr, ds[0]);
            }
            if constexpr (NQ >= 2) {
                if (COMP) dev::dot2_step<T>(ds[1], dc[1], v, v);
                else ds[1] = dev::fma_rn<T>(v, v, ds[1]);
            }
        }
    };
    if constexpr (LPR <= 4) {   // staged tiles of 128 rows, up to 8 entries per row on average
        constexpr int RT = 128, CAP = 1024;
        __shared__ dev::csr_stage<T, RT, CAP> S;
        const dev::xcd_rows R = dev::xcd_rows::of(rows, RT);
        for (long r0 = R.first; r0 < R.end; r0 += R.stride) {
            const int nr = (int)((R.end - r0) < RT ? (R.end - r0) : RT);
            const bool staged = dev::csr_stage_tile<T, RT, CAP, DB>(S, r0, nr, rp, col, val, x);
            const int e0 = staged ? S.rp[0] : 0;
            for (int rr = threadIdx.x / LPR; rr < nr; rr += DB / LPR)
                row(r0 + rr, S.rp[rr] - e0, S.rp[rr + 1] - e0, staged ? S.v : nullptr, S.x);
            __syncthreads();   // the next tile overwrites S
        }
    } else {
        const dev::xcd_rows R = dev::xcd_rows::of(rows, DB / LPR);
        for (long r = R.first + (long)threadIdx.x / LPR; r < R.end; r += R.stride)
            row(r, rp[r], rp[r + 1], nullptr, nullptr);
    }
    if constexpr (NQ > 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part_out<T, COMP>(ds[q], dc[q], part, (size_t)q * gridDim.x + blockIdx.x);
    }
}

// merged dot products of a fused kernel (G partials per term), then its scalar step
template <typename T, bool COMP, int EPI>
__global__ __launch_bounds__(DB) void k_ds_fin(const T* __restrict__ part, unsigned G, dstate* __restrict__ st,
                                               T* __restrict__ coef, double* hist, double delta,
                                               const int* __restrict__ skip)
{
    if (*skip) return;
    constexpr int NQ = epi_terms<EPI>::value;
    double out[NQ];
    merge_columns<T, COMP, NQ>(part, G, out);
    if (threadIdx.x != 0) return;
    if (EPI == EPI_ALPHA) epi_alpha<T>(st, out[0], coef);
    else if (EPI == EPI_BETA) epi_beta<T>(st, out[0], 0, coef);
    else epi_cycle<T>(EPI, st, out[0], out[NQ - 1], coef, hist, delta);
}

// MR part in one pass: partials of (r_0, r_k) and (r_i, r_k), 1 <= i <= k <= L;
// k_ds_gfin solves the normal equations by Cholesky in FP64 (as k_bs_mr); L <= 4.
template <typename T, int L, bool COMP>
__global__ __launch_bounds__(DB) void k_ds_gram(size_t n, bs_pack<T> P, T* __restrict__ part,
                                                const int* __restrict__ skip)
{
    if (*skip) return;
    constexpr int NP = L + L * (L + 1) / 2;
    T s[NP], c[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) s[q] = c[q] = T(0);
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        T r[L + 1];
#pragma unroll
        for (int k = 0; k <= L; ++k) r[k] = P.r[k][i];
        int q = 0;
#pragma unroll
        for (int k = 1; k <= L; ++k) {
#pragma unroll
            for (int i2 = 0; i2 <= k; ++i2) {
                if (COMP) dev::dot2_step<T>(s[q], c[q], r[i2], r[k]);
                else s[q] = dev::fma_rn<T>(r[i2], r[k], s[q]);
                ++q;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) part_out<T, COMP>(s[q], c[q], part, (size_t)q * gridDim.x + blockIdx.x);
}

// the Gram matrix and right-hand side from Gp partials per entry, then Cholesky
template <typename T, int L, bool COMP>
__global__ __launch_bounds__(DB) void k_ds_gfin(const T* __restrict__ part, unsigned Gp, dstate* __restrict__ st,
                                                T* __restrict__ coef, const int* __restrict__ skip)
{
    if (*skip) return;
    constexpr int NP = L + L * (L + 1) / 2;
    double out[NP];
    merge_columns<T, COMP, NP>(part, Gp, out);
    if (threadIdx.x != 0) return;
    double G[L][L], cv[L], yv[L];
    int q = 0;
    for (int k = 1; k <= L; ++k)
        for (int i2 = 0; i2 <= k; ++i2) {
            const double v = out[q++];
            if (i2 == 0) cv[k - 1] = v;
            else {
                G[i2 - 1][k - 1] = v;
                G[k - 1][i2 - 1] = v;
            }
        }
    for (int j = 0; j < L; ++j) {
        double d = G[j][j];
        for (int k = 0; k < j; ++k) d -= G[k][j] * G[k][j];
        if (!(d > 0.0) || !finite(d)) {
            breakdown_stop(st);
            return;
        }
        d = sqrt(d);
        G[j][j] = d;
        for (int i2 = j + 1; i2 < L; ++i2) {
            double v = G[j][i2];
            for (int k = 0; k < j; ++k) v -= G[k][j] * G[k][i2];
            G[j][i2] = v / d;
        }
    }
    for (int i2 = 0; i2 < L; ++i2) {
        double v = cv[i2];
        for (int k = 0; k < i2; ++k) v -= G[k][i2] * yv[k];
        yv[i2] = v / G[i2][i2];
    }
    for (int i2 = L - 1; i2 >= 0; --i2) {
        double v = yv[i2];
        for (int k = i2 + 1; k < L; ++k) v -= G[i2][k] * yv[k];
        yv[i2] = v / G[i2][i2];
    }
    st->omega = yv[L - 1];
    for (int j = 0; j < L; ++j) coef[2 + j] = (T)yv[j];
}

// MR update of x, r_0, u_0 (as k_bs_upd_mr) with partials of (r_0, r~), (r_0, r_0)
// of the new r_0; k_ds_fin runs the end-of-cycle step (monitor check or
// reliable-update trigger, beta)
template <typename T, bool COMP>
__global__ __launch_bounds__(DB) void k_ds_upd_mr(size_t n, int L, bs_pack<T> P, T* __restrict__ x,
                                                  const T* __restrict__ rt, const T* __restrict__ coef,
                                                  T* __restrict__ part, const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ T gm[MAXL];
    if (threadIdx.x < (unsigned)L) gm[threadIdx.x] = coef[2 + threadIdx.x];
    __syncthreads();
    T ds[2] = {T(0), T(0)}, dc[2] = {T(0), T(0)};
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        T sx = x[i], cx = T(0), sr = P.r[0][i], cr = T(0), su = P.u[0][i], cu = T(0);
        T rprev = sr;
        for (int j = 1; j <= L; ++j) {
            const T g = gm[j - 1];
            const T rj = P.r[j][i];
            const T uj = P.u[j][i];
            if (COMP) {
                dev::dot2_step<T>(sx, cx, g, rprev);
                dev::dot2_step<T>(sr, cr, -g, rj);
                dev::dot2_step<T>(su, cu, -g, uj);
            } else {
                sx = dev::fma_rn<T>(g, rprev, sx);
                sr = dev::fma_rn<T>(-g, rj, sr);
                su = dev::fma_rn<T>(-g, uj, su);
            }
            rprev = rj;
        }
        const T r0 = COMP ? sr + cr : sr;
        x[i] = COMP ? sx + cx : sx;
        P.r[0][i] = r0;
        P.u[0][i] = COMP ? su + cu : su;
        if (COMP) {
            dev::dot2_step<T>(ds[0], dc[0], r0, rt[i]);
            dev::dot2_step<T>(ds[1], dc[1], r0, r0);
        } else {
            ds[0] = dev::fma_rn<T>(r0, rt[i], ds[0]);
            ds[1] = dev::fma_rn<T>(r0, r0, ds[1]);
        }
    }
    part_out<T, COMP>(ds[0], dc[0], part, blockIdx.x);
    part_out<T, COMP>(ds[1], dc[1], part, (size_t)gridDim.x + blockIdx.x);
}

// ---- reliable updates (FP32 iterations, FP64 solution and true residual)

// after a cycle: the FP32 recursive residual decides whether the FP64 solution is
// updated now (reduction by delta since the last update, tolerance reached, or the
// iteration cap); stop pauses the cycle graph until the host ran the update
template <typename T>
__global__ void k_rel_check(dstate* st, const T* n4, double delta)
{
    if (st->stop) return;
    st->iters += st->basis;
    st->cycles++;
    const double rn = sqrt(dval(n4)) * st->rscale;
    st->rnorm = rn;
    if (!finite(rn)) {
        st->invalid = 1;
        st->done = 1;
        st->stop = 1;
        return;
    }
    if (rn <= delta * st->rlast || rn <= st->tol || st->iters >= st->max_iters) st->stop = 1;
}

// x64 += xscale x32, x32 = 0
__global__ __launch_bounds__(DB) void k_rel_flush(size_t n, float* __restrict__ x32, double* __restrict__ x64,
                                                  const dstate* __restrict__ st, const int* __restrict__ skip)
{
    if (*skip) return;
    const double g = st->xscale;
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        x64[i] += g * (double)x32[i];
        x32[i] = 0.0f;
    }
}

// the FP64 true residual norm: monitor check, new reference for delta, resume.
// When the update did not lower the true residual the FP32 recurrences have lost
// their accuracy: the iteration restarts from the new residual (r~, u_0, rho,
// alpha, omega as at the start).
// On a (re)start the FP32 residual scale becomes the power of two at or below ||r||:
// the FP32 iteration then runs on the same significands with exponents near 0,
// and with the matrix scaled to |a_ij| < 1 (mscale) A^L r cannot overflow.
__global__ void k_rel_after(dstate* st, const double* n4, double* hist, int first, double mscale)
{
    if (st->done) return;
    const double rn = sqrt(dval(n4));
    if (first && rn > 0.0 && finite(rn)) {
        st->rscale = ldexp(1.0, ilogb(rn));
        st->xscale = st->rscale * mscale;
    }
    st->restart = first || !(rn < st->rlast);
    st->rlast = rn;
    monitor_check(st, hist, rn);
    if (st->restart) {
        st->rho = 1.0;
        st->alpha = 0.0;
        st->omega = 1.0;
    }
    st->stop = st->done;
}

// a breakdown of the FP32 iteration (rho = 0, or a Gram matrix that is not
// numerically positive definite) in a reliable-update solve: clear the stop flags
// so that an FP64 update and a restart from the true residual follow
__global__ void k_rel_resume(dstate* st)
{
    st->breakdown = 0;
    st->done = 0;
    st->stop = 0;
}

// r32 = (float) r64; on a restart also r~ = r32 / ||r32|| and u_0 = 0
__global__ __launch_bounds__(DB) void k_rel_convert(size_t n, const double* __restrict__ r64, float* __restrict__ r32,
                                                    float* __restrict__ rt, float* __restrict__ u0,
                                                    const dstate* __restrict__ st)
{
    if (st->done) return;
    const double inv = st->rlast > 0.0 ? 1.0 / st->rlast : 1.0;
    const double rs = 1.0 / st->rscale;   // a power of two: exact
    const bool restart = st->restart != 0;
    for (size_t i = (size_t)blockIdx.x * DB + threadIdx.x; i < n; i += (size_t)gridDim.x * DB) {
        const float v = (float)(r64[i] * rs);
        r32[i] = v;
        if (restart) {
            rt[i] = (float)(r64[i] * inv);
            u0[i] = 0.0f;
        }
    }
}

// ---------------------------------------------------------------- host side

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
struct dbuf {
    T* p = nullptr;
    dbuf() = default;
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    dbuf(dbuf&& o) noexcept : p(o.p) { o.p = nullptr; }
    ~dbuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count)
    {
        const hipError_t e = hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
        if (e != hipSuccess) return hip_status(e, __FILE__, __LINE__, "device solver workspace");
        return HPB_OK;
    }
};

struct graph_guard {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    graph_guard() = default;
    graph_guard(const graph_guard&) = delete;
    graph_guard& operator=(const graph_guard&) = delete;
    ~graph_guard()
    {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};

// grid of the streaming vector kernels: up to 8 blocks per CU
unsigned vgrid(const hpb_handle_s* h, size_t n, size_t per_cu = 8)
{
    size_t g = (n + DB - 1) / DB;
    const size_t cap = (size_t)(h->num_cu > 0 ? h->num_cu : 256) * per_cu;
    if (g > cap) g = cap;
    return (unsigned)(g < 1 ? 1 : g);
}
// vector kernels that also leave per-block partials: 4 blocks per CU stream as
// fast and halve what the finishing kernel merges
unsigned rgrid(const hpb_handle_s* h, size_t n) { return vgrid(h, n, 4); }

#define DS_CHECK(call)                  \
    do {                                \
        const int _rc = (call);         \
        if (_rc != HPB_OK) return _rc;  \
    } while (0)

template <typename T>
struct problem {
    hpb_handle_t h;
    int rows;
    int64_t nnz;
    const int* rp;
    const int* ci;
    const T* val;
    const T* b;
    T* x;
};

template <typename T>
int end_capture(hipStream_t st, graph_guard& gg, int rc)
{
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != HPB_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_status(e, __FILE__, __LINE__, "hipStreamEndCapture");
    gg.g = g;
    HPB_HIP(hipGraphInstantiate(&gg.e, gg.g, nullptr, nullptr, 0));
    return HPB_OK;
}

// multi-dot / multi-update over more than 64 columns: chunks of 64
template <typename T>
int multidot_chunks(hpb_handle_t h, int k, T* const* V, size_t n, const T* w, bool comp, T* d4, const int* skip)
{
    for (int c0 = 0; c0 < k; c0 += HPB_MAX_MULTIDOT) {
        const int kc = k - c0 < HPB_MAX_MULTIDOT ? k - c0 : HPB_MAX_MULTIDOT;
        DS_CHECK(multidot_async<T>(h, kc, (const T* const*)(V + c0), n, w, comp, d4 + 4 * c0, skip));
    }
    return HPB_OK;
}
template <typename T>
int multiaxpy_chunks(hpb_handle_t h, int k, const T* coef, T* const* V, size_t n, T* y, bool comp, const int* skip)
{
    for (int c0 = 0; c0 < k; c0 += HPB_MAX_MULTIDOT) {
        const int kc = k - c0 < HPB_MAX_MULTIDOT ? k - c0 : HPB_MAX_MULTIDOT;
        DS_CHECK(multiaxpy_devcoef<T>(h, kc, coef + c0, (const T* const*)(V + c0), n, y, comp, skip));
    }
    return HPB_OK;
}

template <typename T>
int read_state(hpb_handle_t h, const dstate* d, dstate* out)
{
    HPB_HIP(hipMemcpyAsync(h->hres, d, sizeof(dstate), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    std::memcpy(out, h->hres, sizeof(dstate));
    return HPB_OK;
}

template <typename T>
int init_state(hpb_handle_t h, dstate* d, const hpb_solver_params& p, int basis, int hist_cap)
{
    dstate s{};
    s.rscale = 1.0;
    s.xscale = 1.0;
    s.max_iters = p.max_iters;
    s.min_iters = p.min_iters;
    s.recalc_freq = p.resid_recalc_freq;
    s.early_exit = p.early_exit;
    s.basis = basis;
    s.hist_cap = hist_cap;
    std::memcpy(h->hres, &s, sizeof s);
    HPB_HIP(hipMemcpyAsync(d, h->hres, sizeof s, hipMemcpyHostToDevice, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    return HPB_OK;
}

// One solver instance: buffers and the captured cycle graph live as long as the
// object, so repeated solves with the same operands (the inner solves of
// iterative refinement) replay the graph without allocating or capturing again.
// FP64 side of the reliable-update mode: the FP32 iteration runs on P (FP32
// values, right-hand side slot unused, x = the FP32 increment), the FP64 solution
// x64 and true residual b64 - A64 x64 are updated every time the FP32 recursive
// residual has dropped by delta.
struct rel64 {
    const double* val;
    const double* b;
    double* x;
    double delta;
    double mscale;   // the FP32 values are mscale * val (a power of two)
};

template <typename T>
class dsolver : public dsolver_base {
public:
    problem<T> P{};
    hpb_solver_params prm{};
    int hcap = 0;
    bool reliable = false;
    rel64 R{};

    int setup(const problem<T>& pr, const hpb_solver_params& p, int hist_cap, const rel64* rel = nullptr)
    {
        P = pr;
        prm = p;
        hcap = hist_cap;
        reliable = rel != nullptr;
        if (rel) R = *rel;
        DS_CHECK(dst.alloc(1));
        if (hcap) DS_CHECK(dhist.alloc(2 * (size_t)hcap));
        if (reliable) {
            DS_CHECK(r64.alloc((size_t)P.rows));
            DS_CHECK(n4d.alloc(4));
        }
        return p.solver == HPB_SOLVER_GMRES ? setup_gmres() : setup_bicgstabl();
    }
    bool matches(const problem<T>& pr, const hpb_solver_params& p, int hist_cap) const
    {
        const bool same_problem = pr.h == P.h && pr.rows == P.rows && pr.nnz == P.nnz && pr.rp == P.rp &&
                                  pr.ci == P.ci && pr.val == P.val && pr.b == P.b && pr.x == P.x;
        const bool same_params = p.solver == prm.solver && p.restart == prm.restart && p.basis == prm.basis &&
                                 p.high_precision == prm.high_precision &&
                                 p.compensated_spmv == prm.compensated_spmv && p.gram_schmidt == prm.gram_schmidt &&
                                 p.resid_recalc_freq == prm.resid_recalc_freq;
        return same_problem && same_params && hist_cap == hcap;
    }
    // monitor settings (tolerances, iteration caps, early exit, recalculation)
    // are read from p on every run; the captured graph does not depend on them
    int run(const hpb_solver_params& p, dstate* out)
    {
        prm.max_iters = p.max_iters;
        prm.min_iters = p.min_iters;
        prm.rel_tol = p.rel_tol;
        prm.abs_tol = p.abs_tol;
        prm.early_exit = p.early_exit;
        prm.resid_recalc_freq = p.resid_recalc_freq;
        DS_CHECK(init_state<T>(P.h, dst.p, prm, prm.solver == HPB_SOLVER_GMRES ? 1 : prm.basis, hcap));
        if (reliable) return run_reliable(out);
        return prm.solver == HPB_SOLVER_GMRES ? run_gmres(out) : run_bicgstabl(out);
    }
    const double* history() const { return dhist.p; }

private:
    dbuf<dstate> dst;
    dbuf<double> dhist;
    graph_guard gg;
    std::vector<dbuf<T>> Vb, Ub;
    std::vector<T*> V;
    dbuf<double> H, cs, sn, g;
    dbuf<T> d4, n4, coef, y, scal, rt, da, db, gram;
    dbuf<double> r64, n4d;
    int gstride = 0;

    double* hist() { return hcap ? dhist.p : nullptr; }

    int setup_gmres()
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const int m = prm.restart;
        const bool comp = prm.high_precision != 0, cspmv = prm.compensated_spmv != 0;
        const bool dgks = prm.gram_schmidt == 2;   // else CGS2: both passes always
        const int ldh = m + 1;
        Vb.resize(m + 1);
        V.assign(m + 1, nullptr);
        for (int j = 0; j <= m; ++j) {
            DS_CHECK(Vb[j].alloc(n));
            V[j] = Vb[j].p;
        }
        DS_CHECK(H.alloc((size_t)ldh * m));
        DS_CHECK(cs.alloc(m));
        DS_CHECK(sn.alloc(m));
        DS_CHECK(g.alloc(m + 1));
        DS_CHECK(n4.alloc(4));
        DS_CHECK(coef.alloc(m + 1));
        DS_CHECK(y.alloc(m));
        DS_CHECK(scal.alloc(m + 1));   // sig_i: the basis is sig_i V_i
        dstate* S = dst.p;
        const int* done = &S->done;
        const int* stop = &S->stop;
        basis_pack<T> bp{};
        for (int j = 0; j <= m; ++j) bp.p[j] = V[j];
        double* hp = hist();
        T* sig = scal.p;
        T* part = reinterpret_cast<T*>(h->part);
        const size_t cus = (size_t)(h->num_cu > 0 ? h->num_cu : 256);
        const size_t part_pairs = h->part_bytes / (2 * sizeof(T));
        // grids over 16-byte vectors; the reductions end in per-block partials
        // merged by one-block kernels (k_gm_hfin, k_gm_nfin)
        const size_t gvec = (n / dev::vec16<T>::W + DB - 1) / DB;
        // basis larger than the 256 MB MALL: stream it (128^3 FP64 GMRES(50), 816 MB:
        // 0.275 s against 0.293 s; 256^3 FP64 GMRES(30): 11.79 s against 12.63 s); a
        // basis that fits is read back from the MALL by the next pass (128^3 FP32
        // GMRES(30), 248 MB: 0.106 s streamed against 0.098 s)
        const bool nt_basis = (double)(m + 1) * (double)n * sizeof(T) > 256.0 * 1024 * 1024;
        // multi-dot blocks per CU over all column chunks: 8 from 2^23 rows (256^3 FP64
        // GMRES(30): 10.99 s, against 11.78 s at 4, 11.10 s at 16), else 4 (128^3:
        // 8 per CU cost 3-6 %, 2 and 1 per CU 2-9 %); updates: 8 per CU (2 and 1 per CU
        // were 4-35 % slower, 16 the same)
        const size_t gd_per_cu = n >= ((size_t)1 << 23) ? 8 : 4;
        const unsigned gupd = (unsigned)std::max<size_t>(1, std::min(gvec, 8 * cus));
        constexpr int KC = 8;
        HPB_HIP(hipStreamSynchronize(st));
        HPB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = [&]() -> int {
            DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, P.x, P.b, V[0], cspmv, done));
            DS_CHECK(nrm2_async<T>(h, n, V[0], n4.p, comp, done));
            hipLaunchKernelGGL((k_gm_start<T>), dim3(1), dim3(1), 0, st, S, n4.p, g.p, m, sig, hp);
            for (int j = 0; j < m; ++j) {
                const int k = j + 1;
                // w = A (sig_j V_j), stored unnormalised in V[j+1]
                DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, V[j], nullptr, V[j + 1], cspmv, stop,
                                      sig + j));
                double* Hj = H.p + (size_t)j * ldh;
                for (int pass = 0; pass < 2; ++pass) {
                    // pass 1 runs under skip1 when the first pass may be accepted
                    const int* skip = (dgks && pass == 1) ? &S->skip1 : stop;
                    const int kk = (dgks && pass == 0) ? k + 1 : k;   // + (w, w)
                    const unsigned nch = (unsigned)((kk + KC - 1) / KC);
                    size_t gd = std::min(gvec, std::max<size_t>(cus, (gd_per_cu * cus) / nch));
                    gd = std::max<size_t>(1, std::min(gd, part_pairs / (size_t)kk));
                    auto dot = [&](auto compc, auto ntc) {
                        constexpr bool C = decltype(compc)::value, NTB = decltype(ntc)::value;
                        hipLaunchKernelGGL((k_gm_dot<T, KC, C, NTB>), dim3((unsigned)gd, nch), dim3(DB), 0, st, n,
                                           kk, bp, V[j + 1], part, skip);
                        hipLaunchKernelGGL((k_gm_hfin<T, C>), dim3(1), dim3(1024), 0, st, part, (unsigned)gd, k, kk,
                                           sig, Hj, coef.p, pass, skip, S);
                    };
                    auto upd = [&](auto compc, auto normc) {
                        constexpr bool C = decltype(compc)::value, NM = decltype(normc)::value;
                        if (nt_basis)
                            hipLaunchKernelGGL((k_gm_upd<T, C, NM, true>), dim3(gupd), dim3(DB), 0, st, n, k, bp,
                                               coef.p, V[j + 1], part, skip);
                        else
                            hipLaunchKernelGGL((k_gm_upd<T, C, NM, false>), dim3(gupd), dim3(DB), 0, st, n, k, bp,
                                               coef.p, V[j + 1], part, skip);
                        if (NM && dgks && pass == 0)
                            hipLaunchKernelGGL((k_gm_dgks<T, C>), dim3(1), dim3(DB), 0, st, part, gupd, j, m, H.p,
                                               ldh, cs.p, sn.p, g.p, sig, S);
                        else if (NM)
                            hipLaunchKernelGGL((k_gm_nfin<T, C>), dim3(1), dim3(DB), 0, st, part, gupd, j, m, H.p,
                                               ldh, cs.p, sn.p, g.p, sig, S, skip);
                    };
                    using BT = std::true_type;
                    using BF = std::false_type;
                    if (comp) {
                        if (nt_basis) dot(BT{}, BT{});
                        else dot(BT{}, BF{});
                    } else {
                        if (nt_basis) dot(BF{}, BT{});
                        else dot(BF{}, BF{});
                    }
                    const bool norm = pass == 1 || dgks;
                    if (comp) {
                        if (norm) upd(BT{}, BT{});
                        else upd(BT{}, BF{});
                    } else {
                        if (norm) upd(BF{}, BT{});
                        else upd(BF{}, BF{});
                    }
                }
            }
            hipLaunchKernelGGL((k_gm_backsub<T>), dim3(1), dim3(1), 0, st, m, H.p, ldh, g.p, sig, y.p, S);
            if (comp)
                hipLaunchKernelGGL((k_gm_xupd<T, true>), dim3(vgrid(h, n)), dim3(DB), 0, st, n, bp, y.p, P.x, S);
            else
                hipLaunchKernelGGL((k_gm_xupd<T, false>), dim3(vgrid(h, n)), dim3(DB), 0, st, n, bp, y.p, P.x, S);
            return HPB_OK;
        }();
        return end_capture<T>(st, gg, rc);
    }

    int run_gmres(dstate* out)
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const bool comp = prm.high_precision != 0;
        DS_CHECK(nrm2_async<T>(h, (size_t)P.rows, P.b, n4.p, comp, nullptr));
        hipLaunchKernelGGL((k_init<T>), dim3(1), dim3(1), 0, st, dst.p, n4.p, prm.rel_tol, prm.abs_tol);
        HPB_LAUNCH_CHECK();
        // every replay starts with the restart residual and the monitor check,
        // so the replay after the last useful cycle ends the solve
        const int max_cycles = prm.max_iters + 2;
        for (int c = 0; c < max_cycles; ++c) {
            HPB_HIP(hipGraphLaunch(gg.e, st));
            DS_CHECK(read_state<T>(h, dst.p, out));
            if (out->done) break;
        }
        return HPB_OK;
    }

    // fused launchers of the BiCGStab(L) cycle (L <= 4, residual recomputed every
    // cycle or never, or reliable updates)
    template <int MODE, int EPI>
    void ds_spmv(const T* x, const T* b, T* y, const T* rtv, const int* skip)
    {
        hpb_handle_t h = P.h;
        const double mean = P.rows > 0 ? (double)P.nnz / P.rows : 1.0;
        const int lpr = mean <= 8.0 ? 2 : (mean <= 12.0 ? 4 : (mean <= 24.0 ? 8 : 16));   // as the library SpMV
        const long per_block = lpr <= 4 ? 128 : DB / lpr;   // staged tiles of 128 rows up to 4 lanes per row
        long grid = (P.rows + per_block - 1) / per_block;
        const long cap = (long)(h->num_cu > 0 ? h->num_cu : 256) * 16;   // the gathers want many waves
        grid = std::max(1L, std::min(grid, cap));
        const bool cr = prm.compensated_spmv != 0, comp = prm.high_precision != 0;
        T* part = reinterpret_cast<T*>(h->part);
#define DS_SPMV(L_, CR_, CO_)                                                                                \
    hipLaunchKernelGGL((k_ds_spmv<T, L_, CR_, CO_, MODE, EPI>), dim3((unsigned)grid), dim3(DB), 0, h->stream,   \
                       P.rows, P.rp, P.ci, P.val, x, b, y, rtv, part, skip)
#define DS_SPMV_L(L_)                                                                                        \
    do {                                                                                                     \
        if (cr) { if (comp) DS_SPMV(L_, true, true); else DS_SPMV(L_, true, false); }                        \
        else { if (comp) DS_SPMV(L_, false, true); else DS_SPMV(L_, false, false); }                         \
    } while (0)
        if (lpr == 2) DS_SPMV_L(2);
        else if (lpr == 4) DS_SPMV_L(4);
        else if (lpr == 8) DS_SPMV_L(8);
        else DS_SPMV_L(16);
#undef DS_SPMV_L
#undef DS_SPMV
        if constexpr (EPI != EPI_NONE) ds_fin<EPI>((unsigned)grid, skip);
    }

    // the scalar step after a fused kernel whose G blocks left partials in part
    template <int EPI>
    void ds_fin(unsigned G, const int* skip)
    {
        hpb_handle_t h = P.h;
        const T* part = reinterpret_cast<const T*>(h->part);
        if (prm.high_precision)
            hipLaunchKernelGGL((k_ds_fin<T, true, EPI>), dim3(1), dim3(DB), 0, h->stream, part, G, dst.p, coef.p,
                               hist(), R.delta, skip);
        else
            hipLaunchKernelGGL((k_ds_fin<T, false, EPI>), dim3(1), dim3(DB), 0, h->stream, part, G, dst.p, coef.p,
                               hist(), R.delta, skip);
    }

    bool fused_ok() const
    {
        return prm.basis <= 4 && (reliable || prm.resid_recalc_freq == 0 || prm.resid_recalc_freq == 1);
    }

    template <int L>
    void ds_gram(bs_pack<T> bp, const int* skip)
    {
        hpb_handle_t h = P.h;
        const size_t n = (size_t)P.rows;
        T* part = reinterpret_cast<T*>(h->part);
        const unsigned G = rgrid(h, n);
        if (prm.high_precision) {
            hipLaunchKernelGGL((k_ds_gram<T, L, true>), dim3(G), dim3(DB), 0, h->stream, n, bp, part, skip);
            hipLaunchKernelGGL((k_ds_gfin<T, L, true>), dim3(1), dim3(DB), 0, h->stream, part, G, dst.p, coef.p,
                               skip);
        } else {
            hipLaunchKernelGGL((k_ds_gram<T, L, false>), dim3(G), dim3(DB), 0, h->stream, n, bp, part, skip);
            hipLaunchKernelGGL((k_ds_gfin<T, L, false>), dim3(1), dim3(DB), 0, h->stream, part, G, dst.p, coef.p,
                               skip);
        }
    }

    int setup_bicgstabl()
    {
        fused = fused_ok();
        if (fused) return setup_bicgstabl_fused();
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const int L = prm.basis;
        const bool comp = prm.high_precision != 0, cspmv = prm.compensated_spmv != 0;
        Vb.resize(L + 1);
        Ub.resize(L + 1);
        V.assign(L + 1, nullptr);
        bs_pack<T> bp{};
        for (int j = 0; j <= L; ++j) {
            DS_CHECK(Vb[j].alloc(n));
            DS_CHECK(Ub[j].alloc(n));
            bp.r[j] = Vb[j].p;
            bp.u[j] = Ub[j].p;
            V[j] = Vb[j].p;
        }
        gstride = 4 * (L + 1);
        DS_CHECK(rt.alloc(n));
        DS_CHECK(n4.alloc(4));
        DS_CHECK(da.alloc(4));
        DS_CHECK(db.alloc(4));
        DS_CHECK(gram.alloc((size_t)gstride * (L + 1)));
        DS_CHECK(coef.alloc(2 + MAXL));
        DS_CHECK(scal.alloc(1));
        dstate* S = dst.p;
        // reliable mode pauses the cycle with stop while the host updates x64
        const int* done = reliable ? &S->stop : &S->done;
        const unsigned vg = vgrid(h, n);
        double* hp = hist();
        HPB_HIP(hipStreamSynchronize(st));
        HPB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = [&]() -> int {
            for (int j = 0; j < L; ++j) {
                DS_CHECK(dot2_async<T>(h, n, V[j], rt.p, da.p, comp, done));
                hipLaunchKernelGGL((k_bs_beta<T>), dim3(1), dim3(1), 0, st, S, da.p, coef.p, j == 0 ? 1 : 0, done);
                hipLaunchKernelGGL((k_bs_upd_u<T>), dim3(vg), dim3(DB), 0, st, n, j, bp, coef.p, done);
                DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, Ub[j].p, nullptr, Ub[j + 1].p, cspmv,
                                      done));
                DS_CHECK(dot2_async<T>(h, n, Ub[j + 1].p, rt.p, db.p, comp, done));
                hipLaunchKernelGGL((k_bs_alpha<T>), dim3(1), dim3(1), 0, st, S, db.p, coef.p, done);
                hipLaunchKernelGGL((k_bs_upd_r<T>), dim3(vg), dim3(DB), 0, st, n, j, bp, P.x, coef.p, done);
                DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, V[j], nullptr, V[j + 1], cspmv, done));
            }
            for (int k = 1; k <= L; ++k)
                DS_CHECK(multidot_async<T>(h, k + 1, (const T* const*)V.data(), n, V[k], comp,
                                           gram.p + (size_t)k * gstride, done));
            hipLaunchKernelGGL((k_bs_mr<T>), dim3(1), dim3(1), 0, st, S, L, gram.p, gstride, coef.p, done);
            if (comp)
                hipLaunchKernelGGL((k_bs_upd_mr<T, true>), dim3(vg), dim3(DB), 0, st, n, L, bp, P.x, coef.p, done);
            else
                hipLaunchKernelGGL((k_bs_upd_mr<T, false>), dim3(vg), dim3(DB), 0, st, n, L, bp, P.x, coef.p, done);
            if (reliable) {
                DS_CHECK(nrm2_async<T>(h, n, V[0], n4.p, comp, done));
                hipLaunchKernelGGL((k_rel_check<T>), dim3(1), dim3(1), 0, st, S, n4.p, R.delta);
                return HPB_OK;
            }
            hipLaunchKernelGGL(k_bs_recalc_flag, dim3(1), dim3(1), 0, st, S);
            DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, P.x, P.b, V[0], cspmv, &S->skip_recalc));
            DS_CHECK(nrm2_async<T>(h, n, V[0], n4.p, comp, done));
            hipLaunchKernelGGL((k_bs_check<T>), dim3(1), dim3(1), 0, st, S, n4.p, hp);
            return HPB_OK;
        }();
        return end_capture<T>(st, gg, rc);
    }

    int setup_bicgstabl_fused()
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const int L = prm.basis;
        const bool comp = prm.high_precision != 0;
        Vb.resize(L + 1);
        Ub.resize(L + 1);
        V.assign(L + 1, nullptr);
        bs_pack<T> bp{};
        for (int j = 0; j <= L; ++j) {
            DS_CHECK(Vb[j].alloc(n));
            DS_CHECK(Ub[j].alloc(n));
            bp.r[j] = Vb[j].p;
            bp.u[j] = Ub[j].p;
            V[j] = Vb[j].p;
        }
        gstride = 4 * (L + 1);
        DS_CHECK(rt.alloc(n));
        DS_CHECK(n4.alloc(4));
        DS_CHECK(da.alloc(4));
        DS_CHECK(db.alloc(4));
        DS_CHECK(gram.alloc((size_t)gstride * (L + 1)));
        DS_CHECK(coef.alloc(2 + MAXL));
        DS_CHECK(scal.alloc(1));
        dstate* S = dst.p;
        const int* skip = reliable ? &S->stop : &S->done;
        const unsigned vg = vgrid(h, n);
        T* part = reinterpret_cast<T*>(h->part);
        const bool recalc = !reliable && prm.resid_recalc_freq == 1;
        HPB_HIP(hipStreamSynchronize(st));
        HPB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = [&]() -> int {
            for (int j = 0; j < L; ++j) {
                hipLaunchKernelGGL((k_bs_upd_u<T>), dim3(vg), dim3(DB), 0, st, n, j, bp, coef.p, skip);
                ds_spmv<0, EPI_ALPHA>(Ub[j].p, nullptr, Ub[j + 1].p, rt.p, skip);
                hipLaunchKernelGGL((k_bs_upd_r<T>), dim3(vg), dim3(DB), 0, st, n, j, bp, P.x, coef.p, skip);
                if (j + 1 < L) ds_spmv<0, EPI_BETA>(V[j], nullptr, V[j + 1], rt.p, skip);
                else ds_spmv<0, EPI_NONE>(V[j], nullptr, V[j + 1], rt.p, skip);
            }
            switch (L) {
            case 1: ds_gram<1>(bp, skip); break;
            case 2: ds_gram<2>(bp, skip); break;
            case 3: ds_gram<3>(bp, skip); break;
            default: ds_gram<4>(bp, skip); break;
            }
            const int epi = reliable ? EPI_REL : EPI_CHECK;
            if (recalc) {
                if (comp)
                    hipLaunchKernelGGL((k_bs_upd_mr<T, true>), dim3(vg), dim3(DB), 0, st, n, L, bp, P.x, coef.p, skip);
                else
                    hipLaunchKernelGGL((k_bs_upd_mr<T, false>), dim3(vg), dim3(DB), 0, st, n, L, bp, P.x, coef.p, skip);
                ds_spmv<1, EPI_CHECK>(P.x, P.b, V[0], rt.p, skip);
            } else {
                const unsigned rg = rgrid(h, n);
                if (comp)
                    hipLaunchKernelGGL((k_ds_upd_mr<T, true>), dim3(rg), dim3(DB), 0, st, n, L, bp, P.x, rt.p, coef.p,
                                       part, skip);
                else
                    hipLaunchKernelGGL((k_ds_upd_mr<T, false>), dim3(rg), dim3(DB), 0, st, n, L, bp, P.x, rt.p,
                                       coef.p, part, skip);
                if (epi == EPI_REL) ds_fin<EPI_REL>(rg, skip);
                else ds_fin<EPI_CHECK>(rg, skip);
            }
            return HPB_OK;
        }();
        return end_capture<T>(st, gg, rc);
    }

    int run_bicgstabl(dstate* out)
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const int L = prm.basis;
        const bool comp = prm.high_precision != 0, cspmv = prm.compensated_spmv != 0;
        const unsigned vg = vgrid(h, n);
        dstate* S = dst.p;
        // ||b||, tol, r_0 = b - A x, r~ = r_0 / ||r_0||, u_0 = 0
        DS_CHECK(nrm2_async<T>(h, n, P.b, n4.p, comp, nullptr));
        hipLaunchKernelGGL((k_init<T>), dim3(1), dim3(1), 0, st, S, n4.p, prm.rel_tol, prm.abs_tol);
        DS_CHECK(csr_apply<T>(h, P.rows, P.nnz, P.rp, P.ci, P.val, P.x, P.b, V[0], cspmv, nullptr));
        DS_CHECK(nrm2_async<T>(h, n, V[0], n4.p, comp, nullptr));
        hipLaunchKernelGGL((k_bs_init<T>), dim3(1), dim3(1), 0, st, S, n4.p, scal.p, hist());
        hipLaunchKernelGGL((k_scale_copy<T>), dim3(vg), dim3(DB), 0, st, n, V[0], rt.p, scal.p, &S->invalid);
        hipLaunchKernelGGL((k_fill<T>), dim3(vg), dim3(DB), 0, st, n, Ub[0].p, T(0));
        HPB_LAUNCH_CHECK();
        DS_CHECK(fused_beta(&S->done));
        DS_CHECK(read_state<T>(h, S, out));
        if (out->done) return HPB_OK;
        const int max_cycles = prm.max_iters / L + 2;
        const int every = 4;
        for (int c = 0; c < max_cycles; c += every) {
            for (int k = 0; k < every; ++k) HPB_HIP(hipGraphLaunch(gg.e, st));
            DS_CHECK(read_state<T>(h, S, out));
            if (out->done) break;
        }
        return HPB_OK;
    }

    // FP64 update: x64 += x32, x32 = 0, r64 = b64 - A64 x64 (compensated rows),
    // monitor check on ||r64||, r32 = (float) r64 (and on the first call r~, u_0)
    int rel_update(int first, int force_restart = 0)
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const unsigned vg = vgrid(h, n);
        dstate* S = dst.p;
        if (!first)
            hipLaunchKernelGGL(k_rel_flush, dim3(vg), dim3(DB), 0, st, n, (float*)P.x, R.x, S, &S->done);
        DS_CHECK(csr_apply<double>(h, P.rows, P.nnz, P.rp, P.ci, R.val, R.x, R.b, r64.p, true, &S->done));
        DS_CHECK(nrm2_async<double>(h, n, r64.p, n4d.p, true, &S->done));
        hipLaunchKernelGGL(k_rel_after, dim3(1), dim3(1), 0, st, S, n4d.p, hist(), first | force_restart, R.mscale);
        hipLaunchKernelGGL(k_rel_convert, dim3(vg), dim3(DB), 0, st, n, r64.p, (float*)V[0], (float*)rt.p,
                           (float*)Ub[0].p, S);
        HPB_LAUNCH_CHECK();
        return fused_beta(&S->done);
    }

    int run_reliable(dstate* out)
    {
        hpb_handle_t h = P.h;
        hipStream_t st = h->stream;
        const size_t n = (size_t)P.rows;
        const int L = prm.basis;
        dstate* S = dst.p;
        DS_CHECK(nrm2_async<double>(h, n, R.b, n4d.p, true, nullptr));
        hipLaunchKernelGGL((k_init<double>), dim3(1), dim3(1), 0, st, S, n4d.p, prm.rel_tol, prm.abs_tol);
        hipLaunchKernelGGL((k_fill<T>), dim3(vgrid(h, n)), dim3(DB), 0, st, n, P.x, T(0));
        DS_CHECK(rel_update(1));
        DS_CHECK(read_state<T>(h, S, out));
        const int max_cycles = prm.max_iters / L + 2;
        const int every = 4;
        constexpr int max_breakdown_restarts = 16;
        int cycles = 0, restarts = 0;
        while (cycles < max_cycles) {
            if (out->done) {
                // a breakdown restarts from an FP64 update; anything else ends the solve
                if (!out->breakdown || out->converged || out->invalid || out->iters >= prm.max_iters ||
                    restarts >= max_breakdown_restarts)
                    break;
                hipLaunchKernelGGL(k_rel_resume, dim3(1), dim3(1), 0, st, S);
                DS_CHECK(rel_update(0, 1));
                DS_CHECK(read_state<T>(h, S, out));
                restarts++;
                updates++;
                continue;
            }
            for (int k = 0; k < every; ++k) HPB_HIP(hipGraphLaunch(gg.e, st));
            cycles += every;
            DS_CHECK(read_state<T>(h, S, out));
            if (!out->done && out->stop) {
                DS_CHECK(rel_update(0));
                DS_CHECK(read_state<T>(h, S, out));
                updates++;
            }
        }
        // keep the progress of an unfinished run (x32 holds the increment since the last update)
        hipLaunchKernelGGL(k_rel_flush, dim3(vgrid(h, n)), dim3(DB), 0, st, n, (float*)P.x, R.x, S, &S->invalid);
        HPB_LAUNCH_CHECK();
        return HPB_OK;
    }

public:
    int updates = 0;
    bool fused = false;   // BiCGStab(L): the cycle graph finishes with beta for the next cycle

    // fused cycle: beta and rho for the first BiCG step from the current r_0
    int fused_beta(const int* skip)
    {
        if (!fused) return HPB_OK;
        hpb_handle_t h = P.h;
        DS_CHECK(dot2_async<T>(h, (size_t)P.rows, V[0], rt.p, da.p, prm.high_precision != 0, skip));
        hipLaunchKernelGGL((k_bs_beta<T>), dim3(1), dim3(1), 0, h->stream, dst.p, da.p, coef.p, 1, skip);
        HPB_LAUNCH_CHECK();
        return HPB_OK;
    }
};

}  // namespace

// Device-resident solve of A x = b (vectors of type T, scalars FP64).  Fills res
// (iters, converged, breakdown, resid_norm, rhs_norm, history_len) and history.
// cache (optional): a solver object kept by the caller between calls with the
// same operands; *cache is replaced when the operands or parameters change and
// must be released with dsolve_release.
template <typename T>
int dsolve(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* ci, const T* val, const T* b, T* x,
           const hpb_solver_params& p, hpb_solver_result* res, double* history, int history_cap,
           dsolver_base** cache)
{
    HPB_REQUIRE(p.solver != HPB_SOLVER_GMRES || (p.restart >= 1 && p.restart < MAXB), HPB_ERR_INVALID,
                "device solver: GMRES restart must be in [1, 255]");
    HPB_REQUIRE(p.solver != HPB_SOLVER_BICGSTABL || (p.basis >= 1 && p.basis <= MAXL), HPB_ERR_UNSUPPORTED,
                "device solver: BiCGStab(L) needs 1 <= L <= 8");
    const int hcap = (history && history_cap > 0) ? history_cap : 0;
    problem<T> P{h, rows, nnz, rp, ci, val, b, x};
    dsolver<T>* ds = nullptr;
    dsolver<T> local;
    if (cache && *cache) {
        ds = static_cast<dsolver<T>*>(*cache);
        if (!ds->matches(P, p, hcap)) {
            delete ds;
            *cache = nullptr;
            ds = nullptr;
        }
    }
    if (!ds) {
        ds = cache ? new dsolver<T>() : &local;
        const int rc = ds->setup(P, p, hcap);
        if (rc != HPB_OK) {
            if (cache) delete ds;
            return rc;
        }
        if (cache) *cache = ds;
    }
    dstate out{};
    DS_CHECK(ds->run(p, &out));
    res->iters = out.iters;
    res->converged = out.converged;
    res->breakdown = out.breakdown;
    res->resid_norm = out.rnorm;
    res->rhs_norm = out.bnorm;
    res->history_len = out.hist_len;
    if (hcap && out.hist_len > 0) {
        HPB_HIP(hipMemcpyAsync(history, ds->history(), sizeof(double) * 2 * (size_t)out.hist_len,
                               hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipStreamSynchronize(h->stream));
    }
    return HPB_OK;
}

void dsolve_release(dsolver_base* cache) { delete cache; }

// FP32 BiCGStab(L) with reliable updates of an FP64 solution: val32 / val64 share
// rp / ci; b64, x64 FP64; r32, x32 FP32 work vectors of the caller.  The inner
// rel_tol is the reduction delta of the FP32 recursive residual that triggers an
// FP64 update; rel_tol is the final tolerance on ||b - A x|| / ||b||.
int dsolve_reliable(hpb_handle_t h, int rows, int64_t nnz, const int* rp, const int* ci, const float* val32,
                    double mscale, const double* val64, const double* b64, double* x64, float* r32, float* x32,
                    const hpb_solver_params& inner, double rel_tol, hpb_solver_result* res, int* updates)
{
    HPB_REQUIRE(inner.solver == HPB_SOLVER_BICGSTABL && inner.basis >= 1 && inner.basis <= MAXL, HPB_ERR_UNSUPPORTED,
                "reliable updates: BiCGStab(L <= 8) inner solver only");
    HPB_REQUIRE(inner.rel_tol > 0.0 && inner.rel_tol < 1.0, HPB_ERR_INVALID,
                "reliable updates: the inner rel_tol (update threshold) must be in (0, 1)");
    problem<float> P{h, rows, nnz, rp, ci, val32, r32, x32};
    hpb_solver_params p = inner;
    const double delta = p.rel_tol;
    p.rel_tol = rel_tol;   // the monitor tests the FP64 true residual against the final tolerance
    rel64 R{val64, b64, x64, delta, mscale};
    dsolver<float> ds;
    DS_CHECK(ds.setup(P, p, 0, &R));
    dstate out{};
    DS_CHECK(ds.run(p, &out));
    res->iters = out.iters;
    res->converged = out.converged;
    res->breakdown = out.breakdown;
    res->resid_norm = out.rlast;
    res->rhs_norm = out.bnorm;
    res->history_len = 0;
    if (updates) *updates = ds.updates;
    return HPB_OK;
}

// 2^-e with max |x_i| 2^-e in [0.5, 1): scaling FP64 values by it before the
// conversion to FP32 changes no significand, and keeps |a_ij| < 1 in the FP32 copy
int absmax_pow2_scale(hpb_handle_t h, size_t n, const double* x, double* scale)
{
    HPB_REQUIRE(h && scale && (n == 0 || x), HPB_ERR_INVALID, "absmax: null argument");
    *scale = 1.0;
    if (n == 0) return HPB_OK;
    auto* d = reinterpret_cast<unsigned long long*>(h->dres);
    HPB_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), h->stream));
    hipLaunchKernelGGL(k_absmax, dim3(vgrid(h, n)), dim3(DB), 0, h->stream, n, x, d);
    HPB_LAUNCH_CHECK();
    HPB_HIP(hipMemcpyAsync(h->hres, d, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    double m;
    std::memcpy(&m, h->hres, sizeof m);
    if (m > 0.0 && std::isfinite(m)) *scale = std::ldexp(1.0, -(std::ilogb(m) + 1));
    return HPB_OK;
}

template int dsolve<double>(hpb_handle_t, int, int64_t, const int*, const int*, const double*, const double*,
                            double*, const hpb_solver_params&, hpb_solver_result*, double*, int, dsolver_base**);
template int dsolve<float>(hpb_handle_t, int, int64_t, const int*, const int*, const float*, const float*, float*,
                           const hpb_solver_params&, hpb_solver_result*, double*, int, dsolver_base**);

}  // namespace hpb
