// hpb_reduce.hip -- single-launch compensated and plain reductions for gfx950.
//
// Replaces gpu_reduction_ogita (source/common/ogita/gpu_reduction_ogita_impl.cuh:35-676,
// 1169-1282) and gpu_reduction (source/common/gpu_reduction_impl.cuh:335-841,
// 1121-1244).  Design:
//   * one launch per reduction: a bounded grid of G <= grid_max blocks streams the
//     vectors with 16-byte loads (double2 / float4), UNROLL loads per input per
//     thread in flight, grid-stride over tiles of BLOCK*UNROLL vectors;
//   * every lane keeps one (s, c) Ogita-Rump-Oishi accumulator; nothing is written
//     per element (the reference zeroes and re-reads an N-sized error array);
//   * wave64 butterfly merge with TwoSum, LDS across the waves of a block;
//   * per-block (s, c) partials go to a handle-owned workspace of 2G words; a
//     one-block kernel queued behind the reduction merges them in block-index
//     order, so the result does not depend on the order in which the blocks ran
//     (reproducible); no floating-point atomics, no in-place multi-pass (the
//     reference's passes read and write the same array from different blocks,
//     impl.cuh:1247-1253);
//   * the norms are safe against overflow and underflow, as the reference's
//     standard-precision norm (cuBLAS nrm2, gpu_vector_operations.h:315-327) is:
//     the launch also takes M = max |x_i|; when M lies outside [lo, hi] (2^-400,
//     2^400 in FP64; 2^-30, 2^30 in FP32) a second launch of the same shape,
//     queued behind it, repeats the sum over x_i 2^k with the power of two 2^k
//     that puts M in [1/2, 1), and scales the result back.  Inside the range the
//     unscaled sum is kept, so those results are unchanged by the scaling.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

namespace hpb {

enum red_op { R_DOT2 = 0, R_DOT, R_SUM2, R_SUM, R_ASUM2, R_ASUM, R_NRM2C, R_NRM2 };

template <int OP> struct red_traits {
    static constexpr bool binary = (OP == R_DOT2 || OP == R_DOT);
    static constexpr bool comp = (OP == R_DOT2 || OP == R_SUM2 || OP == R_ASUM2 || OP == R_NRM2C);
    static constexpr bool nrm = (OP == R_NRM2C || OP == R_NRM2);
};

// Range of M = max |x_i| in which the norms sum unscaled squares: n M^2 stays
// finite for n < 2^48, and the TwoProd error terms of the largest squares stay
// normal numbers.
template <typename T> struct nrm_range;
template <> struct nrm_range<double> { static constexpr double lo = 0x1p-400, hi = 0x1p400; };
template <> struct nrm_range<float> { static constexpr double lo = 0x1p-30, hi = 0x1p30; };

template <typename T> __device__ __forceinline__ T absv(T v);
template <> __device__ __forceinline__ double absv<double>(double v) { return __builtin_fabs(v); }
template <> __device__ __forceinline__ float absv<float>(float v) { return __builtin_fabsf(v); }

template <typename T> __device__ __forceinline__ T sqrtv(T v);
template <> __device__ __forceinline__ double sqrtv<double>(double v) { return __builtin_sqrt(v); }
template <> __device__ __forceinline__ float sqrtv<float>(float v) { return __builtin_sqrtf(v); }

template <typename T> __device__ __forceinline__ T maxv(T a, T b);
template <> __device__ __forceinline__ double maxv<double>(double a, double b) { return __builtin_fmax(a, b); }
template <> __device__ __forceinline__ float maxv<float>(float a, float b) { return __builtin_fmaxf(a, b); }

template <typename T> __device__ __forceinline__ T ldexpv(T v, int k);
template <> __device__ __forceinline__ double ldexpv<double>(double v, int k) { return __builtin_ldexp(v, k); }
template <> __device__ __forceinline__ float ldexpv<float>(float v, int k) { return __builtin_ldexpf(v, k); }

// Block maximum (NaN ignored, as fmax does); result valid in thread 0.
template <typename T, int BLOCK>
__device__ __forceinline__ T block_max(T m)
{
    constexpr int NW = BLOCK / 64;
    __shared__ T lds_m[NW];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = maxv<T>(m, dev::shfl_xor<T>(m, d));
    if (NW > 1) {
        if ((threadIdx.x & 63) == 0) lds_m[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 1; w < NW; ++w) m = maxv<T>(m, lds_m[w]);
        __syncthreads();
    }
    return m;
}

template <typename T, int OP>
__device__ __forceinline__ void red_acc(T& s, T& c, T xv, T yv)
{
    if constexpr (OP == R_DOT2) dev::dot2_step<T>(s, c, xv, yv);
    else if constexpr (OP == R_DOT) s = dev::fma_rn<T>(xv, yv, s);
    else if constexpr (OP == R_SUM2) dev::sum2_step<T>(s, c, xv);
    else if constexpr (OP == R_SUM) s = s + xv;
    else if constexpr (OP == R_ASUM2) dev::sum2_step<T>(s, c, absv<T>(xv));
    else if constexpr (OP == R_ASUM) s = s + absv<T>(xv);
    else if constexpr (OP == R_NRM2C) dev::dot2_step<T>(s, c, xv, xv);
    else s = dev::fma_rn<T>(xv, xv, s);
}

template <typename T> struct bits_of { using type = std::conditional_t<sizeof(T) == 8, unsigned long long, unsigned>; };

template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v)
{
    using U = typename bits_of<T>::type;
    __hip_atomic_store(reinterpret_cast<U*>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T ld_agent(const T* p)
{
    using U = typename bits_of<T>::type;
    U u = __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, u);
}

template <typename T, typename V>
__device__ __forceinline__ T lane_of(const V& v, int w)
{
    if constexpr (std::is_same<T, V>::value) return v;
    else return dev::vget<T>(v, w);
}

template <typename T, int OP>
__device__ __forceinline__ void write_out(T* out, T s, T c)
{
    T r = s + c;
    out[0] = s;
    out[1] = c;
    out[2] = r;
    out[3] = red_traits<OP>::nrm ? sqrtv<T>(r) : r;
}

// Last stage of a norm launch (thread 0): the unscaled result, and the request for
// the rescaling launch when M is outside the safe range (M = 0 and an empty x
// need none; a NaN entry leaves M alone and makes the sum NaN).
template <typename T, int OP>
__device__ __forceinline__ void nrm_finish(T* out, T s, T c, T m, nrm_fix* fix)
{
    write_out<T, OP>(out, s, c);
    const double md = (double)m;
    if (md != 0.0 && !(md >= nrm_range<T>::lo && md <= nrm_range<T>::hi)) {
        fix->amax = md;
        fix->need = 1;
    }
}

// Last stage of the rescaling launch (thread 0): (s, c) is the sum over x_i 2^k;
// for M = inf, k = 0 and s alone is kept (inf, or NaN when x holds a NaN).
template <typename T>
__device__ __forceinline__ void nrm_rescaled_out(T* out, T s, T c, int k, bool inf_mode, nrm_fix* fix)
{
    if (inf_mode) {
        out[0] = s;
        out[1] = T(0);
        out[2] = s;
        out[3] = sqrtv<T>(s);
    } else {
        T r = s + c;
        out[0] = ldexpv<T>(s, -2 * k);
        out[1] = ldexpv<T>(c, -2 * k);
        out[2] = ldexpv<T>(r, -2 * k);
        out[3] = ldexpv<T>(sqrtv<T>(r), -k);
    }
    fix->need = 0;
}

// Fixed-order final stage: merges the G partials in index order (thread t takes
// t, t+BLOCK, ...; then block_reduce).
template <typename T, bool COMP, int BLOCK>
__device__ __forceinline__ void final_merge(const T* part, unsigned G, T& fs, T& fc)
{
    fs = T(0);
    fc = T(0);
    for (unsigned i = threadIdx.x; i < G; i += BLOCK) {
        T ps = part[2 * i];
        T pc = part[2 * i + 1];
        if (COMP) dev::merge<T>(fs, fc, ps, pc);
        else fs = fs + ps;
    }
    dev::block_reduce<T, COMP, BLOCK>(fs, fc);
}

// Returns true in every thread of the block that took the last ticket.  Every
// block publishes its partial with a release increment (an L2 write-back on a
// multi-XCD part); only the last block needs the acquire, taken as a fence after
// it has seen the final count, so the other blocks skip the L2 invalidate that an
// acq_rel increment would cost each of them.
__device__ __forceinline__ bool take_ticket(unsigned* ticket, unsigned G)
{
    __shared__ int am_last;
    if (threadIdx.x == 0) {
        unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == G - 1);
    }
    __syncthreads();
    bool last = am_last != 0;
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}

// NT: stream the full tiles with nontemporal loads (no reuse is expected).
// RESCALE (norms only): the rescaling launch.  It does nothing unless the norm
// launch before it asked for it (fix->need), then sums over x_i 2^k; its merge
// kernel clears the request.
template <typename T, int OP, int BLOCK, int UNROLL, bool VEC, bool NT = false, bool RESCALE = false>
__global__ __launch_bounds__(BLOCK) void k_reduce(size_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                  T* __restrict__ part, T* __restrict__ out,
                                                  const int* __restrict__ skip, nrm_fix* __restrict__ fix)
{
    using tr = red_traits<OP>;
    static_assert(!RESCALE || tr::nrm, "only the norms are rescaled");
    constexpr bool TRACK = tr::nrm && !RESCALE;   // take M = max |x_i| on the way
    // solver graphs pass a device flag: once it is set the launch does nothing
    if (skip != nullptr && *skip) return;
    int k = 0;
    bool inf_mode = false;
    if constexpr (RESCALE) {
        if (fix->need == 0) return;
        const double am = fix->amax;
        if (am > 1.7976931348623157e308) {
            inf_mode = true;
        } else {
            (void)__builtin_frexp(am, &k);   // am = f 2^k, f in [1/2, 1)
            k = -k;
        }
    }
    using V = std::conditional_t<VEC, typename dev::vec16<T>::type, T>;
    constexpr int W = VEC ? dev::vec16<T>::W : 1;
    constexpr size_t TILE = (size_t)BLOCK * UNROLL;

    const unsigned G = gridDim.x;
    const unsigned b = blockIdx.x;
    const unsigned tid = threadIdx.x;
    const size_t nv = n / W;
    const size_t nfull = nv / TILE;
    const V* __restrict__ xv = reinterpret_cast<const V*>(x);
    const V* __restrict__ yv = reinterpret_cast<const V*>(y);

    T s = T(0), c = T(0), m = T(0);
    auto acc = [&](T xe, T ye) {
        if constexpr (TRACK) m = maxv<T>(m, absv<T>(xe));
        if constexpr (RESCALE) xe = ldexpv<T>(xe, k);
        red_acc<T, OP>(s, c, xe, ye);
    };
    for (size_t tile = b; tile < nfull; tile += G) {
        const size_t base = tile * TILE + tid;
        V xa[UNROLL], ya[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            if constexpr (NT) {
                xa[u] = dev::load_nt<V>(xv + base + (size_t)u * BLOCK);
                if constexpr (tr::binary) ya[u] = dev::load_nt<V>(yv + base + (size_t)u * BLOCK);
            } else {
                xa[u] = xv[base + (size_t)u * BLOCK];
                if constexpr (tr::binary) ya[u] = yv[base + (size_t)u * BLOCK];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if constexpr (tr::binary) acc(lane_of<T>(xa[u], w), lane_of<T>(ya[u], w));
                else acc(lane_of<T>(xa[u], w), T(0));
            }
        }
    }
    // The partial tile goes to the block whose turn it is in the grid-stride
    // sequence; the < W trailing scalars of the vector path follow it.
    if (b == (unsigned)(nfull % G)) {
        const size_t base = nfull * TILE + tid;
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const size_t vi = base + (size_t)u * BLOCK;
            if (vi < nv) {
                V a = xv[vi];
                V bb = a;
                if constexpr (tr::binary) bb = yv[vi];
#pragma unroll
                for (int w = 0; w < W; ++w) acc(lane_of<T>(a, w), lane_of<T>(bb, w));
            }
        }
        if constexpr (VEC) {
            const size_t r0 = nv * W;
            if (tid < n - r0) acc(x[r0 + tid], tr::binary ? y[r0 + tid] : T(0));
        }
    }

    dev::block_reduce<T, tr::comp, BLOCK>(s, c);
    if constexpr (TRACK) m = block_max<T, BLOCK>(m);
    if (G == 1) {
        if (tid == 0) {
            if constexpr (TRACK) nrm_finish<T, OP>(out, s, c, m, fix);
            else if constexpr (RESCALE) nrm_rescaled_out<T>(out, s, c, k, inf_mode, fix);
            else write_out<T, OP>(out, s, c);
        }
        return;
    }
    // G > 1: the per-block partials; k_reduce_fin, launched next on the same stream,
    // merges them (the kernel boundary makes them visible to it)
    if (tid == 0) {
        part[2 * b] = s;
        part[2 * b + 1] = c;
        if constexpr (TRACK) part[2 * G + b] = m;
    }
}

// Fixed-order final stage of a G > 1 reduction: one block of the reduction's BLOCK
// threads merges the G partials in block-index order (thread t takes t, t + BLOCK,
// ...; then the block reduction), so the result does not depend on the order in
// which the blocks of k_reduce ran.  A separate launch instead of a last-block
// ticket: on MI355X the agent-scope release each block's ticket needs cost about
// 7 us per 300 us Dot2 at N = 2^27, the second launch about 2 us.
template <typename T, int OP, int BLOCK, bool RESCALE = false>
__global__ __launch_bounds__(BLOCK) void k_reduce_fin(unsigned G, const T* __restrict__ part, T* __restrict__ out,
                                                      const int* __restrict__ skip, nrm_fix* __restrict__ fix)
{
    using tr = red_traits<OP>;
    constexpr bool TRACK = tr::nrm && !RESCALE;
    if (skip != nullptr && *skip) return;
    int k = 0;
    bool inf_mode = false;
    if constexpr (RESCALE) {
        if (fix->need == 0) return;
        const double am = fix->amax;
        if (am > 1.7976931348623157e308) {
            inf_mode = true;
        } else {
            (void)__builtin_frexp(am, &k);
            k = -k;
        }
    }
    T fs, fc;
    final_merge<T, tr::comp, BLOCK>(part, G, fs, fc);
    T fm = T(0);
    if constexpr (TRACK) {
        for (unsigned i = threadIdx.x; i < G; i += BLOCK) fm = maxv<T>(fm, part[2 * G + i]);
        fm = block_max<T, BLOCK>(fm);
    }
    if (threadIdx.x == 0) {
        if constexpr (TRACK) nrm_finish<T, OP>(out, fs, fc, fm, fix);
        else if constexpr (RESCALE) nrm_rescaled_out<T>(out, fs, fc, k, inf_mode, fix);
        else write_out<T, OP>(out, fs, fc);
    }
}

// ---------------------------------------------------------------------------
// min and max (gpu_reduction::min_max, source/common/gpu_reduction.h:61-68,
// gpu_reduction_impl.cuh:82-264): NaN entries are skipped, as the reference's
// comparisons skip them; -0 orders below +0, so the result does not depend on the
// order of the merges.  Same tile walk and partials as the one-input reductions;
// the block that takes the last ticket merges the partials (a blocking call, whose
// copy back outweighs the ticket); out2 = { min, max }, { +inf, -inf } for an empty x.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T min_ord(T a, T b)
{
    if (b != b) return a;
    if (a != a) return b;
    if (a < b) return a;
    if (b < a) return b;
    return __builtin_signbit(a) ? a : b;
}
template <typename T>
__device__ __forceinline__ T max_ord(T a, T b)
{
    if (b != b) return a;
    if (a != a) return b;
    if (a > b) return a;
    if (b > a) return b;
    return __builtin_signbit(a) ? b : a;
}

template <typename T, int BLOCK>
__device__ __forceinline__ void block_minmax(T& lo, T& hi)
{
    constexpr int NW = BLOCK / 64;
    __shared__ T lds_lo[NW], lds_hi[NW];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min_ord<T>(lo, dev::shfl_xor<T>(lo, d));
        hi = max_ord<T>(hi, dev::shfl_xor<T>(hi, d));
    }
    if (NW > 1) {
        if ((threadIdx.x & 63) == 0) {
            lds_lo[threadIdx.x >> 6] = lo;
            lds_hi[threadIdx.x >> 6] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 1; w < NW; ++w) {
                lo = min_ord<T>(lo, lds_lo[w]);
                hi = max_ord<T>(hi, lds_hi[w]);
            }
        __syncthreads();
    }
}

template <typename T, int BLOCK, int UNROLL, bool VEC>
__global__ __launch_bounds__(BLOCK) void k_minmax(size_t n, const T* __restrict__ x, T* __restrict__ part,
                                                  unsigned* __restrict__ ticket, T* __restrict__ out)
{
    using V = std::conditional_t<VEC, typename dev::vec16<T>::type, T>;
    constexpr int W = VEC ? dev::vec16<T>::W : 1;
    constexpr size_t TILE = (size_t)BLOCK * UNROLL;
    const unsigned G = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
    const size_t nv = n / W;
    const size_t nfull = nv / TILE;
    const V* __restrict__ xv = reinterpret_cast<const V*>(x);
    T lo = __builtin_huge_val(), hi = -__builtin_huge_val();
    auto acc = [&](T v) {
        lo = min_ord<T>(lo, v);
        hi = max_ord<T>(hi, v);
    };
    for (size_t tile = b; tile < nfull; tile += G) {
        const size_t base = tile * TILE + tid;
        V xa[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) xa[u] = xv[base + (size_t)u * BLOCK];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int w = 0; w < W; ++w) acc(lane_of<T>(xa[u], w));
    }
    if (b == (unsigned)(nfull % G)) {
        const size_t base = nfull * TILE + tid;
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const size_t vi = base + (size_t)u * BLOCK;
            if (vi < nv) {
                V a = xv[vi];
#pragma unroll
                for (int w = 0; w < W; ++w) acc(lane_of<T>(a, w));
            }
        }
        if constexpr (VEC) {
            const size_t r0 = nv * W;
            if (tid < n - r0) acc(x[r0 + tid]);
        }
    }
    block_minmax<T, BLOCK>(lo, hi);
    if (G == 1) {
        if (tid == 0) {
            out[0] = lo;
            out[1] = hi;
        }
        return;
    }
    if (tid == 0) {
        st_agent<T>(&part[2 * b], lo);
        st_agent<T>(&part[2 * b + 1], hi);
    }
    if (!take_ticket(ticket, G)) return;
    T flo = __builtin_huge_val(), fhi = -__builtin_huge_val();
    for (unsigned i = tid; i < G; i += BLOCK) {
        flo = min_ord<T>(flo, ld_agent<T>(&part[2 * i]));
        fhi = max_ord<T>(fhi, ld_agent<T>(&part[2 * i + 1]));
    }
    block_minmax<T, BLOCK>(flo, fhi);
    if (tid == 0) {
        out[0] = flo;
        out[1] = fhi;
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename T> struct dtype_id;
template <> struct dtype_id<double> { static constexpr int v = 0; };
template <> struct dtype_id<float> { static constexpr int v = 1; };

struct red_shape {
    int block, unroll, grid, vecw, flags;
};

template <typename T>
static red_shape reduce_shape(const hpb_handle_s* h, size_t n, bool aligned, bool binary)
{
    const hpb_tuning_t& t = binary ? h->tune_red[dtype_id<T>::v] : h->tune_red1[dtype_id<T>::v];
    red_shape r;
    r.block = t.block;
    r.unroll = t.unroll;
    r.flags = t.flags;
    r.vecw = aligned ? dev::vec16<T>::W : 1;
    const size_t nv = n / (size_t)r.vecw;
    const size_t tile = (size_t)t.block * (size_t)t.unroll;
    size_t ntiles = (nv + tile - 1) / tile;
    if (ntiles < 1) ntiles = 1;
    r.grid = (int)(ntiles < (size_t)t.grid_max ? ntiles : (size_t)t.grid_max);
    return r;
}

template <typename T, int OP, int BLOCK, int UNROLL, bool RS>
static void launch_bu(const red_shape& sh, hipStream_t st, size_t n, const T* x, const T* y, T* part, T* out,
                      const int* skip, nrm_fix* fix)
{
    if (sh.vecw > 1 && (sh.flags & HPB_TUNE_NONTEMPORAL) && UNROLL >= 2)
        hipLaunchKernelGGL((k_reduce<T, OP, BLOCK, UNROLL, true, true, RS>), dim3(sh.grid), dim3(BLOCK), 0, st, n, x,
                           y, part, out, skip, fix);
    else if (sh.vecw > 1)
        hipLaunchKernelGGL((k_reduce<T, OP, BLOCK, UNROLL, true, false, RS>), dim3(sh.grid), dim3(BLOCK), 0, st, n,
                           x, y, part, out, skip, fix);
    else
        hipLaunchKernelGGL((k_reduce<T, OP, BLOCK, UNROLL, false, false, RS>), dim3(sh.grid), dim3(BLOCK), 0, st, n,
                           x, y, part, out, skip, fix);
    if (sh.grid > 1)
        hipLaunchKernelGGL((k_reduce_fin<T, OP, BLOCK, RS>), dim3(1), dim3(BLOCK), 0, st, (unsigned)sh.grid,
                           (const T*)part, out, skip, fix);
}

template <typename T, int OP, bool RS = false>
static void launch_red(const red_shape& sh, hipStream_t st, size_t n, const T* x, const T* y, T* part, T* out,
                       const int* skip, nrm_fix* fix)
{
    switch (sh.block * 16 + sh.unroll) {
    case 256 * 16 + 1: launch_bu<T, OP, 256, 1, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 256 * 16 + 2: launch_bu<T, OP, 256, 2, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 256 * 16 + 8: launch_bu<T, OP, 256, 8, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 512 * 16 + 1: launch_bu<T, OP, 512, 1, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 512 * 16 + 2: launch_bu<T, OP, 512, 2, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 512 * 16 + 4: launch_bu<T, OP, 512, 4, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    case 512 * 16 + 8: launch_bu<T, OP, 512, 8, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    default: launch_bu<T, OP, 256, 4, RS>(sh, st, n, x, y, part, out, skip, fix); break;
    }
}

static inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// dev_out != nullptr: asynchronous, result left in device memory.
// otherwise: result copied to host_out (4 values), stream synchronised.
template <typename T, int OP>
static int reduce_entry(hpb_handle_t h, size_t n, const T* x, const T* y, T* dev_out, T* host_out,
                        const int* skip = nullptr)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h, HPB_ERR_INVALID, "reduction: null handle");
    HPB_REQUIRE(n == 0 || x, HPB_ERR_INVALID, "reduction: null x");
    HPB_REQUIRE(!red_traits<OP>::binary || n == 0 || y, HPB_ERR_INVALID, "reduction: null y");
    HPB_REQUIRE(dev_out || host_out, HPB_ERR_INVALID, "reduction: null output");
    const bool al = aligned16(x) && (!red_traits<OP>::binary || aligned16(y));
    red_shape sh = reduce_shape<T>(h, n, al, red_traits<OP>::binary);
    T* out = dev_out ? dev_out : reinterpret_cast<T*>(h->dres);
    T* part = reinterpret_cast<T*>(h->part);
    nrm_fix* fix = reinterpret_cast<nrm_fix*>(h->nrmfix);
    launch_red<T, OP>(sh, h->stream, n, x, y, part, out, skip, fix);
    HPB_LAUNCH_CHECK();
    if constexpr (red_traits<OP>::nrm) {
        // asynchronous: the rescaling launch is always queued (it returns at once
        // unless the norm launch asked for it); blocking: only when asked
        if (dev_out) {
            launch_red<T, OP, true>(sh, h->stream, n, x, y, part, out, skip, fix);
            HPB_LAUNCH_CHECK();
            return HPB_OK;
        }
        HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 4 * sizeof(T), hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipMemcpyAsync(reinterpret_cast<char*>(h->hres) + 64, h->nrmfix, sizeof(nrm_fix),
                               hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipStreamSynchronize(h->stream));
        const nrm_fix* hf = reinterpret_cast<const nrm_fix*>(reinterpret_cast<const char*>(h->hres) + 64);
        if (hf->need) {
            launch_red<T, OP, true>(sh, h->stream, n, x, y, part, out, skip, fix);
            HPB_LAUNCH_CHECK();
            HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 4 * sizeof(T), hipMemcpyDeviceToHost, h->stream));
            HPB_HIP(hipStreamSynchronize(h->stream));
        }
        const T* r = reinterpret_cast<const T*>(h->hres);
        for (int i = 0; i < 4; ++i) host_out[i] = r[i];
        return HPB_OK;
    }
    if (!dev_out) {
        HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 4 * sizeof(T), hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipStreamSynchronize(h->stream));
        const T* r = reinterpret_cast<const T*>(h->hres);
        for (int i = 0; i < 4; ++i) host_out[i] = r[i];
    }
    return HPB_OK;
}

template <typename T>
static int minmax_entry(hpb_handle_t h, size_t n, const T* x, T* host_out2)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && host_out2 && (n == 0 || x), HPB_ERR_INVALID, "hpb_minmax: null argument");
    const bool al = aligned16(x);
    // 512 threads x 4 16-byte vectors per tile, at most one block per CU
    const size_t nv = n / (size_t)(al ? dev::vec16<T>::W : 1);
    size_t grid = (nv + 2047) / 2048;
    const size_t gmax = h->num_cu > 0 ? (size_t)h->num_cu : 256;
    grid = grid < 1 ? 1 : (grid > gmax ? gmax : grid);
    T* out = reinterpret_cast<T*>(h->dres);
    T* part = reinterpret_cast<T*>(h->part);
    const dim3 g((unsigned)grid), blk(512);
    if (al) hipLaunchKernelGGL((k_minmax<T, 512, 4, true>), g, blk, 0, h->stream, n, x, part, h->ticket, out);
    else hipLaunchKernelGGL((k_minmax<T, 512, 4, false>), g, blk, 0, h->stream, n, x, part, h->ticket, out);
    HPB_LAUNCH_CHECK();
    HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 2 * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    host_out2[0] = reinterpret_cast<const T*>(h->hres)[0];
    host_out2[1] = reinterpret_cast<const T*>(h->hres)[1];
    return HPB_OK;
}

// Internal asynchronous launchers for the solver graphs (see hpb_internal.hpp).
template <typename T>
int dot2_async(hpb_handle_t h, size_t n, const T* x, const T* y, T* dev_out4, bool comp, const int* skip)
{
    return comp ? reduce_entry<T, R_DOT2>(h, n, x, y, dev_out4, nullptr, skip)
                : reduce_entry<T, R_DOT>(h, n, x, y, dev_out4, nullptr, skip);
}
template <typename T>
int nrm2_async(hpb_handle_t h, size_t n, const T* x, T* dev_out4, bool comp, const int* skip)
{
    return comp ? reduce_entry<T, R_NRM2C>(h, n, x, nullptr, dev_out4, nullptr, skip)
                : reduce_entry<T, R_NRM2>(h, n, x, nullptr, dev_out4, nullptr, skip);
}
template int dot2_async<double>(hpb_handle_t, size_t, const double*, const double*, double*, bool, const int*);
template int dot2_async<float>(hpb_handle_t, size_t, const float*, const float*, float*, bool, const int*);
template int nrm2_async<double>(hpb_handle_t, size_t, const double*, double*, bool, const int*);
template int nrm2_async<float>(hpb_handle_t, size_t, const float*, float*, bool, const int*);

}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_reduce_config(hpb_handle_t h, int dtype, size_t n, int aligned, int binary, int* cfg4)
{
    HPB_REQUIRE(h && cfg4, HPB_ERR_INVALID, "hpb_reduce_config: null argument");
    HPB_REQUIRE(dtype == 0 || dtype == 1, HPB_ERR_INVALID, "hpb_reduce_config: dtype must be 0 or 1");
    red_shape s = dtype == 0 ? reduce_shape<double>(h, n, aligned != 0, binary != 0)
                             : reduce_shape<float>(h, n, aligned != 0, binary != 0);
    cfg4[0] = s.block;
    cfg4[1] = s.unroll;
    cfg4[2] = s.grid;
    cfg4[3] = s.vecw;
    return HPB_OK;
}

#define HPB_RED_BIN(NAME, OP)                                                                        \
    int hpb_##NAME##_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out4)  \
    {                                                                                                \
        return reduce_entry<double, OP>(h, n, x, y, nullptr, out4);                                  \
    }                                                                                                \
    int hpb_##NAME##_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out4)     \
    {                                                                                                \
        return reduce_entry<float, OP>(h, n, x, y, nullptr, out4);                                   \
    }                                                                                                \
    int hpb_##NAME##_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* d) \
    {                                                                                                \
        HPB_REQUIRE(d, HPB_ERR_INVALID, "async reduction: null device output");                    \
        return reduce_entry<double, OP>(h, n, x, y, d, nullptr);                                     \
    }                                                                                                \
    int hpb_##NAME##_f32_async(hpb_handle_t h, size_t n, const float* x, const float* y, float* d)  \
    {                                                                                                \
        HPB_REQUIRE(d, HPB_ERR_INVALID, "async reduction: null device output");                    \
        return reduce_entry<float, OP>(h, n, x, y, d, nullptr);                                      \
    }

#define HPB_RED_UN(NAME, OP)                                                                         \
    int hpb_##NAME##_f64(hpb_handle_t h, size_t n, const double* x, double* out4)                  \
    {                                                                                                \
        return reduce_entry<double, OP>(h, n, x, nullptr, nullptr, out4);                            \
    }                                                                                                \
    int hpb_##NAME##_f32(hpb_handle_t h, size_t n, const float* x, float* out4)                    \
    {                                                                                                \
        return reduce_entry<float, OP>(h, n, x, nullptr, nullptr, out4);                             \
    }

#define HPB_RED_UN_ASYNC(NAME, OP)                                                                   \
    int hpb_##NAME##_f64_async(hpb_handle_t h, size_t n, const double* x, double* d)               \
    {                                                                                                \
        HPB_REQUIRE(d, HPB_ERR_INVALID, "async reduction: null device output");                    \
        return reduce_entry<double, OP>(h, n, x, nullptr, d, nullptr);                               \
    }                                                                                                \
    int hpb_##NAME##_f32_async(hpb_handle_t h, size_t n, const float* x, float* d)                 \
    {                                                                                                \
        HPB_REQUIRE(d, HPB_ERR_INVALID, "async reduction: null device output");                    \
        return reduce_entry<float, OP>(h, n, x, nullptr, d, nullptr);                                \
    }

HPB_RED_BIN(dot2, R_DOT2)
HPB_RED_BIN(dot, R_DOT)
HPB_RED_UN(sum2, R_SUM2)
HPB_RED_UN(sum, R_SUM)
HPB_RED_UN(asum2, R_ASUM2)
HPB_RED_UN(asum, R_ASUM)
HPB_RED_UN(nrm2c, R_NRM2C)
HPB_RED_UN(nrm2, R_NRM2)
HPB_RED_UN_ASYNC(nrm2c, R_NRM2C)
HPB_RED_UN_ASYNC(nrm2, R_NRM2)

int hpb_minmax_f64(hpb_handle_t h, size_t n, const double* x, double* out2)
{
    return minmax_entry<double>(h, n, x, out2);
}
int hpb_minmax_f32(hpb_handle_t h, size_t n, const float* x, float* out2)
{
    return minmax_entry<float>(h, n, x, out2);
}

}  // extern "C"
