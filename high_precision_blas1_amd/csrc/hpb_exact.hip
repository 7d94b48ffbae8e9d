// hpb_exact.hip -- exact dot product and exact sum on the device.
//
// Replaces the reference's exact oracles: the host GMP loop of
// dot_product_gmp::dot_exact (source/high_prec/dot_product_gmp.hpp:108-130) and its
// GPU form through CUMP (source/high_prec/dot_product_cump.hpp:134-151,
// cump_blas_kernels.cuh:381-444), both of which form sum x_j*y_j in 512-1024-bit
// floating point and return the value as one double (mpf get_d, truncating).
// Here the value is exact and returned as (hi, lo): hi = RN(exact), lo = RN(exact - hi).
//
// Design:
//   * a fixed-point long accumulator of XL 32-bit digits held in int64 words;
//     bit 0 is 2^-1074, so every finite double is an integer multiple of it;
//   * each thread streams its elements into two floating-point expansions of K
//     doubles (TwoSum cascades, error-free; two, so that the dependency chains of
//     consecutive terms interleave); what falls out of the last component is added
//     to the block's accumulator in LDS with 64-bit integer atomics (for the
//     generator's pairs up to kappa 1e16 this never happens at K = 6), and the
//     expansions themselves are added at the end;
//   * products are split exactly with TwoProd (p + e = x*y); FP32 products are
//     exact in FP64 and added as one double;
//   * each block carries its accumulator to 32-bit digits and stores it; a second
//     launch adds the digits of all blocks (one block per digit; integer sums, so
//     the order does not matter), and its last block carries and rounds once to
//     nearest even.  (Formerly the last block of the first launch added all the records
//     itself, one digit per thread over every block in turn: at 2^27 elements
//     that serial walk was half the time.)
// Domain (reported, never silently wrong): finite inputs; for FP64 dots every
// nonzero product has |x*y| >= 2^-969 (below that TwoProd's error term can
// underflow) and every |x*y| (|x| for sums) < 2^960.  Out-of-domain inputs make the
// call fail with a flag instead of a value.
#include <hip/hip_runtime.h>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

namespace hpb {
namespace {

constexpr int XL = 72;    // 72 * 32 = 2304 bits >= 2098 + carry headroom
constexpr int XF_NONFINITE = 1;
constexpr int XF_RANGE = 2;
constexpr int XBLOCK = 256;
constexpr int XGRID_MAX = 1024;

enum { X_DOT64 = 0, X_DOT32 = 1, X_SUM64 = 2 };

// Add the double v (finite) to the digit array: v = m * 2^(off - 1074) with a
// 53-bit integer m placed at bit off.  ATOMIC selects LDS atomics (shared by the
// block) or plain adds (single-thread use).
template <bool ATOMIC>
__device__ __forceinline__ void acc_add(long long* acc, double v)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const int ex = (int)((bits >> 52) & 0x7ff);
    if (v == 0.0 || ex == 0x7ff) return;
    unsigned long long m = bits & ((1ull << 52) - 1);
    int off = 0;
    if (ex != 0) {
        m |= 1ull << 52;
        off = ex - 1;
    }
    const int li = off >> 5, sh = off & 31;
    const unsigned long long lo = m << sh;
    const unsigned long long hi = sh ? (m >> (64 - sh)) : 0ull;
    long long p0 = (long long)(lo & 0xffffffffull);
    long long p1 = (long long)(lo >> 32);
    long long p2 = (long long)hi;
    if (bits >> 63) {
        p0 = -p0;
        p1 = -p1;
        p2 = -p2;
    }
    if constexpr (ATOMIC) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&acc[li]), (unsigned long long)p0);
        atomicAdd(reinterpret_cast<unsigned long long*>(&acc[li + 1]), (unsigned long long)p1);
        if (p2) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[li + 2]), (unsigned long long)p2);
    } else {
        acc[li] += p0;
        acc[li + 1] += p1;
        acc[li + 2] += p2;
    }
}

// Carry so that digits 0..XL-2 lie in [0, 2^32); the top word keeps the sign.
__device__ __forceinline__ void acc_carry(long long* d)
{
    long long carry = 0;
    for (int i = 0; i < XL - 1; ++i) {
        const long long v = d[i] + carry;
        carry = v >> 32;   // arithmetic shift: floor division
        d[i] = v - (carry << 32);
    }
    d[XL - 1] += carry;
}

__device__ __forceinline__ int bit_at(const unsigned* mag, int b) { return (int)((mag[b >> 5] >> (b & 31)) & 1u); }

// Round the carried accumulator to the nearest double, ties to even.  The
// magnitude is formed in mag[] (two's complement when negative).
__device__ double acc_round(const long long* d, unsigned* mag)
{
    const bool neg = d[XL - 1] < 0;
    unsigned long long borrow = 1;
    for (int i = 0; i < XL; ++i) {
        const unsigned dig = (unsigned)(d[i] & 0xffffffffll);
        if (neg) {
            const unsigned long long v = (unsigned long long)(~dig) + borrow;
            mag[i] = (unsigned)v;
            borrow = v >> 32;
        } else {
            mag[i] = dig;
        }
    }
    int top = -1;
    for (int i = XL - 1; i >= 0; --i)
        if (mag[i]) {
            top = i;
            break;
        }
    if (top < 0) return 0.0;
    const int msb = top * 32 + (31 - __builtin_clz(mag[top]));
    int ulp = msb - 52;
    if (ulp < 0) ulp = 0;   // subnormal results
    if (msb - 1074 > 1023) return neg ? -__builtin_inf() : __builtin_inf();
    unsigned long long q = 0;
    for (int b = msb; b >= ulp; --b) q = (q << 1) | (unsigned long long)bit_at(mag, b);
    int half = 0, sticky = 0;
    if (ulp > 0) {
        half = bit_at(mag, ulp - 1);
        const int sb = ulp - 1;   // bits below the half bit: [0, sb)
        const int full = sb >> 5;
        for (int i = 0; i < full && !sticky; ++i)
            if (mag[i]) sticky = 1;
        if (!sticky && (sb & 31)) sticky = (mag[full] & ((1u << (sb & 31)) - 1u)) != 0;
    }
    if (half && (sticky || (q & 1ull))) q += 1;
    const double r = ldexp((double)q, ulp - 1074);
    return neg ? -r : r;
}

// Adds v to the expansion a[0..K) with a TwoSum cascade (error-free); what falls
// out of the last term goes to the block's accumulator.  A zero error ends the
// cascade: the remaining terms would not change.
template <int K>
__device__ __forceinline__ void fpe_add(double (&a)[K], double v, long long* acc)
{
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (v == 0.0) return;
        double t;
        a[j] = dev::two_sum<double>(a[j], v, t);
        v = t;
    }
    if (v != 0.0) acc_add<true>(acc, v);
}

constexpr double TWO_M969 = 0x1p-969;
constexpr double TWO_P960 = 0x1p+960;

// One element into the thread's two expansions a[0], a[1]: for FP64 dots the
// product goes to a[0] and its TwoProd error to a[1] (two independent dependency
// chains); FP32 products and summands alternate by element (odd).  The range flags
// are formed without branches and the element is added regardless: a non-finite or
// out-of-range term only spoils a value that its flag then withholds (non-finite
// digits are skipped by acc_add).
template <int K, int MODE>
__device__ __forceinline__ void x_element(double (&a)[2][K], long long* acc, int& flag, double x, double y, int odd)
{
    if constexpr (MODE == X_DOT64) {
        const double p = x * y;
        const double e = __builtin_fma(x, y, -p);
        const double ap = __builtin_fabs(p);
        const int big = !(ap < TWO_P960) ? ((__builtin_isfinite(x) && __builtin_isfinite(y)) ? XF_RANGE : XF_NONFINITE)
                                         : 0;
        const int tiny = (ap < TWO_M969 && x != 0.0 && y != 0.0) ? XF_RANGE : 0;
        flag |= big | tiny;
        fpe_add<K>(a[0], p, acc);
        fpe_add<K>(a[1], e, acc);
    } else if constexpr (MODE == X_DOT32) {
        const double p = x * y;   // 24 x 24 bits: exact
        flag |= __builtin_isfinite(p) ? 0 : XF_NONFINITE;
        fpe_add<K>(a[odd], p, acc);
    } else {
        flag |= !(__builtin_fabs(x) < TWO_P960) ? (__builtin_isfinite(x) ? XF_RANGE : XF_NONFINITE) : 0;
        fpe_add<K>(a[odd], x, acc);
    }
}

template <int MODE> struct x_elem_t { using type = double; };
template <> struct x_elem_t<X_DOT32> { using type = float; };

// Per-block digit records.  Each thread streams its elements (16-byte vector loads
// when VEC, two vectors in flight, their elements formed in an unrolled loop) into
// two K-term expansions; their terms and the rare overflow of the cascades go to the
// block's accumulator in LDS (64-bit integer atomics: the order of the additions
// does not matter).  The block carries its accumulator to 32-bit digits and stores
// them digit-major: part[w G + block] for w < XL, the input flags in
// part[XL G + block].
template <int K, int MODE, bool VEC>
__global__ __launch_bounds__(XBLOCK) void k_exact(size_t n, const void* __restrict__ xv, const void* __restrict__ yv,
                                                  long long* __restrict__ part)
{
    using T = typename x_elem_t<MODE>::type;
    __shared__ long long acc[XL];
    __shared__ int sflag;
    for (int i = threadIdx.x; i < XL; i += XBLOCK) acc[i] = 0;
    if (threadIdx.x == 0) sflag = 0;
    __syncthreads();

    double a[2][K];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j) a[q][j] = 0.0;
    int flag = 0;
    const T* x = static_cast<const T*>(xv);
    const T* y = static_cast<const T*>(yv);   // null for sums
    auto load_y = [&](size_t i) -> double {
        if constexpr (MODE == X_SUM64) return 0.0;
        else return (double)y[i];
    };
    const size_t stride = (size_t)gridDim.x * XBLOCK;
    if constexpr (VEC) {
        using V = typename dev::vec16<T>::type;
        constexpr int W = dev::vec16<T>::W;
        const V* xq = reinterpret_cast<const V*>(x);
        const V* yq = reinterpret_cast<const V*>(y);
        const size_t nv = n / W;
        size_t i = (size_t)blockIdx.x * XBLOCK + threadIdx.x;
        for (; i + stride < nv; i += 2 * stride) {
            // nontemporal: read once (5399 against 5104 GB/s at kappa 1e4, the same at
            // kappa 1e24 and above, where the expansions set the pace)
            const V x0 = dev::load_nt<V>(xq + i), x1 = dev::load_nt<V>(xq + i + stride);
            double xs[2 * W], ys[2 * W];
#pragma unroll
            for (int q = 0; q < W; ++q) {
                xs[q] = (double)dev::vget<T>(x0, q);
                xs[W + q] = (double)dev::vget<T>(x1, q);
                ys[q] = ys[W + q] = 0.0;
            }
            if constexpr (MODE != X_SUM64) {
                const V y0 = dev::load_nt<V>(yq + i), y1 = dev::load_nt<V>(yq + i + stride);
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    ys[q] = (double)dev::vget<T>(y0, q);
                    ys[W + q] = (double)dev::vget<T>(y1, q);
                }
            }
#pragma unroll
            for (int u = 0; u < 2 * W; ++u) x_element<K, MODE>(a, acc, flag, xs[u], ys[u], u & 1);
        }
        if (i < nv) {
            const V x0 = xq[i];
            V y0{};
            if constexpr (MODE != X_SUM64) y0 = yq[i];
#pragma unroll
            for (int q = 0; q < W; ++q)
                x_element<K, MODE>(a, acc, flag, (double)dev::vget<T>(x0, q), (double)dev::vget<T>(y0, q), q & 1);
        }
        const size_t r0 = nv * W;   // the last n mod W elements
        if (blockIdx.x == 0 && threadIdx.x < n - r0)
            x_element<K, MODE>(a, acc, flag, (double)x[r0 + threadIdx.x], load_y(r0 + threadIdx.x), 0);
    } else {
        size_t i = (size_t)blockIdx.x * XBLOCK + threadIdx.x;
        for (; i + stride < n; i += 2 * stride) {
            x_element<K, MODE>(a, acc, flag, (double)x[i], load_y(i), 0);
            x_element<K, MODE>(a, acc, flag, (double)x[i + stride], load_y(i + stride), 1);
        }
        if (i < n) x_element<K, MODE>(a, acc, flag, (double)x[i], load_y(i), 0);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j) acc_add<true>(acc, a[q][j]);
    if (flag) atomicOr(&sflag, flag);
    __syncthreads();
    if (threadIdx.x == 0) acc_carry(acc);
    __syncthreads();
    const unsigned G = gridDim.x;
    for (int w = threadIdx.x; w < XL; w += XBLOCK) part[(size_t)w * G + blockIdx.x] = acc[w];
    if (threadIdx.x == 0) part[(size_t)XL * G + blockIdx.x] = sflag;
}

// ---- block-parallel carry and rounding of the digit sums (the last block of
// k_exact_digits; XL <= blockDim.x, thread w owns word w)

// Carry so that words 0..XL-2 lie in [0, 2^32) and the top word keeps the sign (the
// unique normal form, as acc_carry): rounds in which every word passes its carry to
// the next at once, until none is left (a carry moves one word per round).
__device__ __forceinline__ void acc_carry_par(long long* d, long long* q)
{
    const int w = threadIdx.x;
    for (int round = 0; round < 2 * XL; ++round) {
        const long long c = (w < XL - 1) ? (d[w] >> 32) : 0;   // floor division
        if (w < XL) q[w] = c;
        __syncthreads();
        if (w < XL - 1) d[w] -= c * 4294967296ll;
        if (w > 0 && w < XL) d[w] += q[w - 1];
        __syncthreads();
        const int outside = (w < XL - 1) && (d[w] < 0 || d[w] >= 4294967296ll);
        if (!__syncthreads_or(outside)) break;
    }
}

// RN(value) of the carried words d (value = sum d_w 2^(32 w - 1074)), ties to even;
// d is left unchanged (e, q, mag and sh are scratch in LDS).  Every thread returns it.
__device__ double acc_round_par(const long long* d, long long* e, long long* q, unsigned* mag, int* sh)
{
    const int w = threadIdx.x;
    const bool neg = d[XL - 1] < 0;
    if (w < XL) e[w] = neg ? -d[w] : d[w];   // |value| once carried again
    __syncthreads();
    if (neg) acc_carry_par(e, q);
    if (w < XL) mag[w] = (unsigned)(e[w] & 0xffffffffll);
    if (w == 0) sh[0] = -1;
    __syncthreads();
    if (w < XL && mag[w] != 0u) atomicMax(&sh[0], w);
    __syncthreads();
    const int top = sh[0];
    if (top < 0) return 0.0;
    const int msb = top * 32 + (31 - __builtin_clz(mag[top]));
    const int ulp = msb > 52 ? msb - 52 : 0;   // ulp 0: subnormal or exact, no rounding
    // sticky: any bit below the half bit ulp - 1
    const int sb = ulp > 0 ? ulp - 1 : 0, full = sb >> 5;
    const int below = (w < full && mag[w] != 0u) ? 1 : 0;
    const int sticky_words = __syncthreads_or(below);
    if (msb - 1074 > 1023) return neg ? -__builtin_inf() : __builtin_inf();
    const int a = ulp >> 5, s = ulp & 31, nb = msb - ulp + 1;   // nb <= 53
    const unsigned long long w0 = mag[a], w1 = a + 1 < XL ? mag[a + 1] : 0u, w2 = a + 2 < XL ? mag[a + 2] : 0u;
    const unsigned long long lo64 = w0 | (w1 << 32);
    unsigned long long qv = s == 0 ? lo64 : ((lo64 >> s) | (w2 << (64 - s)));
    qv &= (1ull << nb) - 1ull;
    if (ulp > 0) {
        const int half = (int)((mag[sb >> 5] >> (sb & 31)) & 1u);
        int sticky = sticky_words;
        if (!sticky && (sb & 31)) sticky = (mag[full] & ((1u << (sb & 31)) - 1u)) != 0u;
        if (half && (sticky || (qv & 1ull))) qv += 1;
    }
    const double r = ldexp((double)qv, ulp - 1074);
    return neg ? -r : r;
}

// Digit sums over the G block records, one block per word (word XL: the flags,
// OR-ed).  Each carried digit is < 2^32 (the top word small and signed), so G <= 2^20
// blocks cannot overflow int64; integer sums, so the order does not matter.  The
// block that takes the last ticket (each block's sum published by a release
// increment) carries the sums, rounds to nearest even, subtracts, rounds again:
// out[0] = hi, out[1] = lo, out[2] = flags (as a double).
__global__ __launch_bounds__(256) void k_exact_digits(unsigned G, const long long* __restrict__ part,
                                                      long long* __restrict__ sums, unsigned* __restrict__ ticket,
                                                      double* __restrict__ out)
{
    static_assert(XL <= 256, "one word per thread");
    __shared__ long long red[256];
    __shared__ int am_last;
    __shared__ long long acc[XL], e[XL], q[XL];
    __shared__ unsigned mag[XL];
    __shared__ int sh[1];
    const unsigned w = blockIdx.x;
    const long long* p = part + (size_t)w * G;
    long long s = 0;
    for (unsigned k = threadIdx.x; k < G; k += 256) s = w < (unsigned)XL ? s + p[k] : (s | p[k]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) red[threadIdx.x] = w < (unsigned)XL ? red[threadIdx.x] + red[threadIdx.x + m]
                                                                     : (red[threadIdx.x] | red[threadIdx.x + m]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[w] = red[0];
        am_last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u;
    }
    __syncthreads();
    if (!am_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < (unsigned)XL) acc[threadIdx.x] = sums[threadIdx.x];
    const int sflag = (int)sums[XL];
    __syncthreads();
    double hi = __builtin_nan(""), lo = __builtin_nan("");
    if (sflag == 0) {
        acc_carry_par(acc, q);
        hi = acc_round_par(acc, e, q, mag, sh);
        if (__builtin_isfinite(hi)) {
            __syncthreads();
            if (threadIdx.x == 0) acc_add<false>(acc, -hi);
            __syncthreads();
            acc_carry_par(acc, q);
            lo = acc_round_par(acc, e, q, mag, sh);
        } else {
            lo = 0.0;
        }
    }
    if (threadIdx.x == 0) {
        out[0] = hi;
        out[1] = lo;
        out[2] = (double)sflag;
    }
}

constexpr int XK = 6;   // terms per expansion (two expansions per thread)

template <int MODE>
int exact_launch(hpb_handle_t h, size_t n, const void* x, const void* y, double* dev_out3)
{
    using T = typename x_elem_t<MODE>::type;
    size_t want = (n + (size_t)XBLOCK * 16 - 1) / ((size_t)XBLOCK * 16);   // >= 16 elements per thread
    if (want < 1) want = 1;
    size_t grid = want < (size_t)XGRID_MAX ? want : (size_t)XGRID_MAX;
    // records ((XL + 1) G words) and the digit sums (XL + 1 words) in the workspace
    const size_t gcap = h->part_bytes / ((XL + 1) * sizeof(long long)) - 1;
    if (grid > gcap) grid = gcap;
    const unsigned G = (unsigned)grid;
    long long* part = reinterpret_cast<long long*>(h->part);
    long long* sums = part + (size_t)(XL + 1) * G;
    const bool vec = ((uintptr_t)x & 15u) == 0 && (MODE == X_SUM64 || ((uintptr_t)y & 15u) == 0);
    if (vec)
        hipLaunchKernelGGL((k_exact<XK, MODE, true>), dim3(G), dim3(XBLOCK), 0, h->stream, n, x, y, part);
    else
        hipLaunchKernelGGL((k_exact<XK, MODE, false>), dim3(G), dim3(XBLOCK), 0, h->stream, n, x, y, part);
    hipLaunchKernelGGL(k_exact_digits, dim3(XL + 1), dim3(256), 0, h->stream, G, part, sums, h->ticket + 2, dev_out3);
    (void)sizeof(T);
    return HPB_OK;
}

template <int MODE>
int exact_entry(hpb_handle_t h, size_t n, const void* x, const void* y, double* out2, bool async)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && out2, HPB_ERR_INVALID, "exact: null argument");
    HPB_REQUIRE(n == 0 || (x && (MODE == X_SUM64 || y)), HPB_ERR_INVALID, "exact: null vector");
    // per-block element counts stay far below 2^30 (integer headroom of the digits)
    HPB_REQUIRE(n < (1ull << 38), HPB_ERR_INVALID, "exact: n must be < 2^38");
    double* dres = async ? out2 : static_cast<double*>(h->dres);
    if (n == 0) {
        if (async) {
            const double z[3] = {0.0, 0.0, 0.0};
            HPB_HIP(hipMemcpyAsync(out2, z, sizeof z, hipMemcpyHostToDevice, h->stream));
            HPB_HIP(hipStreamSynchronize(h->stream));
        } else {
            out2[0] = out2[1] = 0.0;
        }
        return HPB_OK;
    }
    exact_launch<MODE>(h, n, x, y, dres);
    HPB_LAUNCH_CHECK();
    if (async) return HPB_OK;
    double r[3];
    HPB_HIP(hipMemcpyAsync(h->hres, dres, sizeof r, hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    const double* hr = static_cast<const double*>(h->hres);
    out2[0] = hr[0];
    out2[1] = hr[1];
    const int fl = (int)hr[2];
    HPB_REQUIRE(!(fl & XF_NONFINITE), HPB_ERR_INVALID, "exact: non-finite input");
    HPB_REQUIRE(!(fl & XF_RANGE), HPB_ERR_UNSUPPORTED,
                "exact: a product (or element) lies outside the exact range [2^-969, 2^960)");
    return HPB_OK;
}

}  // namespace
}  // namespace hpb

extern "C" {

int hpb_exact_dot_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out2)
{
    return hpb::exact_entry<hpb::X_DOT64>(h, n, x, y, out2, false);
}
int hpb_exact_dot_f32(hpb_handle_t h, size_t n, const float* x, const float* y, double* out2)
{
    return hpb::exact_entry<hpb::X_DOT32>(h, n, x, y, out2, false);
}
int hpb_exact_sum_f64(hpb_handle_t h, size_t n, const double* x, double* out2)
{
    return hpb::exact_entry<hpb::X_SUM64>(h, n, x, nullptr, out2, false);
}
int hpb_exact_dot_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out3)
{
    return hpb::exact_entry<hpb::X_DOT64>(h, n, x, y, dev_out3, true);
}

}  // extern "C"
