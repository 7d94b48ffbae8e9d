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
//   * each thread streams its elements into a small floating-point expansion of
//     K doubles (a TwoSum cascade, error-free); what falls out of the last
//     component is added to the block's accumulator in LDS with 64-bit integer
//     atomics, and the expansion itself is added at the end;
//   * products are split exactly with TwoProd (p + e = x*y); FP32 products are
//     exact in FP64 and added as one double;
//   * each block carries its accumulator to 32-bit digits and stores it; the block
//     that takes the last ticket adds the digits of all blocks (integer sums, so
//     the order does not matter), carries, and rounds once to nearest even.
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
constexpr int XP = 80;    // words per block record: XL digits, flag, padding
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

template <int K>
__device__ __forceinline__ void fpe_add(double (&a)[K], double v, long long* acc)
{
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double t;
        a[j] = dev::two_sum<double>(a[j], v, t);
        v = t;
    }
    if (v != 0.0) acc_add<true>(acc, v);
}

constexpr double TWO_M969 = 0x1p-969;
constexpr double TWO_P960 = 0x1p+960;

template <int K, int MODE>
__device__ __forceinline__ void x_element(double (&a)[K], long long* acc, int& flag, const void* xv, const void* yv,
                                          size_t i)
{
    if constexpr (MODE == X_DOT64) {
        const double x = static_cast<const double*>(xv)[i];
        const double y = static_cast<const double*>(yv)[i];
        const double p = x * y;
        const double e = __builtin_fma(x, y, -p);
        const double ap = __builtin_fabs(p);
        if (!(ap < TWO_P960)) {
            flag |= (__builtin_isfinite(x) && __builtin_isfinite(y)) ? XF_RANGE : XF_NONFINITE;
            return;
        }
        if (ap < TWO_M969 && x != 0.0 && y != 0.0) flag |= XF_RANGE;
        fpe_add<K>(a, p, acc);
        fpe_add<K>(a, e, acc);
    } else if constexpr (MODE == X_DOT32) {
        const double x = (double)static_cast<const float*>(xv)[i];
        const double y = (double)static_cast<const float*>(yv)[i];
        const double p = x * y;   // 24 x 24 bits: exact
        if (!__builtin_isfinite(p)) {
            flag |= XF_NONFINITE;
            return;
        }
        fpe_add<K>(a, p, acc);
    } else {
        const double x = static_cast<const double*>(xv)[i];
        if (!(__builtin_fabs(x) < TWO_P960)) {
            flag |= __builtin_isfinite(x) ? XF_RANGE : XF_NONFINITE;
            return;
        }
        fpe_add<K>(a, x, acc);
    }
}

__device__ __forceinline__ void st_agent_ll(long long* p, long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ld_agent_ll(const long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// out[0] = hi, out[1] = lo, out[2] = flags (as a double).
template <int K, int MODE>
__global__ __launch_bounds__(XBLOCK) void k_exact(size_t n, const void* __restrict__ xv, const void* __restrict__ yv,
                                                  long long* __restrict__ part, unsigned* __restrict__ ticket,
                                                  double* __restrict__ out)
{
    __shared__ long long acc[XL];
    __shared__ int sflag;
    __shared__ int am_last;
    __shared__ unsigned mag[XL];
    for (int i = threadIdx.x; i < XL; i += XBLOCK) acc[i] = 0;
    if (threadIdx.x == 0) sflag = 0;
    __syncthreads();

    double a[K];
#pragma unroll
    for (int j = 0; j < K; ++j) a[j] = 0.0;
    int flag = 0;
    const size_t stride = (size_t)gridDim.x * XBLOCK;
    size_t i = (size_t)blockIdx.x * XBLOCK + threadIdx.x;
    // two elements in flight per thread
    for (; i + stride < n; i += 2 * stride) {
        x_element<K, MODE>(a, acc, flag, xv, yv, i);
        x_element<K, MODE>(a, acc, flag, xv, yv, i + stride);
    }
    if (i < n) x_element<K, MODE>(a, acc, flag, xv, yv, i);
#pragma unroll
    for (int j = 0; j < K; ++j) acc_add<true>(acc, a[j]);
    if (flag) atomicOr(&sflag, flag);
    __syncthreads();

    const unsigned G = gridDim.x;
    long long* rec = part + (size_t)blockIdx.x * XP;
    if (threadIdx.x == 0) acc_carry(acc);
    __syncthreads();
    for (int w = threadIdx.x; w < XL; w += XBLOCK) st_agent_ll(&rec[w], acc[w]);
    if (threadIdx.x == 0) st_agent_ll(&rec[XL], (long long)sflag);
    // every wave's stores complete before the barrier (the barrier alone does not wait
    // for them), so the ticket's release below covers the whole record
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == G - 1);
    }
    __syncthreads();
    if (!am_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    // Digit sums over all blocks: each carried digit is < 2^32 (top word small
    // and signed), so G <= 2^20 blocks cannot overflow int64.
    for (int w = threadIdx.x; w <= XL; w += XBLOCK) {
        long long s = 0;
        if (w < XL) {
            for (unsigned b = 0; b < G; ++b) s += ld_agent_ll(&part[(size_t)b * XP + w]);
            acc[w] = s;
        } else {
            for (unsigned b = 0; b < G; ++b) s |= ld_agent_ll(&part[(size_t)b * XP + XL]);
            sflag = (int)s;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double hi = __builtin_nan(""), lo = __builtin_nan("");
        if (sflag == 0) {
            acc_carry(acc);
            hi = acc_round(acc, mag);
            if (__builtin_isfinite(hi)) {
                acc_add<false>(acc, -hi);
                acc_carry(acc);
                lo = acc_round(acc, mag);
            } else {
                lo = 0.0;
            }
        }
        out[0] = hi;
        out[1] = lo;
        out[2] = (double)sflag;
    }
}

constexpr int XK = 4;   // expansion length per thread

template <int MODE>
int exact_launch(hpb_handle_t h, size_t n, const void* x, const void* y, double* dev_out3)
{
    size_t want = (n + 2 * (size_t)XBLOCK * 8 - 1) / (2 * (size_t)XBLOCK * 8);   // >= 8 elements per thread
    if (want < 1) want = 1;
    int grid = (int)(want < (size_t)XGRID_MAX ? want : (size_t)XGRID_MAX);
    if ((size_t)grid * XP * sizeof(long long) > h->part_bytes) grid = (int)(h->part_bytes / (XP * sizeof(long long)));
    hipLaunchKernelGGL((k_exact<XK, MODE>), dim3(grid), dim3(XBLOCK), 0, h->stream, n, x, y,
                       reinterpret_cast<long long*>(h->part), h->ticket + 2, dev_out3);
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
