// hpb_acc256.hpp -- exact 256-bit fixed-point accumulator (two's complement,
// least significant bit 2^-106), host and device.
//
// Used by the GenDot generator: every generated entry is a multiple of 2^-53, so
// every product x_i*y_i = p + e (TwoProd) is a multiple of 2^-106 and the running
// dot is held exactly.  Integer addition is associative, so partial sums may be
// combined in any order and the result is still bit-reproducible.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define HPB_ACC_HD __host__ __device__ __forceinline__
#else
#define HPB_ACC_HD inline
#endif

namespace hpb {

struct acc256 {
    uint64_t w[4];
};

HPB_ACC_HD void acc_zero(acc256& a)
{
    a.w[0] = a.w[1] = a.w[2] = a.w[3] = 0;
}

HPB_ACC_HD void acc_add(acc256& a, const acc256& b)
{
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        uint64_t s = a.w[i] + b.w[i];
        uint64_t c1 = s < a.w[i];
        uint64_t s2 = s + carry;
        uint64_t c2 = s2 < s;
        a.w[i] = s2;
        carry = c1 | c2;
    }
}

HPB_ACC_HD void acc_neg(acc256& a)
{
    uint64_t carry = 1;
    for (int i = 0; i < 4; ++i) {
        uint64_t v = ~a.w[i] + carry;
        carry = (carry && v == 0) ? 1 : 0;
        a.w[i] = v;
    }
}

HPB_ACC_HD uint64_t dbits(double v)
{
#ifdef __HIP_DEVICE_COMPILE__
    return (uint64_t)__double_as_longlong(v);
#else
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return b;
#endif
}

// Add v exactly; v must be a multiple of 2^-106 with |v| < 2^149.
HPB_ACC_HD void acc_add_double(acc256& a, double v)
{
    if (v == 0.0) return;
    const uint64_t bits = dbits(v);
    const bool neg = (bits >> 63) != 0;
    const int ex = (int)((bits >> 52) & 0x7ff);
    uint64_t m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
    int sh = ex - 1075 + 106;
    if (sh < 0) {
        m >>= -sh;
        sh = 0;
    }
    acc256 t;
    acc_zero(t);
    const int li = sh >> 6, bs = sh & 63;
    t.w[li] = m << bs;
    if (bs && li + 1 < 4) t.w[li + 1] = m >> (64 - bs);
    if (neg) acc_neg(t);
    acc_add(a, t);
}

HPB_ACC_HD int clz64(uint64_t v)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __clzll((long long)v);
#else
    return __builtin_clzll(v);
#endif
}

// Nearest double, ties to even.
HPB_ACC_HD double acc_round(const acc256& a)
{
    acc256 m = a;
    const bool neg = (a.w[3] >> 63) != 0;
    if (neg) acc_neg(m);
    int top = -1;
    for (int i = 3; i >= 0; --i)
        if (m.w[i]) {
            top = i;
            break;
        }
    if (top < 0) return 0.0;
    const int msb = top * 64 + 63 - clz64(m.w[top]);
    int ulp = msb - 52;
    if (ulp < 0) ulp = 0;
    // q = m >> ulp (at most 53 bits)
    uint64_t q;
    {
        const int li = ulp >> 6, bs = ulp & 63;
        q = m.w[li] >> bs;
        if (bs && li + 1 < 4) q |= m.w[li + 1] << (64 - bs);
        const int nb = msb - ulp + 1;
        if (nb < 64) q &= (1ull << nb) - 1;
    }
    int half = 0, sticky = 0;
    if (ulp > 0) {
        const int hb = ulp - 1;
        half = (int)((m.w[hb >> 6] >> (hb & 63)) & 1ull);
        // any bit below hb
        const int li = hb >> 6, bs = hb & 63;
        if (bs && (m.w[li] & ((1ull << bs) - 1))) sticky = 1;
        for (int i = 0; i < li && !sticky; ++i)
            if (m.w[i]) sticky = 1;
    }
    if (half && (sticky || (q & 1))) q += 1;
    double r = (double)q;
#ifdef __HIP_DEVICE_COMPILE__
    r = ldexp(r, ulp - 106);
#else
    r = __builtin_ldexp(r, ulp - 106);
#endif
    return neg ? -r : r;
}

}  // namespace hpb
