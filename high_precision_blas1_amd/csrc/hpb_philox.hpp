// hpb_philox.hpp -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on host and
// device so that seeded vectors are bit-identical wherever they are made.
// No reference counterpart: the reference draws from cuRAND seeded by
// std::random_device (gpu_vector_operations.h:496-512), which is not reproducible.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define HPB_HD __host__ __device__ __forceinline__
#else
#define HPB_HD inline
#endif

namespace hpb {

struct philox_out {
    uint32_t v[4];
};

HPB_HD philox_out philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    philox_out o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

// 53 random bits from two words.
HPB_HD uint64_t k53(uint32_t hi, uint32_t lo) { return ((((uint64_t)hi) << 32) | lo) >> 11; }
// uniform in [0, 1)
HPB_HD double unif53(uint64_t k) { return (double)k * 0x1p-53; }
// uniform in (0, 1]
HPB_HD double unif53_open0(uint64_t k) { return (double)(k + 1) * 0x1p-53; }
// 24 random bits -> (0, 1]
HPB_HD float unif24_open0(uint32_t w) { return (float)((w >> 8) + 1u) * 0x1p-24f; }

}  // namespace hpb
