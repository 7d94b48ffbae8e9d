// hpb_device.hpp -- device-side building blocks for gfx950 (wave64).
//
// Error-free transformations and the (s, c) accumulator merges used by every
// compensated reduction in the library.
//   TwoSum (Knuth) ........ source/common/ogita/gpu_reduction_ogita_impl_functions.cuh:115-123
//   TwoProd (FMA) ......... gpu_reduction_ogita_impl_functions.cuh:89-113
//   per-element Dot2 step . gpu_reduction_ogita_impl.cuh:498-501  (c += pi + t)
//   pair merge ............ gpu_reduction_ogita_impl.cuh:532-534  (s = TwoSum(s,s'), c = (c+t)+c')
// The tree shape is our own (wave64 butterfly + LDS across waves + fixed-order
// final stage), not the reference's 1024-thread / warp-32 tree.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpb {
namespace dev {

template <typename T> __device__ __forceinline__ T fma_rn(T a, T b, T c);
template <> __device__ __forceinline__ double fma_rn<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> __device__ __forceinline__ float fma_rn<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// s = fl(a + b), a + b = s + t exactly (no overflow).
template <typename T>
__device__ __forceinline__ T two_sum(T a, T b, T& t)
{
    T s = a + b;
    T bs = s - a;
    T as = s - bs;
    t = (b - bs) + (a - as);
    return s;
}

// p = fl(a b), a b = p + e exactly (no underflow below the FMA's reach).
template <typename T>
__device__ __forceinline__ T two_prod(T a, T b, T& e)
{
    T p = a * b;
    e = fma_rn<T>(a, b, -p);
    return p;
}

// Compensated accumulator state of one lane.
template <typename T>
struct sc_t {
    T s, c;
};

// ORO Dot2 step: (s, c) += x*y.
template <typename T>
__device__ __forceinline__ void dot2_step(T& s, T& c, T x, T y)
{
    T pi, t;
    T p = two_prod<T>(x, y, pi);
    s = two_sum<T>(s, p, t);
    c = c + (pi + t);
}

// ORO Sum2 step: (s, c) += v.
template <typename T>
__device__ __forceinline__ void sum2_step(T& s, T& c, T v)
{
    T t;
    s = two_sum<T>(s, v, t);
    c = c + t;
}

// Merge (s2, c2) into (s, c): Sum2 over the s parts, plain sum of the c parts.
template <typename T>
__device__ __forceinline__ void merge(T& s, T& c, T s2, T c2)
{
    T t;
    s = two_sum<T>(s, s2, t);
    c = (c + t) + c2;
}

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m)
{
    return __shfl_xor(v, m, 64);
}

// Butterfly over the 64 lanes of a wave; every lane ends with the wave's total.
// Level order: lane distance 32, 16, ..., 1; each lane merges its partner into
// its own value.
template <typename T, bool COMP>
__device__ __forceinline__ void wave_reduce(T& s, T& c)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        T s2 = shfl_xor<T>(s, m);
        if (COMP) {
            T c2 = shfl_xor<T>(c, m);
            merge<T>(s, c, s2, c2);
        } else {
            s = s + s2;
        }
    }
}

// Block reduction: wave butterflies, then lane 0 of wave 0 merges the wave
// totals in wave order.  Result valid in thread 0 only.  Uses 2*NW words of LDS.
template <typename T, bool COMP, int BLOCK>
__device__ __forceinline__ void block_reduce(T& s, T& c)
{
    constexpr int NW = BLOCK / 64;
    __shared__ T lds_s[NW];
    __shared__ T lds_c[NW];
    wave_reduce<T, COMP>(s, c);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (NW > 1) {
        if (lane == 0) {
            lds_s[wid] = s;
            lds_c[wid] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < NW; ++w) {
                if (COMP) merge<T>(s, c, lds_s[w], lds_c[w]);
                else s = s + lds_s[w];
            }
        }
        __syncthreads();
    }
}

// 16-byte vector type per element type.
template <typename T> struct vec16;
template <> struct vec16<double> { using type = double2; static constexpr int W = 2; };
template <> struct vec16<float> { using type = float4; static constexpr int W = 4; };

template <typename T>
__device__ __forceinline__ T vget(const typename vec16<T>::type& v, int i);
template <> __device__ __forceinline__ double vget<double>(const double2& v, int i) { return i == 0 ? v.x : v.y; }
template <> __device__ __forceinline__ float vget<float>(const float4& v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// Nontemporal (streaming) loads and stores; 16-byte vectors stay single
// dwordx4 instructions with the nt bit.
template <typename V>
__device__ __forceinline__ V load_nt(const V* p)
{
    return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ double2 load_nt<double2>(const double2* p)
{
    double2 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
template <>
__device__ __forceinline__ float4 load_nt<float4>(const float4* p)
{
    float4 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    v.z = __builtin_nontemporal_load(&p->z);
    v.w = __builtin_nontemporal_load(&p->w);
    return v;
}
template <typename V>
__device__ __forceinline__ void store_nt(V* p, const V& v)
{
    __builtin_nontemporal_store(v, p);
}
template <>
__device__ __forceinline__ void store_nt<double2>(double2* p, const double2& v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}
template <>
__device__ __forceinline__ void store_nt<float4>(float4* p, const float4& v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
    __builtin_nontemporal_store(v.z, &p->z);
    __builtin_nontemporal_store(v.w, &p->w);
}

// One lane's share of a CSR row product: entries beg + lane, beg + lane + LPR, ...
// accumulated into (s, c) in that order (Dot2 steps when COMP, fma otherwise; NEG
// negates the values, for residual rows).  The lane's first two entries issue both
// column/value loads and then both gathers of x before any arithmetic, so a row of
// up to 2 LPR entries costs three dependent memory round trips (row pointer,
// column/value, x) rather than two per entry; longer rows continue one entry at a
// time.  The order of accumulation is that of the plain loop.
template <typename T, int LPR, bool COMP, bool NEG>
__device__ __forceinline__ void csr_row_acc(int beg, int end, int lane, const int* __restrict__ col,
                                            const T* __restrict__ val, const T* __restrict__ x, T& s, T& c)
{
    const int k0 = beg + lane, k1 = k0 + LPR;
    const bool v0 = k0 < end, v1 = k1 < end;
    int c0 = 0, c1 = 0;
    T a0 = T(0), a1 = T(0);
    if (v0) {
        c0 = col[k0];
        a0 = val[k0];
    }
    if (v1) {
        c1 = col[k1];
        a1 = val[k1];
    }
    T x0 = T(0), x1 = T(0);
    if (v0) x0 = x[c0];
    if (v1) x1 = x[c1];
    if (NEG) {
        a0 = -a0;
        a1 = -a1;
    }
    if (v0) {
        if (COMP) dot2_step<T>(s, c, a0, x0);
        else s = fma_rn<T>(a0, x0, s);
    }
    if (v1) {
        if (COMP) dot2_step<T>(s, c, a1, x1);
        else s = fma_rn<T>(a1, x1, s);
    }
    for (int k = k1 + LPR; k < end; k += LPR) {
        const T a = NEG ? -val[k] : val[k];
        const T xv = x[col[k]];
        if (COMP) dot2_step<T>(s, c, a, xv);
        else s = fma_rn<T>(a, xv, s);
    }
}

// Staged CSR tiles.  A block forms the rows of a tile of RT consecutive rows in two
// steps: every thread of the block reads consecutive entries of the tile (coalesced
// column and value loads, the gather of x right behind them) into LDS; then the rows
// are formed from LDS with the lanes-per-row structure of csr_row_acc, the same
// operations in the same order, so the same bits.  A tile of more than CAP entries is
// formed directly from global memory instead.  On the 256^3 7-point matrix (4 lanes
// per row, RT 128, CAP 1024) the SpMV takes 378 us against 410 us direct: a lane's
// strided 4- and 8-byte loads of its own row entries were the cost, not the gather.
// Streaming the columns and values in and y out nontemporally, with a few tiles per
// block (2 lanes per row), takes it to 320-330 us.
template <typename T, int RT, int CAP>
struct csr_stage {
    T v[CAP];
    T x[CAP];
    int rp[RT + 1];
};

// Loads the row pointers of rows [r0, r0 + nr) into S.rp and, when the tile's
// entries fit, their values and the gathered x into S.v / S.x (indexed from the
// tile's first entry, S.rp[0]); returns whether they were staged (the same in every
// thread).  Every thread of the block calls it: it holds barriers.  NT: the columns
// and values, read once, are loaded nontemporally so that they do not displace the
// x lines the neighbouring tiles gather again.
template <typename T, int RT, int CAP, int BLOCK, bool NT = false>
__device__ __forceinline__ bool csr_stage_tile(csr_stage<T, RT, CAP>& S, long r0, int nr, const int* __restrict__ rp,
                                               const int* __restrict__ col, const T* __restrict__ val,
                                               const T* __restrict__ x)
{
    static_assert(CAP % BLOCK == 0, "csr_stage_tile: CAP must be a multiple of the block");
    constexpr int PER = CAP / BLOCK;
    for (int i = threadIdx.x; i <= nr; i += BLOCK) S.rp[i] = rp[r0 + i];
    __syncthreads();
    const int e0 = S.rp[0], ne = S.rp[nr] - e0;
    if (ne > CAP) return false;
    int c[PER];
    T v[PER], xv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + j * BLOCK;
        c[j] = 0;
        v[j] = T(0);
        if (k < ne) {
            if constexpr (NT) {
                c[j] = __builtin_nontemporal_load(col + e0 + k);
                v[j] = __builtin_nontemporal_load(val + e0 + k);
            } else {
                c[j] = col[e0 + k];
                v[j] = val[e0 + k];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + j * BLOCK;
        xv[j] = k < ne ? x[c[j]] : T(0);
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + j * BLOCK;
        if (k < ne) {
            S.v[k] = v[j];
            S.x[k] = xv[j];
        }
    }
    __syncthreads();
    return true;
}

// csr_row_acc over staged operands: tile entries [beg, end) (relative to the tile's
// first entry), lane's entries beg + lane, beg + lane + LPR, ... in that order.
template <typename T, int LPR, bool COMP, bool NEG>
__device__ __forceinline__ void csr_row_acc_staged(int beg, int end, int lane, const T* sv, const T* sx, T& s, T& c)
{
    for (int k = beg + lane; k < end; k += LPR) {
        const T a = NEG ? -sv[k] : sv[k];
        if (COMP) dot2_step<T>(s, c, a, sx[k]);
        else s = fma_rn<T>(a, sx[k], s);
    }
}

// XCD-aware row ranges for the row kernels (SpMV, its fused forms).  Consecutive
// workgroups are dispatched round-robin over the XCD_COUNT XCDs of the device, each
// with its own L2; workgroup b runs on XCD b % XCD_COUNT.  Giving each XCD one
// contiguous slice of the rows keeps the x gathers of neighbouring rows (stencil
// planes apart) in the same L2.  With a grid that is not a multiple of XCD_COUNT the
// rows are strided over all blocks in order.  Every row is computed the same way in
// either mapping.
constexpr int XCD_COUNT = 8;   // MI355X: 8 XCDs of 32 CUs
struct xcd_rows {
    long first, end, stride;   // rows first + k * stride < end of this block (per row slot 0)
    __device__ static xcd_rows of(long rows, int rows_per_block) { return of(rows, rows_per_block, gridDim.x, blockIdx.x); }
    // the same for block b of a grid of G blocks (a persistent kernel that runs a phase
    // on part of its grid)
    __device__ static xcd_rows of(long rows, int rows_per_block, long G, long b)
    {
        xcd_rows R;
        if (G % XCD_COUNT != 0 || G < XCD_COUNT) {
            R.first = b * rows_per_block;
            R.end = rows;
            R.stride = G * rows_per_block;
            return R;
        }
        const long gx = G / XCD_COUNT, x = b % XCD_COUNT, local = b / XCD_COUNT;
        // slices of whole blocks' worth of rows
        const long chunk = ((rows + XCD_COUNT - 1) / XCD_COUNT + rows_per_block - 1) / rows_per_block * rows_per_block;
        const long lo = x * chunk;
        R.end = lo + chunk < rows ? lo + chunk : rows;
        R.first = lo + local * rows_per_block;
        R.stride = gx * rows_per_block;
        return R;
    }
};

}  // namespace dev
}  // namespace hpb
