// hpb_zdot.hip -- compensated and plain complex dot products x^H y.
//
// Replaces the complex instantiations of gpu_reduction_ogita::dot
// (source/common/ogita/gpu_reduction_ogita_impl_functions.cuh:126-195): for
// a = x_i, b = y_i the product conj(a) b is split into its four real products,
//   Re = aR bR + aI bI,  Im = aR bI - aI bR,
// each formed with TwoProd, the two terms of a component added with TwoSum, and the
// component's error term = e1 + e2 + t (in that order, as in the reference); the
// components are then accumulated like the real Dot2: s = TwoSum(s, p), c += (pi + t).
// Vectors are interleaved (re, im) pairs of T, n complex elements.
//
// Launch structure as the real reductions (hpb_reduce.hip): a bounded grid strides
// over tiles of BLOCK * UNROLL complex elements (all loads of a tile issued before
// the arithmetic), each lane keeps (s, c) for both components, wave64 butterflies
// and LDS across waves; per-block partials, then a one-block launch merges them in
// block-index order (lane t takes partials t, t + BLOCK, ...; reproducible).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

namespace hpb {
namespace {

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };

template <typename T, bool COMP>
__device__ __forceinline__ void zacc(T& sr, T& cr, T& si, T& ci, T ar, T ai, T br, T bi)
{
    if constexpr (COMP) {
        T e1, e2, e3, e4, t1, t2, t;
        const T p1 = dev::two_prod<T>(ar, br, e1);
        const T p2 = dev::two_prod<T>(ai, bi, e2);
        const T p3 = dev::two_prod<T>(ar, bi, e3);
        const T p4 = dev::two_prod<T>(-ai, br, e4);
        const T pr = dev::two_sum<T>(p1, p2, t1);
        const T pir = (e1 + e2) + t1;
        const T pi = dev::two_sum<T>(p3, p4, t2);
        const T pii = (e3 + e4) + t2;
        sr = dev::two_sum<T>(sr, pr, t);
        cr = cr + (pir + t);
        si = dev::two_sum<T>(si, pi, t);
        ci = ci + (pii + t);
    } else {
        sr = dev::fma_rn<T>(ar, br, sr);
        sr = dev::fma_rn<T>(ai, bi, sr);
        si = dev::fma_rn<T>(ar, bi, si);
        si = dev::fma_rn<T>(-ai, br, si);
    }
}

template <typename T, bool NT>
__device__ __forceinline__ typename cvec<T>::type zload(const typename cvec<T>::type* p)
{
    if constexpr (NT) {
        typename cvec<T>::type v;
        v.x = __builtin_nontemporal_load(&p->x);
        v.y = __builtin_nontemporal_load(&p->y);
        return v;
    } else {
        return *p;
    }
}

// Per-block partials part[4 b .. 4 b + 3] = { s_re, c_re, s_im, c_im }.
template <typename T, bool COMP, int BLOCK, int UNROLL, bool NT>
__global__ __launch_bounds__(BLOCK) void k_zdot(size_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                T* __restrict__ part)
{
    using V = typename cvec<T>::type;
    constexpr size_t TILE = (size_t)BLOCK * UNROLL;
    const unsigned G = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
    const size_t nfull = n / TILE;
    const V* xv = reinterpret_cast<const V*>(x);
    const V* yv = reinterpret_cast<const V*>(y);
    T sr = T(0), cr = T(0), si = T(0), ci = T(0);
    for (size_t tile = b; tile < nfull; tile += G) {
        V xa[UNROLL], ya[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            xa[u] = zload<T, NT>(xv + tile * TILE + tid + (size_t)u * BLOCK);
            ya[u] = zload<T, NT>(yv + tile * TILE + tid + (size_t)u * BLOCK);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) zacc<T, COMP>(sr, cr, si, ci, xa[u].x, xa[u].y, ya[u].x, ya[u].y);
    }
    if (b == (unsigned)(nfull % G)) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const size_t e = nfull * TILE + tid + (size_t)u * BLOCK;
            if (e < n) {
                const V a = xv[e], bb = yv[e];
                zacc<T, COMP>(sr, cr, si, ci, a.x, a.y, bb.x, bb.y);
            }
        }
    }
    dev::block_reduce<T, COMP, BLOCK>(sr, cr);
    dev::block_reduce<T, COMP, BLOCK>(si, ci);
    if (tid == 0) {
        part[4 * b] = sr;
        part[4 * b + 1] = cr;
        part[4 * b + 2] = si;
        part[4 * b + 3] = ci;
    }
}

// One block: out[0..5] = { s_re, c_re, s_im, c_im, fl(s_re + c_re), fl(s_im + c_im) }
// from the G partials.
template <typename T, bool COMP, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_zdot_fin(unsigned G, const T* __restrict__ part, T* __restrict__ out)
{
    const unsigned tid = threadIdx.x;
    T fr = T(0), gr = T(0), fi = T(0), gi = T(0);
    for (unsigned i = tid; i < G; i += BLOCK) {
        if (COMP) {
            dev::merge<T>(fr, gr, part[4 * i], part[4 * i + 1]);
            dev::merge<T>(fi, gi, part[4 * i + 2], part[4 * i + 3]);
        } else {
            fr = fr + part[4 * i];
            fi = fi + part[4 * i + 2];
        }
    }
    dev::block_reduce<T, COMP, BLOCK>(fr, gr);
    dev::block_reduce<T, COMP, BLOCK>(fi, gi);
    if (tid == 0) {
        out[0] = fr;
        out[1] = COMP ? gr : T(0);
        out[2] = fi;
        out[3] = COMP ? gi : T(0);
        out[4] = COMP ? fr + gr : fr;
        out[5] = COMP ? fi + gi : fi;
    }
}

constexpr int ZBLOCK = 256;
constexpr int ZUNROLL = 4;
// Nontemporal loads: at n = 2^26 complex FP64 (2 GiB read) the call takes 290.7 us
// (7389 GB/s) against 332.1 us with ordinary loads; the same sweep over grids of
// 192-2048 blocks of 256 or 512 threads found nothing faster than 256 blocks x 256.
constexpr bool ZNT = true;

template <typename T, bool COMP>
int zdot_entry(hpb_handle_t h, size_t n, const T* x, const T* y, T* dev_out, T* host_out)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (dev_out || host_out), HPB_ERR_INVALID, "complex dot: null argument");
    HPB_REQUIRE(n == 0 || (x && y), HPB_ERR_INVALID, "complex dot: null vector");
    HPB_REQUIRE(((uintptr_t)x % (2 * sizeof(T))) == 0 && ((uintptr_t)y % (2 * sizeof(T))) == 0, HPB_ERR_INVALID,
                "complex dot: vectors must be aligned to one complex element");
    const size_t tile = (size_t)ZBLOCK * ZUNROLL;
    size_t grid = (n + tile - 1) / tile;
    const size_t cap = (size_t)(h->num_cu > 0 ? h->num_cu : 256);
    if (grid > cap) grid = cap;
    if (grid < 1) grid = 1;
    T* out = dev_out ? dev_out : reinterpret_cast<T*>(h->dres);
    T* part = reinterpret_cast<T*>(h->part);
    hipLaunchKernelGGL((k_zdot<T, COMP, ZBLOCK, ZUNROLL, ZNT>), dim3((unsigned)grid), dim3(ZBLOCK), 0, h->stream, n, x,
                       y, part);
    hipLaunchKernelGGL((k_zdot_fin<T, COMP, ZBLOCK>), dim3(1), dim3(ZBLOCK), 0, h->stream, (unsigned)grid, part, out);
    HPB_LAUNCH_CHECK();
    if (!dev_out) {
        HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 6 * sizeof(T), hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipStreamSynchronize(h->stream));
        const T* r = reinterpret_cast<const T*>(h->hres);
        for (int i = 0; i < 6; ++i) host_out[i] = r[i];
    }
    return HPB_OK;
}

}  // namespace
}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_zdot2_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out6)
{
    return zdot_entry<double, true>(h, n, x, y, nullptr, out6);
}
int hpb_zdot2_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out6)
{
    return zdot_entry<float, true>(h, n, x, y, nullptr, out6);
}
int hpb_zdot_f64(hpb_handle_t h, size_t n, const double* x, const double* y, double* out6)
{
    return zdot_entry<double, false>(h, n, x, y, nullptr, out6);
}
int hpb_zdot_f32(hpb_handle_t h, size_t n, const float* x, const float* y, float* out6)
{
    return zdot_entry<float, false>(h, n, x, y, nullptr, out6);
}
int hpb_zdot2_f64_async(hpb_handle_t h, size_t n, const double* x, const double* y, double* dev_out6)
{
    HPB_REQUIRE(dev_out6, HPB_ERR_INVALID, "complex dot: null device output");
    return zdot_entry<double, true>(h, n, x, y, dev_out6, nullptr);
}

}  // extern "C"
