// hpb_multi.hip -- fused multi-vector kernels for Gram-Schmidt.
//
// The reference orthogonalises with modified Gram-Schmidt, one dot and one axpy
// per basis vector, each a separate pass over the vectors and a host round trip
// (source/numerical_algos/lin_solvers/gmres.h:426-452).  These kernels let a
// classical Gram-Schmidt pass read w once for k basis vectors:
//   multidot:  out_j = V_j . w,  j < k <= 64   (compensated or plain)
//   multiaxpy: y := y + sum_j a_j V_j          (compensated or plain), one pass
// Algorithmic bytes: multidot (k + 1) n sizeof(T); multiaxpy (k + 2) n sizeof(T).
//
// multidot grid: blockIdx.y picks a chunk of KC columns, blockIdx.x strides over
// tiles of w; every thread keeps KC (s, c) accumulators.  Per-block partials go
// to the handle workspace and a one-block kernel merges each column's G partials
// in a fixed order.  (No last-block ticket: its agent-scope release in every
// block costs an L2 write-back each on a multi-XCD part, more than the second
// launch for grids of more than a few hundred blocks.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include <atomic>
#include <mutex>
#include <type_traits>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"

namespace hpb {

template <typename T>
struct ptr_pack {
    const T* p[HPB_MAX_MULTIDOT];
};
template <typename T>
struct coef_pack {
    T a[HPB_MAX_MULTIDOT];
};

// NT: nontemporal loads of the basis columns, each read once.  w is always loaded
// through the caches: the blocks of the other column chunks read the same w tiles
// and find them in L2 (at n = 2^24, k = 64, streaming w as well cost 10 % of the
// rate).  HOIST (one column chunk): the kc column loads of a tile are all issued
// before the first Dot2 step; UNR2 (with HOIST, k <= 4): two tiles' loads are issued
// together.  The order of the arithmetic is unchanged by either.
template <typename T, int KC, bool COMP, int BLOCK, bool VEC, bool NT = false, bool HOIST = false, bool UNR2 = false>
__global__ __launch_bounds__(BLOCK) void k_multidot(size_t n, int k, ptr_pack<T> V, const T* __restrict__ w,
                                                    T* __restrict__ part, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip) return;   // solver graphs: flag set, launch does nothing
    using Vt = std::conditional_t<VEC, typename dev::vec16<T>::type, T>;
    constexpr int W = VEC ? dev::vec16<T>::W : 1;
    const unsigned G = gridDim.x;
    const int j0 = blockIdx.y * KC;
    const int kc = (k - j0) < KC ? (k - j0) : KC;
    const size_t nv = n / W;
    T s[KC], c[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        s[j] = T(0);
        c[j] = T(0);
    }
    const Vt* wv = reinterpret_cast<const Vt*>(w);
    auto load_col = [&](int j, size_t vi) -> Vt {
        if constexpr (NT && VEC) return dev::load_nt<Vt>(reinterpret_cast<const Vt*>(V.p[j0 + j]) + vi);
        else return reinterpret_cast<const Vt*>(V.p[j0 + j])[vi];
    };
    auto col_steps = [&](int j, const Vt& a, const Vt& b) {   // column j's Dot2 steps of one tile
#pragma unroll
        for (int q = 0; q < W; ++q) {
            T av, bv;
            if constexpr (VEC) {
                av = dev::vget<T>(a, q);
                bv = dev::vget<T>(b, q);
            } else {
                av = a;
                bv = b;
            }
            if (COMP) dev::dot2_step<T>(s[j], c[j], bv, av);
            else s[j] = dev::fma_rn<T>(bv, av, s[j]);
        }
    };
    size_t vi = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t vstride = (size_t)G * BLOCK;
    if constexpr (HOIST && UNR2) {
        for (; vi + vstride < nv; vi += 2 * vstride) {
            const Vt a0 = wv[vi], a1 = wv[vi + vstride];
            Vt b0[KC], b1[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) {
                    b0[j] = load_col(j, vi);
                    b1[j] = load_col(j, vi + vstride);
                }
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) col_steps(j, a0, b0[j]);
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) col_steps(j, a1, b1[j]);
        }
    }
    for (; vi < nv; vi += vstride) {
        const Vt a = wv[vi];
        if constexpr (HOIST) {
            Vt bh[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) bh[j] = load_col(j, vi);
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) col_steps(j, a, bh[j]);
        } else {
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) col_steps(j, a, load_col(j, vi));
        }
    }
    if constexpr (VEC) {
        const size_t r0 = nv * W;
        if (blockIdx.x == 0 && threadIdx.x < n - r0) {
            const size_t i = r0 + threadIdx.x;
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (j < kc) {
                    if (COMP) dev::dot2_step<T>(s[j], c[j], V.p[j0 + j][i], w[i]);
                    else s[j] = dev::fma_rn<T>(V.p[j0 + j][i], w[i], s[j]);
                }
        }
    }
    // per-block totals, column by column
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        if (j < kc) {
            T ss = s[j], cc = c[j];
            dev::block_reduce<T, COMP, BLOCK>(ss, cc);
            if (threadIdx.x == 0) {
                part[2 * ((size_t)(j0 + j) * G + blockIdx.x)] = ss;
                part[2 * ((size_t)(j0 + j) * G + blockIdx.x) + 1] = cc;
            }
        }
    }
}

// Merge of the G partials of each column, one wave per column (lane l takes
// partials l, l+64, ...; then the wave butterfly): a fixed order, so results
// are reproducible.  out[4j..4j+3] = { s, c, s + c, s + c } (plain: c = 0).
template <typename T, bool COMP>
__global__ __launch_bounds__(1024) void k_multidot_fin(const T* __restrict__ part, unsigned G, int k,
                                                       T* __restrict__ out, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = wid; j < k; j += 1024 / 64) {
        T fs = T(0), fc = T(0);
        for (unsigned b = lane; b < G; b += 64) {
            const T ps = part[2 * ((size_t)j * G + b)];
            const T pc = part[2 * ((size_t)j * G + b) + 1];
            if (COMP) dev::merge<T>(fs, fc, ps, pc);
            else fs = fs + ps;
        }
        dev::wave_reduce<T, COMP>(fs, fc);
        if (lane == 0) {
            const T r = fs + fc;
            out[4 * j] = fs;
            out[4 * j + 1] = COMP ? fc : T(0);
            out[4 * j + 2] = COMP ? r : fs;
            out[4 * j + 3] = COMP ? r : fs;
        }
    }
}

// y := y + sum_j a_j V_j (compensated: Dot2 over the k+1 terms of each element)
template <typename T, bool COMP, bool VEC, bool DEVCOEF>
__global__ __launch_bounds__(256) void k_multiaxpy(size_t n, int k, coef_pack<T> a, const T* __restrict__ da,
                                                   ptr_pack<T> V, T* __restrict__ y, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip) return;
    using Vt = std::conditional_t<VEC, typename dev::vec16<T>::type, T>;
    constexpr int W = VEC ? dev::vec16<T>::W : 1;
    const size_t nv = n / W;
    __shared__ T sa[HPB_MAX_MULTIDOT];
    if (threadIdx.x < (unsigned)k) sa[threadIdx.x] = DEVCOEF ? da[threadIdx.x] : a.a[threadIdx.x];
    __syncthreads();
    Vt* yv = reinterpret_cast<Vt*>(y);
    for (size_t vi = (size_t)blockIdx.x * 256 + threadIdx.x; vi < nv; vi += (size_t)gridDim.x * 256) {
        Vt yy = yv[vi];
        T s[W], c[W];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            if constexpr (VEC) s[q] = dev::vget<T>(yy, q);
            else s[q] = yy;
            c[q] = T(0);
        }
        for (int j = 0; j < k; ++j) {
            const Vt b = reinterpret_cast<const Vt*>(V.p[j])[vi];
            const T aj = sa[j];
#pragma unroll
            for (int q = 0; q < W; ++q) {
                T bv;
                if constexpr (VEC) bv = dev::vget<T>(b, q);
                else bv = b;
                if (COMP) dev::dot2_step<T>(s[q], c[q], aj, bv);
                else s[q] = dev::fma_rn<T>(aj, bv, s[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const T r = COMP ? s[q] + c[q] : s[q];
            if constexpr (VEC) reinterpret_cast<T*>(&yy)[q] = r;
            else yy = r;
        }
        yv[vi] = yy;
    }
    if constexpr (VEC) {
        const size_t r0 = nv * W;
        if (blockIdx.x == 0 && threadIdx.x < n - r0) {
            const size_t i = r0 + threadIdx.x;
            T s = y[i], c = T(0);
            for (int j = 0; j < k; ++j) {
                if (COMP) dev::dot2_step<T>(s, c, sa[j], V.p[j][i]);
                else s = dev::fma_rn<T>(sa[j], V.p[j][i], s);
            }
            y[i] = COMP ? s + c : s;
        }
    }
}

static inline bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Checks the arguments and launches the multi-dot and its merge on stream st with
// launch shape tu on a device of ncu CUs; part holds the per-block partials, at
// most max_pairs (s, c) pairs (the grid is capped so that g k <= max_pairs).
template <typename T>
static int multidot_launch(hipStream_t st, int ncu, const hpb_tuning_t& tu, int k, const T* const* V, size_t n,
                           const T* w, int comp, T* part, size_t max_pairs, T* out, const int* skip)
{
    HPB_REQUIRE(V && (n == 0 || w), HPB_ERR_INVALID, "hpb_multidot: null argument");
    HPB_REQUIRE(k >= 1 && k <= HPB_MAX_MULTIDOT, HPB_ERR_INVALID, "hpb_multidot: need 1 <= k <= 64");
    HPB_REQUIRE(out, HPB_ERR_INVALID, "hpb_multidot: null output");
    ptr_pack<T> pk;
    bool vec = al16(w);
    for (int j = 0; j < k; ++j) {
        HPB_REQUIRE(n == 0 || V[j], HPB_ERR_INVALID, "hpb_multidot: null basis vector");
        pk.p[j] = V[j];
        vec = vec && al16(V[j]);
    }
    for (int j = k; j < HPB_MAX_MULTIDOT; ++j) pk.p[j] = nullptr;
    // launch shape (tuning kind 3): columns per chunk (4 or 8), blocks over all
    // chunks, nontemporal loads
    const int KCr = tu.unroll == 4 ? 4 : 8;
    const bool nt = (tu.flags & HPB_TUNE_NONTEMPORAL) != 0;
    const int nch = (k + KCr - 1) / KCr;
    const size_t W = vec ? dev::vec16<T>::W : 1;
    const size_t nv = n / W;
    size_t g = (nv + 255) / 256;
    const size_t cus = (size_t)(ncu > 0 ? ncu : 256);
    size_t cap = (size_t)tu.grid_max / (size_t)nch;
    if (cap < cus) cap = cus;   // at least one block per CU per chunk
    // one chunk: few long-running blocks, as for a single Dot2 (n = 2^24, GB/s at 2 / 1
    // blocks per CU against the 8 per CU of the table: k = 1: 6250 / 5494 / 5132; k = 2:
    // 6263 / 5817 / 5553; k = 4: 6510 / 6864 / 6070; k = 8: 6636 / 6146 / 6506)
    if (nch == 1) cap = std::min(cap, (k == 3 || k == 4 ? 1 : 2) * cus);
    if (cap * (size_t)k > max_pairs) cap = max_pairs / (size_t)k;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    dim3 grid((unsigned)g, (unsigned)nch);
#define HPB_MDOT(KC_, C_, V_, NT_)                                                                               \
    do {                                                                                                        \
        if constexpr (V_) {                                                                                     \
            if (k <= 4)                                                                                         \
                hipLaunchKernelGGL((k_multidot<T, 4, C_, 256, V_, NT_, true, true>), grid, dim3(256), 0, st, n, k, \
                                   pk, w, part, skip);                                                          \
            else if (nch == 1)                                                                                  \
                hipLaunchKernelGGL((k_multidot<T, KC_, C_, 256, V_, NT_, true>), grid, dim3(256), 0, st, n, k, pk, \
                                   w, part, skip);                                                              \
            else                                                                                                \
                hipLaunchKernelGGL((k_multidot<T, KC_, C_, 256, V_, NT_>), grid, dim3(256), 0, st, n, k, pk, w,    \
                                   part, skip);                                                                 \
        } else {                                                                                                \
            hipLaunchKernelGGL((k_multidot<T, KC_, C_, 256, V_, NT_>), grid, dim3(256), 0, st, n, k, pk, w, part,  \
                               skip);                                                                           \
        }                                                                                                       \
    } while (0)
#define HPB_MDOT_KC(C_, V_, NT_)                  \
    do {                                          \
        if (KCr == 4) HPB_MDOT(4, C_, V_, NT_);   \
        else HPB_MDOT(8, C_, V_, NT_);            \
    } while (0)
    if (comp) {
        if (vec) { if (nt) HPB_MDOT_KC(true, true, true); else HPB_MDOT_KC(true, true, false); }
        else HPB_MDOT_KC(true, false, false);
        hipLaunchKernelGGL((k_multidot_fin<T, true>), dim3(1), dim3(1024), 0, st, part, (unsigned)g, k, out, skip);
    } else {
        if (vec) { if (nt) HPB_MDOT_KC(false, true, true); else HPB_MDOT_KC(false, true, false); }
        else HPB_MDOT_KC(false, false, false);
        hipLaunchKernelGGL((k_multidot_fin<T, false>), dim3(1), dim3(1024), 0, st, part, (unsigned)g, k, out, skip);
    }
#undef HPB_MDOT_KC
#undef HPB_MDOT
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename T>
static int multidot_entry(hpb_handle_t h, int k, const T* const* V, size_t n, const T* w, int comp, T* dev_out,
                          T* host_out, const int* skip = nullptr)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h, HPB_ERR_INVALID, "hpb_multidot: null handle");
    HPB_REQUIRE(dev_out || host_out, HPB_ERR_INVALID, "hpb_multidot: null output");
    T* out = dev_out ? dev_out : reinterpret_cast<T*>(h->dres);
    const int rc = multidot_launch<T>(h->stream, h->num_cu, h->tune_multi[sizeof(T) == 8 ? 0 : 1], k, V, n, w, comp,
                                      reinterpret_cast<T*>(h->part), (size_t)HPB_MAX_GRID * HPB_MAX_MULTIDOT, out,
                                      skip);
    if (rc != HPB_OK) return rc;
    if (!dev_out) {
        HPB_HIP(hipMemcpyAsync(h->hres, h->dres, 4 * sizeof(T) * k, hipMemcpyDeviceToHost, h->stream));
        HPB_HIP(hipStreamSynchronize(h->stream));
        const T* r = reinterpret_cast<const T*>(h->hres);
        for (int i = 0; i < 4 * k; ++i) host_out[i] = r[i];
    }
    return HPB_OK;
}

template <typename T>
static int multiaxpy_entry(hpb_handle_t h, int k, const T* a, const T* dev_a, const T* const* V, size_t n, T* y,
                           int comp, const int* skip = nullptr)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && V && (a || dev_a) && (n == 0 || y), HPB_ERR_INVALID, "hpb_multiaxpy: null argument");
    HPB_REQUIRE(k >= 1 && k <= HPB_MAX_MULTIDOT, HPB_ERR_INVALID, "hpb_multiaxpy: need 1 <= k <= 64");
    if (n == 0) return HPB_OK;
    ptr_pack<T> pk;
    coef_pack<T> ca;
    bool vec = al16(y);
    for (int j = 0; j < HPB_MAX_MULTIDOT; ++j) {
        pk.p[j] = j < k ? V[j] : nullptr;
        ca.a[j] = (j < k && a) ? a[j] : T(0);
        if (j < k) {
            HPB_REQUIRE(V[j], HPB_ERR_INVALID, "hpb_multiaxpy: null basis vector");
            vec = vec && al16(V[j]);
        }
    }
    const size_t W = vec ? dev::vec16<T>::W : 1;
    size_t g = (n / W + 255) / 256;
    const size_t cap = (size_t)(h->num_cu > 0 ? h->num_cu : 256) * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    const bool dc = dev_a != nullptr;
#define HPB_MAXPY(C, VV, D) \
    hipLaunchKernelGGL((k_multiaxpy<T, C, VV, D>), dim3((unsigned)g), dim3(256), 0, h->stream, n, k, ca, dev_a, pk, y, skip)
    if (comp) {
        if (vec) { if (dc) HPB_MAXPY(true, true, true); else HPB_MAXPY(true, true, false); }
        else { if (dc) HPB_MAXPY(true, false, true); else HPB_MAXPY(true, false, false); }
    } else {
        if (vec) { if (dc) HPB_MAXPY(false, true, true); else HPB_MAXPY(false, true, false); }
        else { if (dc) HPB_MAXPY(false, false, true); else HPB_MAXPY(false, false, false); }
    }
#undef HPB_MAXPY
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename T>
int multidot_async(hpb_handle_t h, int k, const T* const* V, size_t n, const T* w, bool comp, T* dev_out4k,
                   const int* skip)
{
    return multidot_entry<T>(h, k, V, n, w, comp ? 1 : 0, dev_out4k, nullptr, skip);
}
template <typename T>
int multiaxpy_devcoef(hpb_handle_t h, int k, const T* dev_a, const T* const* V, size_t n, T* y, bool comp,
                      const int* skip)
{
    return multiaxpy_entry<T>(h, k, nullptr, dev_a, V, n, y, comp ? 1 : 0, skip);
}
template int multidot_async<double>(hpb_handle_t, int, const double* const*, size_t, const double*, bool, double*,
                                    const int*);
template int multidot_async<float>(hpb_handle_t, int, const float* const*, size_t, const float*, bool, float*,
                                   const int*);
template int multiaxpy_devcoef<double>(hpb_handle_t, int, const double*, const double* const*, size_t, double*, bool,
                                       const int*);
template int multiaxpy_devcoef<float>(hpb_handle_t, int, const float*, const float* const*, size_t, float*, bool,
                                      const int*);

// ---------------------------------------------------------------------------
// No-allocation form (hpb_multidot_f64_noalloc): the per-block partials go to the
// file-scope device array `scratch` instead of a handle's workspace, for callers
// that cannot allocate (graph capture, serving loops).  Each device holds its own
// copy of the array, so its address is looked up once per device with that device
// current (hipGetSymbolAddress) and kept; host code uses only that address.  All
// calls on one device share the array: calls that overlap on one device race.
// Every launch overwrites the partials it reads, so nothing needs clearing.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_PAIRS = 16384;   // (s, c) pairs: 256 KiB per device
constexpr int NOALLOC_DEVICES = 64;
static __device__ __align__(256) double scratch[2 * NOALLOC_PAIRS];

struct noalloc_dev {
    std::atomic<int> ready{0};
    double* ptr = nullptr;
    int ncu = 0;
};
static noalloc_dev g_noalloc[NOALLOC_DEVICES];
static std::mutex g_noalloc_mu;

// the current device's scratch address and CU count (looked up on its first use)
static int noalloc_scratch(double** ptr, int* ncu)
{
    int dev = 0;
    HPB_HIP(hipGetDevice(&dev));
    HPB_REQUIRE(dev >= 0 && dev < NOALLOC_DEVICES, HPB_ERR_UNSUPPORTED, "hpb_multidot_noalloc: device index >= 64");
    noalloc_dev& d = g_noalloc[dev];
    if (!d.ready.load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (!d.ready.load(std::memory_order_relaxed)) {
            void* p = nullptr;
            HPB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            int cu = 0;
            HPB_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
            d.ptr = static_cast<double*>(p);
            d.ncu = cu;
            d.ready.store(1, std::memory_order_release);
        }
    }
    *ptr = d.ptr;
    *ncu = d.ncu;
    return HPB_OK;
}

int noalloc_init()
{
    double* p = nullptr;
    int cu = 0;
    return noalloc_scratch(&p, &cu);
}

}  // namespace hpb

using namespace hpb;

extern "C" {

int hpb_multidot_f64_noalloc(void* stream, int k, const double* const* V, size_t n, const double* w,
                             int compensated, double* dev_out4k)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(k >= 1 && k <= HPB_MAX_MULTIDOT, HPB_ERR_INVALID,
                "hpb_multidot_f64_noalloc: need 1 <= k <= 64 (the scratch array holds the partials of 64 columns)");
    HPB_REQUIRE(dev_out4k, HPB_ERR_INVALID, "hpb_multidot_f64_noalloc: null device output");
    double* part = nullptr;
    int ncu = 0;
    const int rc = noalloc_scratch(&part, &ncu);
    if (rc != HPB_OK) return rc;
    // the launch shape of a new handle's default tuning (hpb_create)
    const hpb_tuning_t tu = {256, 8, 8 * (ncu > 0 ? ncu : 256), HPB_TUNE_NONTEMPORAL};
    return multidot_launch<double>((hipStream_t)stream, ncu, tu, k, V, n, w, compensated, part, NOALLOC_PAIRS,
                                   dev_out4k, nullptr);
}

int hpb_multidot_f64(hpb_handle_t h, int k, const double* const* V, size_t n, const double* w, int compensated,
                     double* out4k)
{ return multidot_entry<double>(h, k, V, n, w, compensated, nullptr, out4k); }
int hpb_multidot_f32(hpb_handle_t h, int k, const float* const* V, size_t n, const float* w, int compensated,
                     float* out4k)
{ return multidot_entry<float>(h, k, V, n, w, compensated, nullptr, out4k); }
int hpb_multidot_f64_async(hpb_handle_t h, int k, const double* const* V, size_t n, const double* w, int compensated,
                           double* dev_out4k)
{
    HPB_REQUIRE(dev_out4k, HPB_ERR_INVALID, "hpb_multidot_async: null device output");
    return multidot_entry<double>(h, k, V, n, w, compensated, dev_out4k, nullptr);
}
int hpb_multidot_f32_async(hpb_handle_t h, int k, const float* const* V, size_t n, const float* w, int compensated,
                           float* dev_out4k)
{
    HPB_REQUIRE(dev_out4k, HPB_ERR_INVALID, "hpb_multidot_async: null device output");
    return multidot_entry<float>(h, k, V, n, w, compensated, dev_out4k, nullptr);
}
int hpb_multiaxpy_f64(hpb_handle_t h, int k, const double* a, const double* const* V, size_t n, double* y,
                      int compensated)
{ return multiaxpy_entry<double>(h, k, a, nullptr, V, n, y, compensated); }
int hpb_multiaxpy_f32(hpb_handle_t h, int k, const float* a, const float* const* V, size_t n, float* y,
                      int compensated)
{ return multiaxpy_entry<float>(h, k, a, nullptr, V, n, y, compensated); }
int hpb_multiaxpy_dev_f64(hpb_handle_t h, int k, const double* dev_a, const double* const* V, size_t n, double* y,
                          int compensated)
{
    HPB_REQUIRE(dev_a, HPB_ERR_INVALID, "hpb_multiaxpy_dev: null device coefficients");
    return multiaxpy_entry<double>(h, k, nullptr, dev_a, V, n, y, compensated);
}
int hpb_multiaxpy_dev_f32(hpb_handle_t h, int k, const float* dev_a, const float* const* V, size_t n, float* y,
                          int compensated)
{
    HPB_REQUIRE(dev_a, HPB_ERR_INVALID, "hpb_multiaxpy_dev: null device coefficients");
    return multiaxpy_entry<float>(h, k, nullptr, dev_a, V, n, y, compensated);
}

}  // extern "C"
