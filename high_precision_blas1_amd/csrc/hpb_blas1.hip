// hpb_blas1.hip -- elementwise BLAS1 kernels (gfx950).
//
// Replaces the one-element-per-thread kernels of
// source/common/gpu_vector_operations_impl.cuh:226-528 and the cuBLAS copy/swap/
// scal/axpy calls of gpu_vector_operations.h:433-462.  One generic streaming
// kernel: 16-byte loads of every input, UNROLL vectors per thread, grid-stride
// over a bounded grid; every input vector of a tile is loaded before anything is
// stored, so in-place forms (z on both sides) are safe.
//
// Rounding is fixed explicitly with fma, not left to the compiler's contraction:
//   a*x + b*y               -> fma(a, x, b*y)
//   a*x + b*y + c*z         -> fma(c, z, fma(a, x, b*y))
//   a*x + b*y + c*w + d*z   -> fma(d, z, fma(c, w, fma(a, x, b*y)))
//   y + a*x                 -> fma(a, x, y)      (correctly rounded axpy)
//   mul_x*x + scalar        -> fma(mul_x, x, scalar)
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hpb_device.hpp"
#include "hpb_internal.hpp"
#include "hpb_philox.hpp"

namespace hpb {

using dev::fma_rn;

template <typename T> struct f_fill { T a; __device__ T operator()(T, T, T, T) const { return a; } };
template <typename T> struct f_add_mul_scalar {
    T scalar, mul;
    __device__ T operator()(T x, T, T, T) const { return fma_rn<T>(mul, x, scalar); }
};
template <typename T> struct f_scal { T a; __device__ T operator()(T x, T, T, T) const { return a * x; } };
template <typename T> struct f_copy { __device__ T operator()(T x, T, T, T) const { return x; } };
template <typename T> struct f_abs {
    __device__ T operator()(T x, T, T, T) const { return x < T(0) ? -x : (x == T(0) ? T(0) : x); }
};
template <typename T> struct f_axpy { T a; __device__ T operator()(T x, T y, T, T) const { return fma_rn<T>(a, x, y); } };
template <typename T> struct f_axpby {
    T a, b;
    __device__ T operator()(T x, T y, T, T) const { return fma_rn<T>(a, x, b * y); }
};
template <typename T> struct f_axpbypcz {
    T a, b, c;
    __device__ T operator()(T x, T y, T z, T) const { return fma_rn<T>(c, z, fma_rn<T>(a, x, b * y)); }
};
template <typename T> struct f_axpbypcwpdz {
    T a, b, c, d;
    __device__ T operator()(T x, T y, T w, T z) const
    {
        return fma_rn<T>(d, z, fma_rn<T>(c, w, fma_rn<T>(a, x, b * y)));
    }
};
template <typename T> struct f_mulpw { T a, b; __device__ T operator()(T x, T y, T, T) const { return (a * x) * (b * y); } };
template <typename T> struct f_mulpw_ip { T b; __device__ T operator()(T x, T y, T, T) const { return x * (b * y); } };
template <typename T> struct f_divpw { T a, b; __device__ T operator()(T x, T y, T, T) const { return (a * x) / (b * y); } };
template <typename T> struct f_divpw_ip { T b; __device__ T operator()(T x, T y, T, T) const { return x / (b * y); } };
// z = a*x + b*y: a x + b y = s + t + e1 + e2 exactly (TwoProd, TwoProd, TwoSum);
// the correction (e1 + e2) + t is rounded twice before the final s + correction,
// so z is faithful (one of the two doubles around a x + b y), not always the
// correctly rounded value.
template <typename T> struct f_axpbyc {
    T a, b;
    __device__ T operator()(T x, T y, T, T) const
    {
        T e1, e2, t;
        T p1 = dev::two_prod<T>(a, x, e1);
        T p2 = dev::two_prod<T>(b, y, e2);
        T s = dev::two_sum<T>(p1, p2, t);
        return s + ((e1 + e2) + t);
    }
};

template <typename T, class F, int NIN, int BLOCK, int UNROLL, bool VEC, bool NT = false>
__global__ __launch_bounds__(BLOCK) void k_map(size_t n, F f, const T* i0, const T* i1, const T* i2, const T* i3,
                                               T* out)
{
    using V = std::conditional_t<VEC, typename dev::vec16<T>::type, T>;
    constexpr int W = VEC ? dev::vec16<T>::W : 1;
    constexpr size_t TILE = (size_t)BLOCK * UNROLL;
    const size_t nv = n / W;
    const size_t ntiles = (nv + TILE - 1) / TILE;
    const V* p0 = reinterpret_cast<const V*>(i0);
    const V* p1 = reinterpret_cast<const V*>(i1);
    const V* p2 = reinterpret_cast<const V*>(i2);
    const V* p3 = reinterpret_cast<const V*>(i3);
    V* po = reinterpret_cast<V*>(out);
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const size_t base = tile * TILE + threadIdx.x;
        V a0[UNROLL], a1[UNROLL], a2[UNROLL], a3[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const size_t vi = base + (size_t)u * BLOCK;
            if (vi < nv) {
                if constexpr (NT) {
                    if constexpr (NIN > 0) a0[u] = dev::load_nt<V>(p0 + vi);
                    if constexpr (NIN > 1) a1[u] = dev::load_nt<V>(p1 + vi);
                    if constexpr (NIN > 2) a2[u] = dev::load_nt<V>(p2 + vi);
                    if constexpr (NIN > 3) a3[u] = dev::load_nt<V>(p3 + vi);
                } else {
                    if constexpr (NIN > 0) a0[u] = p0[vi];
                    if constexpr (NIN > 1) a1[u] = p1[vi];
                    if constexpr (NIN > 2) a2[u] = p2[vi];
                    if constexpr (NIN > 3) a3[u] = p3[vi];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const size_t vi = base + (size_t)u * BLOCK;
            if (vi < nv) {
                V r;
                if constexpr (VEC) {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        T v0 = NIN > 0 ? dev::vget<T>(a0[u], w) : T(0);
                        T v1 = NIN > 1 ? dev::vget<T>(a1[u], w) : T(0);
                        T v2 = NIN > 2 ? dev::vget<T>(a2[u], w) : T(0);
                        T v3 = NIN > 3 ? dev::vget<T>(a3[u], w) : T(0);
                        reinterpret_cast<T*>(&r)[w] = f(v0, v1, v2, v3);
                    }
                } else {
                    r = f(NIN > 0 ? a0[u] : T(0), NIN > 1 ? a1[u] : T(0), NIN > 2 ? a2[u] : T(0),
                          NIN > 3 ? a3[u] : T(0));
                }
                if constexpr (NT) dev::store_nt<V>(po + vi, r);
                else po[vi] = r;
            }
        }
    }
    if constexpr (VEC) {
        // trailing < W scalars
        const size_t r0 = nv * W;
        if (blockIdx.x == 0 && threadIdx.x < n - r0) {
            const size_t i = r0 + threadIdx.x;
            out[i] = f(NIN > 0 ? i0[i] : T(0), NIN > 1 ? i1[i] : T(0), NIN > 2 ? i2[i] : T(0), NIN > 3 ? i3[i] : T(0));
        }
    }
}

template <typename T> struct dtid;
template <> struct dtid<double> { static constexpr int v = 0; };
template <> struct dtid<float> { static constexpr int v = 1; };

static inline bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <typename T, class F, int NIN, int BLOCK, int UNROLL>
static void map_launch_bu(hipStream_t st, int grid, bool vec, bool nt, size_t n, const F& f, const T* i0,
                          const T* i1, const T* i2, const T* i3, T* out)
{
    if (vec && nt)
        hipLaunchKernelGGL((k_map<T, F, NIN, BLOCK, UNROLL, true, true>), dim3(grid), dim3(BLOCK), 0, st, n, f, i0, i1,
                           i2, i3, out);
    else if (vec)
        hipLaunchKernelGGL((k_map<T, F, NIN, BLOCK, UNROLL, true>), dim3(grid), dim3(BLOCK), 0, st, n, f, i0, i1, i2,
                           i3, out);
    else
        hipLaunchKernelGGL((k_map<T, F, NIN, BLOCK, UNROLL, false>), dim3(grid), dim3(BLOCK), 0, st, n, f, i0, i1,
                           i2, i3, out);
}

template <typename T, class F, int NIN>
static int map_entry(hpb_handle_t h, size_t n, const F& f, const T* i0, const T* i1, const T* i2, const T* i3,
                     T* out)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h, HPB_ERR_INVALID, "elementwise: null handle");
    if (n == 0) return HPB_OK;
    HPB_REQUIRE(out, HPB_ERR_INVALID, "elementwise: null output");
    HPB_REQUIRE((NIN < 1 || i0) && (NIN < 2 || i1) && (NIN < 3 || i2) && (NIN < 4 || i3), HPB_ERR_INVALID,
                "elementwise: null input");
    bool vec = al16(out) && (NIN < 1 || al16(i0)) && (NIN < 2 || al16(i1)) && (NIN < 3 || al16(i2)) &&
               (NIN < 4 || al16(i3));
    const hpb_tuning_t& t = h->tune_map[dtid<T>::v];
    const size_t W = vec ? dev::vec16<T>::W : 1;
    const size_t nv = n / W;
    const size_t tile = (size_t)t.block * (size_t)t.unroll;
    size_t ntiles = (nv + tile - 1) / tile;
    if (ntiles < 1) ntiles = 1;
    const int grid = (int)(ntiles < (size_t)t.grid_max ? ntiles : (size_t)t.grid_max);
    const bool nt = (t.flags & HPB_TUNE_NONTEMPORAL) != 0;
    if (t.block == 512) {
        if (t.unroll >= 4) map_launch_bu<T, F, NIN, 512, 4>(h->stream, grid, vec, nt, n, f, i0, i1, i2, i3, out);
        else map_launch_bu<T, F, NIN, 512, 2>(h->stream, grid, vec, nt, n, f, i0, i1, i2, i3, out);
    } else {
        if (t.unroll >= 4) map_launch_bu<T, F, NIN, 256, 4>(h->stream, grid, vec, nt, n, f, i0, i1, i2, i3, out);
        else if (t.unroll == 1) map_launch_bu<T, F, NIN, 256, 1>(h->stream, grid, vec, nt, n, f, i0, i1, i2, i3, out);
        else map_launch_bu<T, F, NIN, 256, 2>(h->stream, grid, vec, nt, n, f, i0, i1, i2, i3, out);
    }
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

// swap: two outputs
template <typename T>
__global__ __launch_bounds__(256) void k_swap(size_t n, T* x, T* y)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        T a = x[i];
        x[i] = y[i];
        y[i] = a;
    }
}

// any non-finite -> *flag = 1 (flag zeroed by k_zero_flag before)
template <typename T>
__global__ __launch_bounds__(256) void k_valid(size_t n, const T* x, int* flag)
{
    int bad = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        T v = x[i];
        bad |= !(v - v == T(0));   // false for inf and nan
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_zero_flag(int* flag) { *flag = 0; }

template <typename T>
__global__ void k_set_value(T* x, size_t at, T val) { x[at] = val; }

template <typename T> __device__ __forceinline__ T rnd_value(uint64_t seed, uint32_t sid, size_t i, T a, T b);
template <>
__device__ __forceinline__ double rnd_value<double>(uint64_t seed, uint32_t sid, size_t i, double a, double b)
{
    philox_out o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), sid, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u = unif53_open0(k53(o.v[0], o.v[1]));
    return fma_rn<double>(b - a, u, a);
}
template <>
__device__ __forceinline__ float rnd_value<float>(uint64_t seed, uint32_t sid, size_t i, float a, float b)
{
    philox_out o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), sid, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    float u = unif24_open0(o.v[0]);
    return fma_rn<float>(b - a, u, a);
}

template <typename T>
__global__ __launch_bounds__(256) void k_random(size_t n, uint64_t seed, uint32_t sid, T a, T b, T* x)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        x[i] = rnd_value<T>(seed, sid, i, a, b);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert(size_t n, const S* x, D* y)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) y[i] = (D)x[i];
}

// y32 := (float)(s * x64)   (scaled FP64 -> FP32 residual of iterative refinement)
__global__ __launch_bounds__(256) void k_scale_to_f32(size_t n, double s, const double* __restrict__ x,
                                                      float* __restrict__ y)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) y[i] = (float)(s * x[i]);
}

// y64 := fma(a, (double)x32, y64)   (FP64 refinement update from an FP32 correction)
__global__ __launch_bounds__(256) void k_axpy_f32_f64(size_t n, double a, const float* __restrict__ x,
                                                      double* __restrict__ y)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        y[i] = __builtin_fma(a, (double)x[i], y[i]);
}

static int simple_grid(hpb_handle_t h, size_t n)
{
    size_t g = (n + 255) / 256;
    size_t cap = (size_t)(h->num_cu > 0 ? h->num_cu : 256) * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
static int valid_entry(hpb_handle_t h, size_t n, const T* x, int* out)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && out && (n == 0 || x), HPB_ERR_INVALID, "hpb_valid: null argument");
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, h->stream, h->dflag);
    if (n) hipLaunchKernelGGL((k_valid<T>), dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, x, h->dflag);
    HPB_LAUNCH_CHECK();
    int flag = 0;
    HPB_HIP(hipMemcpyAsync(h->hres, h->dflag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    flag = *reinterpret_cast<int*>(h->hres);
    *out = flag ? 0 : 1;
    return HPB_OK;
}

template <typename T>
static int set_value_entry(hpb_handle_t h, size_t n, size_t at, T val, T* x)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && x, HPB_ERR_INVALID, "hpb_set_value: null argument");
    HPB_REQUIRE(at < n, HPB_ERR_INVALID, "hpb_set_value: index out of range");
    hipLaunchKernelGGL((k_set_value<T>), dim3(1), dim3(1), 0, h->stream, x, at, val);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename T>
static int get_value_entry(hpb_handle_t h, size_t n, size_t at, const T* x, T* out)
{
    HPB_REQUIRE(h && x && out, HPB_ERR_INVALID, "hpb_get_value: null argument");
    HPB_REQUIRE(at < n, HPB_ERR_INVALID, "hpb_get_value: index out of range");
    HPB_HIP(hipMemcpyAsync(h->hres, x + at, sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HPB_HIP(hipStreamSynchronize(h->stream));
    *out = *reinterpret_cast<const T*>(h->hres);
    return HPB_OK;
}

template <typename T>
static int random_entry(hpb_handle_t h, size_t n, uint64_t seed, uint32_t sid, T a, T b, T* x)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (n == 0 || x), HPB_ERR_INVALID, "hpb_random_uniform: null argument");
    if (!n) return HPB_OK;
    hipLaunchKernelGGL((k_random<T>), dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, seed, sid, a, b, x);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename T>
static int swap_entry(hpb_handle_t h, size_t n, T* x, T* y)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (n == 0 || (x && y)), HPB_ERR_INVALID, "hpb_swap: null argument");
    if (!n) return HPB_OK;
    hipLaunchKernelGGL((k_swap<T>), dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, x, y);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

template <typename S, typename D>
static int convert_entry(hpb_handle_t h, size_t n, const S* x, D* y)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (n == 0 || (x && y)), HPB_ERR_INVALID, "hpb_convert: null argument");
    if (!n) return HPB_OK;
    hipLaunchKernelGGL((k_convert<S, D>), dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, x, y);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

}  // namespace hpb

using namespace hpb;

#define NUL(T) ((const T*)nullptr)

#define HPB_BOTH(DECL_D, BODY_D, DECL_F, BODY_F) \
    DECL_D { BODY_D }                             \
    DECL_F { BODY_F }

extern "C" {

int hpb_fill_f64(hpb_handle_t h, size_t n, double a, double* x)
{ return map_entry<double, f_fill<double>, 0>(h, n, f_fill<double>{a}, NUL(double), NUL(double), NUL(double), NUL(double), x); }
int hpb_fill_f32(hpb_handle_t h, size_t n, float a, float* x)
{ return map_entry<float, f_fill<float>, 0>(h, n, f_fill<float>{a}, NUL(float), NUL(float), NUL(float), NUL(float), x); }

int hpb_add_mul_scalar_f64(hpb_handle_t h, size_t n, double scalar, double mul_x, double* x)
{ return map_entry<double, f_add_mul_scalar<double>, 1>(h, n, f_add_mul_scalar<double>{scalar, mul_x}, x, NUL(double), NUL(double), NUL(double), x); }
int hpb_add_mul_scalar_f32(hpb_handle_t h, size_t n, float scalar, float mul_x, float* x)
{ return map_entry<float, f_add_mul_scalar<float>, 1>(h, n, f_add_mul_scalar<float>{scalar, mul_x}, x, NUL(float), NUL(float), NUL(float), x); }

int hpb_scal_f64(hpb_handle_t h, size_t n, double a, double* x)
{ return map_entry<double, f_scal<double>, 1>(h, n, f_scal<double>{a}, x, NUL(double), NUL(double), NUL(double), x); }
int hpb_scal_f32(hpb_handle_t h, size_t n, float a, float* x)
{ return map_entry<float, f_scal<float>, 1>(h, n, f_scal<float>{a}, x, NUL(float), NUL(float), NUL(float), x); }

int hpb_copy_f64(hpb_handle_t h, size_t n, const double* x, double* y)
{ return map_entry<double, f_copy<double>, 1>(h, n, f_copy<double>{}, x, NUL(double), NUL(double), NUL(double), y); }
int hpb_copy_f32(hpb_handle_t h, size_t n, const float* x, float* y)
{ return map_entry<float, f_copy<float>, 1>(h, n, f_copy<float>{}, x, NUL(float), NUL(float), NUL(float), y); }

int hpb_swap_f64(hpb_handle_t h, size_t n, double* x, double* y) { return swap_entry<double>(h, n, x, y); }
int hpb_swap_f32(hpb_handle_t h, size_t n, float* x, float* y) { return swap_entry<float>(h, n, x, y); }

int hpb_abs_f64(hpb_handle_t h, size_t n, const double* x, double* y)
{ return map_entry<double, f_abs<double>, 1>(h, n, f_abs<double>{}, x, NUL(double), NUL(double), NUL(double), y); }
int hpb_abs_f32(hpb_handle_t h, size_t n, const float* x, float* y)
{ return map_entry<float, f_abs<float>, 1>(h, n, f_abs<float>{}, x, NUL(float), NUL(float), NUL(float), y); }

int hpb_assign_mul_f64(hpb_handle_t h, size_t n, double a, const double* x, double* y)
{ return map_entry<double, f_scal<double>, 1>(h, n, f_scal<double>{a}, x, NUL(double), NUL(double), NUL(double), y); }
int hpb_assign_mul_f32(hpb_handle_t h, size_t n, float a, const float* x, float* y)
{ return map_entry<float, f_scal<float>, 1>(h, n, f_scal<float>{a}, x, NUL(float), NUL(float), NUL(float), y); }

int hpb_axpy_f64(hpb_handle_t h, size_t n, double a, const double* x, double* y)
{ return map_entry<double, f_axpy<double>, 2>(h, n, f_axpy<double>{a}, x, y, NUL(double), NUL(double), y); }
int hpb_axpy_f32(hpb_handle_t h, size_t n, float a, const float* x, float* y)
{ return map_entry<float, f_axpy<float>, 2>(h, n, f_axpy<float>{a}, x, y, NUL(float), NUL(float), y); }

int hpb_axpby_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z)
{ return map_entry<double, f_axpby<double>, 2>(h, n, f_axpby<double>{a, b}, x, y, NUL(double), NUL(double), z); }
int hpb_axpby_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z)
{ return map_entry<float, f_axpby<float>, 2>(h, n, f_axpby<float>{a, b}, x, y, NUL(float), NUL(float), z); }

int hpb_axpbypcz_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double c, double* z)
{ return map_entry<double, f_axpbypcz<double>, 3>(h, n, f_axpbypcz<double>{a, b, c}, x, y, z, NUL(double), z); }
int hpb_axpbypcz_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float c, float* z)
{ return map_entry<float, f_axpbypcz<float>, 3>(h, n, f_axpbypcz<float>{a, b, c}, x, y, z, NUL(float), z); }

int hpb_axpbypcwpdz_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double c,
                        const double* w, double d, double* z)
{ return map_entry<double, f_axpbypcwpdz<double>, 4>(h, n, f_axpbypcwpdz<double>{a, b, c, d}, x, y, w, z, z); }
int hpb_axpbypcwpdz_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float c,
                        const float* w, float d, float* z)
{ return map_entry<float, f_axpbypcwpdz<float>, 4>(h, n, f_axpbypcwpdz<float>{a, b, c, d}, x, y, w, z, z); }

int hpb_assign_mul4_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double c,
                        const double* v, double d, const double* w, double* z)
{ return map_entry<double, f_axpbypcwpdz<double>, 4>(h, n, f_axpbypcwpdz<double>{a, b, c, d}, x, y, v, w, z); }
int hpb_assign_mul4_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float c,
                        const float* v, float d, const float* w, float* z)
{ return map_entry<float, f_axpbypcwpdz<float>, 4>(h, n, f_axpbypcwpdz<float>{a, b, c, d}, x, y, v, w, z); }

int hpb_mul_pointwise_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z)
{ return map_entry<double, f_mulpw<double>, 2>(h, n, f_mulpw<double>{a, b}, x, y, NUL(double), NUL(double), z); }
int hpb_mul_pointwise_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z)
{ return map_entry<float, f_mulpw<float>, 2>(h, n, f_mulpw<float>{a, b}, x, y, NUL(float), NUL(float), z); }
int hpb_mul_pointwise_inplace_f64(hpb_handle_t h, size_t n, double* x, double b, const double* y)
{ return map_entry<double, f_mulpw_ip<double>, 2>(h, n, f_mulpw_ip<double>{b}, x, y, NUL(double), NUL(double), x); }
int hpb_mul_pointwise_inplace_f32(hpb_handle_t h, size_t n, float* x, float b, const float* y)
{ return map_entry<float, f_mulpw_ip<float>, 2>(h, n, f_mulpw_ip<float>{b}, x, y, NUL(float), NUL(float), x); }

int hpb_div_pointwise_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z)
{ return map_entry<double, f_divpw<double>, 2>(h, n, f_divpw<double>{a, b}, x, y, NUL(double), NUL(double), z); }
int hpb_div_pointwise_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z)
{ return map_entry<float, f_divpw<float>, 2>(h, n, f_divpw<float>{a, b}, x, y, NUL(float), NUL(float), z); }
int hpb_div_pointwise_inplace_f64(hpb_handle_t h, size_t n, double* x, double b, const double* y)
{ return map_entry<double, f_divpw_ip<double>, 2>(h, n, f_divpw_ip<double>{b}, x, y, NUL(double), NUL(double), x); }
int hpb_div_pointwise_inplace_f32(hpb_handle_t h, size_t n, float* x, float b, const float* y)
{ return map_entry<float, f_divpw_ip<float>, 2>(h, n, f_divpw_ip<float>{b}, x, y, NUL(float), NUL(float), x); }

int hpb_axpbyc_f64(hpb_handle_t h, size_t n, double a, const double* x, double b, const double* y, double* z)
{ return map_entry<double, f_axpbyc<double>, 2>(h, n, f_axpbyc<double>{a, b}, x, y, NUL(double), NUL(double), z); }
int hpb_axpbyc_f32(hpb_handle_t h, size_t n, float a, const float* x, float b, const float* y, float* z)
{ return map_entry<float, f_axpbyc<float>, 2>(h, n, f_axpbyc<float>{a, b}, x, y, NUL(float), NUL(float), z); }

int hpb_valid_f64(hpb_handle_t h, size_t n, const double* x, int* out) { return valid_entry<double>(h, n, x, out); }
int hpb_valid_f32(hpb_handle_t h, size_t n, const float* x, int* out) { return valid_entry<float>(h, n, x, out); }

int hpb_set_value_f64(hpb_handle_t h, size_t n, size_t at, double v, double* x) { return set_value_entry<double>(h, n, at, v, x); }
int hpb_set_value_f32(hpb_handle_t h, size_t n, size_t at, float v, float* x) { return set_value_entry<float>(h, n, at, v, x); }
int hpb_get_value_f64(hpb_handle_t h, size_t n, size_t at, const double* x, double* o) { return get_value_entry<double>(h, n, at, x, o); }
int hpb_get_value_f32(hpb_handle_t h, size_t n, size_t at, const float* x, float* o) { return get_value_entry<float>(h, n, at, x, o); }

int hpb_random_uniform_f64(hpb_handle_t h, size_t n, uint64_t seed, uint32_t sid, double a, double b, double* x)
{ return random_entry<double>(h, n, seed, sid, a, b, x); }
int hpb_random_uniform_f32(hpb_handle_t h, size_t n, uint64_t seed, uint32_t sid, float a, float b, float* x)
{ return random_entry<float>(h, n, seed, sid, a, b, x); }

int hpb_convert_f32_to_f64(hpb_handle_t h, size_t n, const float* x, double* y) { return convert_entry<float, double>(h, n, x, y); }
int hpb_convert_f64_to_f32(hpb_handle_t h, size_t n, const double* x, float* y) { return convert_entry<double, float>(h, n, x, y); }

int hpb_scale_convert_f64_to_f32(hpb_handle_t h, size_t n, double s, const double* x, float* y)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (n == 0 || (x && y)), HPB_ERR_INVALID, "hpb_scale_convert_f64_to_f32: null argument");
    if (!n) return HPB_OK;
    hipLaunchKernelGGL(k_scale_to_f32, dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, s, x, y);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

int hpb_axpy_f32_to_f64(hpb_handle_t h, size_t n, double a, const float* x, double* y)
{
    HPB_CLEAR_STALE();
    HPB_REQUIRE(h && (n == 0 || (x && y)), HPB_ERR_INVALID, "hpb_axpy_f32_to_f64: null argument");
    if (!n) return HPB_OK;
    hipLaunchKernelGGL(k_axpy_f32_f64, dim3(simple_grid(h, n)), dim3(256), 0, h->stream, n, a, x, y);
    HPB_LAUNCH_CHECK();
    return HPB_OK;
}

}  // extern "C"
