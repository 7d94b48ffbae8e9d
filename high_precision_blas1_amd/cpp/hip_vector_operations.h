// hip_vector_operations.h -- the VectorOperations concept on the gfx950 library.
//
// Drop-in for gpu_vector_operations<T> (source/common/gpu_vector_operations.h:86-538):
// same types (scalar_type, vector_type = T* device pointer, norm_type/Tsc,
// location), lifetime calls (init/start_use/stop_use/free_vector and the _rank1
// forms), host staging (view/set/get/get_buffer/debug_view), the precision
// switch (use_high_precision / use_standard_precision / get_current_precision)
// and every reduction and elementwise method, with the same argument meaning.
// The constructor takes an hpb handle instead of a cublas_wrap*.
//
// High precision selects the Ogita-Rump-Oishi kernels (Dot2, compensated nrm2,
// compensated sum of |x|); standard precision the plain kernels.  Host-scalar
// results synchronise the handle's stream like the reference; the *_dev methods
// leave the result in device memory (used by graph-captured solver code).
// Errors throw std::runtime_error with the library's message (the reference's
// CUDA_SAFE_CALL convention, utils/cuda_safe_call.h:143-147).
#pragma once
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "hpb.h"

namespace hpb {

inline void check(int status, const char* what)
{
    if (status != 0) throw std::runtime_error(std::string(what) + ": " + hpb_last_error());
}

// Per-type entry points of the C ABI.
template <typename T> struct abi;

template <> struct abi<double> {
#define HPB_F(name) static constexpr auto name = &hpb_##name##_f64;
#define HPB_FA(name) static constexpr auto name##_async = &hpb_##name##_f64_async;
    HPB_F(dot2) HPB_F(dot) HPB_FA(dot2) HPB_FA(dot) HPB_F(nrm2c) HPB_F(nrm2) HPB_FA(nrm2c)
    HPB_FA(nrm2) HPB_F(asum2) HPB_F(asum) HPB_F(sum2) HPB_F(sum) HPB_F(fill) HPB_F(add_mul_scalar) HPB_F(scal)
    HPB_F(copy) HPB_F(swap) HPB_F(abs) HPB_F(assign_mul) HPB_F(axpy) HPB_F(axpby) HPB_F(axpbypcz)
    HPB_F(axpbypcwpdz) HPB_F(assign_mul4) HPB_F(mul_pointwise) HPB_F(mul_pointwise_inplace) HPB_F(div_pointwise)
    HPB_F(div_pointwise_inplace) HPB_F(axpbyc) HPB_F(valid) HPB_F(set_value) HPB_F(get_value)
    HPB_F(random_uniform) HPB_F(csr_spmv) HPB_F(csr_spmvc) HPB_F(multidot) HPB_F(multiaxpy) HPB_FA(multidot)
    HPB_F(multiaxpy_dev) HPB_F(csr_trsv) HPB_F(csr_ilu0) HPB_F(minmax) HPB_F(exact_dot)
    HPB_F(csr_transpose)
#undef HPB_F
#undef HPB_FA
};
template <> struct abi<float> {
#define HPB_F(name) static constexpr auto name = &hpb_##name##_f32;
#define HPB_FA(name) static constexpr auto name##_async = &hpb_##name##_f32_async;
    HPB_F(dot2) HPB_F(dot) HPB_FA(dot2) HPB_FA(dot) HPB_F(nrm2c) HPB_F(nrm2) HPB_FA(nrm2c)
    HPB_FA(nrm2) HPB_F(asum2) HPB_F(asum) HPB_F(sum2) HPB_F(sum) HPB_F(fill) HPB_F(add_mul_scalar) HPB_F(scal)
    HPB_F(copy) HPB_F(swap) HPB_F(abs) HPB_F(assign_mul) HPB_F(axpy) HPB_F(axpby) HPB_F(axpbypcz)
    HPB_F(axpbypcwpdz) HPB_F(assign_mul4) HPB_F(mul_pointwise) HPB_F(mul_pointwise_inplace) HPB_F(div_pointwise)
    HPB_F(div_pointwise_inplace) HPB_F(axpbyc) HPB_F(valid) HPB_F(set_value) HPB_F(get_value)
    HPB_F(random_uniform) HPB_F(csr_spmv) HPB_F(csr_spmvc) HPB_F(multidot) HPB_F(multiaxpy) HPB_FA(multidot)
    HPB_F(multiaxpy_dev) HPB_F(csr_trsv) HPB_F(csr_ilu0) HPB_F(minmax) HPB_F(exact_dot)
    HPB_F(csr_transpose)
#undef HPB_F
#undef HPB_FA
};

template <typename T>
struct hip_vector_operations {
    typedef T scalar_type;
    typedef T* vector_type;
    typedef T norm_type;
    typedef T Tsc;
    bool location;

    hip_vector_operations(size_t sz, hpb_handle_t handle) : location(true), sz_(sz), h_(handle)
    {
        if (!h_) throw std::invalid_argument("hip_vector_operations: null hpb handle");
        x_host_.resize(sz_);
    }

    hpb_handle_t handle() const { return h_; }

    // precision switch (gpu_vector_operations.h:129-150)
    void use_high_precision() const { high_ = true; }
    void use_standard_precision() const { high_ = false; }
    bool get_current_precision() const { return high_; }

    // lifetime (gpu_vector_operations.h:152-192)
    void init_vector(vector_type& x) const { x = nullptr; }
    void init_vector_rank1(vector_type& x) const { x = nullptr; }
    void free_vector(vector_type& x) const
    {
        if (x) hpb_free(x);
        x = nullptr;
    }
    void start_use_vector(vector_type& x) const
    {
        if (!x) check(hpb_malloc((void**)&x, sz_ * sizeof(T)), "start_use_vector");
    }
    void start_use_vector_rank1(vector_type& x) const
    {
        if (!x) check(hpb_malloc((void**)&x, (sz_ + 1) * sizeof(T)), "start_use_vector_rank1");
    }
    void stop_use_vector(vector_type&) const {}
    size_t get_vector_size() const { return sz_; }
    size_t get_l2_size() const { return (size_t)std::sqrt(Tsc(sz_)); }
    bool device_location() const { return location; }

    // host staging (gpu_vector_operations.h:197-252)
    vector_type view(const vector_type& x) const
    {
        if (!x) return nullptr;
        check(hpb_memcpy_d2h(h_, x_host_.data(), x, sz_ * sizeof(T)), "view");
        return x_host_.data();
    }
    void set(vector_type& x) const
    {
        if (x) check(hpb_memcpy_h2d(h_, x, x_host_.data(), sz_ * sizeof(T)), "set");
    }
    void set(const vector_type& x_host, vector_type& x) const
    {
        if (x) check(hpb_memcpy_h2d(h_, x, x_host, sz_ * sizeof(T)), "set");
    }
    void get(const vector_type& x, vector_type& x_host) const
    {
        if (x) check(hpb_memcpy_d2h(h_, x_host, x, sz_ * sizeof(T)), "get");
    }
    vector_type get_buffer() const { return x_host_.data(); }
    void debug_view(const vector_type& vec, const std::string& f_name) const
    {
        vector_type xh = view(vec);
        std::ofstream f(f_name);
        for (size_t j = 0; j < sz_; ++j) f << std::scientific << std::setprecision(16) << xh[j] << (j + 1 < sz_ ? "\n" : "");
    }

    // reductions (gpu_vector_operations.h:255-403)
    bool check_is_valid_number(const vector_type& x) const
    {
        int ok = 0;
        check(abi<T>::valid(h_, sz_, x, &ok), "check_is_valid_number");
        return ok != 0;
    }
    scalar_type scalar_prod(const vector_type& x, const vector_type& y) const
    {
        T o[4];
        check((high_ ? abi<T>::dot2 : abi<T>::dot)(h_, sz_, x, y, o), "scalar_prod");
        return o[2];
    }
    void scalar_prod(const vector_type& x, const vector_type& y, scalar_type* result) const
    {
        result[0] = scalar_prod(x, y);
    }
    // device result: dev_out4 = {s, c, s + c, sqrt(s + c)} (no synchronisation)
    void scalar_prod_dev(const vector_type& x, const vector_type& y, scalar_type* dev_out4) const
    {
        check((high_ ? abi<T>::dot2_async : abi<T>::dot_async)(h_, sz_, x, y, dev_out4), "scalar_prod_dev");
    }
    scalar_type absolute_sum(const vector_type& x) const
    {
        T o[4];
        check((high_ ? abi<T>::asum2 : abi<T>::asum)(h_, sz_, x, o), "absolute_sum");
        return o[2];
    }
    void absolute_sum(const vector_type& x, scalar_type* result) const { result[0] = absolute_sum(x); }
    Tsc norm(const vector_type& x) const
    {
        T o[4];
        check((high_ ? abi<T>::nrm2c : abi<T>::nrm2)(h_, sz_, x, o), "norm");
        return o[3];
    }
    void norm(const vector_type& x, Tsc* result) const { result[0] = norm(x); }
    void norm_dev(const vector_type& x, Tsc* dev_out4) const
    {
        check((high_ ? abi<T>::nrm2c_async : abi<T>::nrm2_async)(h_, sz_, x, dev_out4), "norm_dev");
    }
    Tsc norm_l2(const vector_type& x) const { return norm(x) / std::sqrt(Tsc(sz_)); }
    // norms of x extended by one entry val_x (gpu_vector_operations.h:342-391),
    // computed without the reference's per-call allocation: ||(x, v)||^2 = ||x||^2 + v^2.
    Tsc norm_rank1(const vector_type& x, const scalar_type val_x) const
    {
        T o[4];
        check((high_ ? abi<T>::nrm2c : abi<T>::nrm2)(h_, sz_, x, o), "norm_rank1");
        return rank1_sqrt(o[0], o[1], val_x);
    }
    Tsc norm_rank1_l2(const vector_type& x, const scalar_type val_x) const
    {
        T o[4];
        check((high_ ? abi<T>::nrm2c : abi<T>::nrm2)(h_, sz_, x, o), "norm_rank1_l2");
        return rank1_sqrt(o[0], o[1], Tsc(sz_) * val_x) / std::sqrt(Tsc(sz_));
    }
    Tsc normalize(vector_type& x) const
    {
        Tsc nr = norm(x);
        check(abi<T>::scal(h_, sz_, T(1) / nr, x), "normalize");
        return nr;
    }

    // elementwise (gpu_vector_operations.h:404-493, impl.cuh:226-528)
    void assign_scalar(const scalar_type scalar, vector_type& x) const
    {
        check(abi<T>::fill(h_, sz_, scalar, x), "assign_scalar");
    }
    void add_mul_scalar(const scalar_type scalar, const scalar_type mul_x, vector_type& x) const
    {
        check(abi<T>::add_mul_scalar(h_, sz_, scalar, mul_x, x), "add_mul_scalar");
    }
    void scale(const T alpha, vector_type& x) const { check(abi<T>::scal(h_, sz_, alpha, x), "scale"); }
    void assign(const vector_type& x, vector_type& y) const { check(abi<T>::copy(h_, sz_, x, y), "assign"); }
    void swap(vector_type& x, vector_type& y) const { check(abi<T>::swap(h_, sz_, x, y), "swap"); }
    void make_abs_copy(const vector_type& x, vector_type& y) const { check(abi<T>::abs(h_, sz_, x, y), "make_abs_copy"); }
    void make_abs(vector_type& x) const { check(abi<T>::abs(h_, sz_, x, x), "make_abs"); }
    void assign_mul(const scalar_type mul_x, const vector_type& x, vector_type& y) const
    {
        check(abi<T>::assign_mul(h_, sz_, mul_x, x, y), "assign_mul");
    }
    // y := y + mul_x*x (the reference calls cublas axpy here)
    void add_mul(scalar_type mul_x, const vector_type& x, vector_type& y) const
    {
        check(abi<T>::axpy(h_, sz_, mul_x, x, y), "add_mul");
    }
    // z := mul_x*x + mul_y*y
    void assign_mul(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                    vector_type& z) const
    {
        check(abi<T>::axpby(h_, sz_, mul_x, x, mul_y, y, z), "assign_mul");
    }
    // y := mul_x*x + mul_y*y
    void add_mul(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, vector_type& y) const
    {
        check(abi<T>::axpby(h_, sz_, mul_x, x, mul_y, y, y), "add_mul");
    }
    // z := mul_x*x + mul_y*y + mul_z*z
    void add_mul(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                 const scalar_type mul_z, vector_type& z) const
    {
        check(abi<T>::axpbypcz(h_, sz_, mul_x, x, mul_y, y, mul_z, z), "add_mul");
    }
    void mul_pointwise(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                       vector_type& z) const
    {
        check(abi<T>::mul_pointwise(h_, sz_, mul_x, x, mul_y, y, z), "mul_pointwise");
    }
    void mul_pointwise(vector_type& x, const scalar_type mul_y, const vector_type& y) const
    {
        check(abi<T>::mul_pointwise_inplace(h_, sz_, x, mul_y, y), "mul_pointwise");
    }
    void div_pointwise(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                       vector_type& z) const
    {
        check(abi<T>::div_pointwise(h_, sz_, mul_x, x, mul_y, y, z), "div_pointwise");
    }
    void div_pointwise(vector_type& x, const scalar_type mul_y, const vector_type& y) const
    {
        check(abi<T>::div_pointwise_inplace(h_, sz_, x, mul_y, y), "div_pointwise");
    }
    // z := mul_x*x + mul_y*y + mul_w*w + mul_z*z
    void add_mul(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                 const scalar_type mul_w, const vector_type& w, const scalar_type mul_z, vector_type& z) const
    {
        check(abi<T>::axpbypcwpdz(h_, sz_, mul_x, x, mul_y, y, mul_w, w, mul_z, z), "add_mul");
    }
    // z := mul_x*x + mul_y*y + mul_v*v + mul_w*w
    void assign_mul(scalar_type mul_x, const vector_type& x, const scalar_type mul_y, const vector_type& y,
                    scalar_type mul_v, const vector_type& v, const scalar_type mul_w, const vector_type& w,
                    vector_type& z) const
    {
        check(abi<T>::assign_mul4(h_, sz_, mul_x, x, mul_y, y, mul_v, v, mul_w, w, z), "assign_mul");
    }
    // compensated z := mul_x*x + mul_y*y (error-free split, faithful result)
    void assign_mul_compensated(const scalar_type mul_x, const vector_type& x, const scalar_type mul_y,
                                const vector_type& y, vector_type& z) const
    {
        check(abi<T>::axpbyc(h_, sz_, mul_x, x, mul_y, y, z), "assign_mul_compensated");
    }
    // fused Gram-Schmidt helpers: out_j = (V_j, w) for j < k, reading w once;
    // y := y + sum_j a_j V_j in one pass (compensated in high precision)
    void multi_scalar_prod(int k, const vector_type* V, const vector_type& w, scalar_type* out) const
    {
        T buf[4 * 64];
        for (int j0 = 0; j0 < k; j0 += 64) {
            const int kk = (k - j0) < 64 ? (k - j0) : 64;
            check(abi<T>::multidot(h_, kk, (const T* const*)(V + j0), sz_, w, high_ ? 1 : 0, buf), "multi_scalar_prod");
            for (int j = 0; j < kk; ++j) out[j0 + j] = buf[4 * j + 2];
        }
    }
    void multi_add_mul(int k, const scalar_type* a, const vector_type* V, vector_type& y) const
    {
        for (int j0 = 0; j0 < k; j0 += 64) {
            const int kk = (k - j0) < 64 ? (k - j0) : 64;
            check(abi<T>::multiaxpy(h_, kk, a + j0, (const T* const*)(V + j0), sz_, y, high_ ? 1 : 0), "multi_add_mul");
        }
    }
    void set_value_at_point(scalar_type val_x, size_t at, vector_type& x) const
    {
        check(abi<T>::set_value(h_, sz_, at, val_x, x), "set_value_at_point");
    }
    void set_value_at_point(scalar_type val_x, size_t at, vector_type& x, size_t sz_l) const
    {
        check(abi<T>::set_value(h_, sz_l, at, val_x, x), "set_value_at_point");
    }
    T get_value_at_point(size_t at, const vector_type& x) const
    {
        T v;
        check(abi<T>::get_value(h_, sz_, at, x, &v), "get_value_at_point");
        return v;
    }
    // assign_random (gpu_vector_operations.h:496-523), seeded: values in (0, 1] or (a, b]
    void assign_random(vector_type& vec, uint64_t seed = 0, uint32_t stream_id = 0) const
    {
        check(abi<T>::random_uniform(h_, sz_, seed, stream_id, T(0), T(1), vec), "assign_random");
    }
    void assign_random(vector_type& vec, scalar_type a, scalar_type b, uint64_t seed = 0, uint32_t stream_id = 0) const
    {
        check(abi<T>::random_uniform(h_, sz_, seed, stream_id, a, b, vec), "assign_random");
    }

private:
    static Tsc rank1_sqrt(T s, T c, T v)
    {
        // (s, c) + v^2 with one more Dot2 step
        T p = v * v;
        T e = std::fma(v, v, -p);
        T ss = s + p;
        T bs = ss - s;
        T t = (p - bs) + (s - (ss - bs));
        return std::sqrt(ss + (c + (e + t)));
    }

    size_t sz_;
    hpb_handle_t h_;
    mutable bool high_ = false;
    mutable std::vector<T> x_host_;
};

}  // namespace hpb
