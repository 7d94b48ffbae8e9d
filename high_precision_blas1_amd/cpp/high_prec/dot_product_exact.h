// dot_product_exact.h -- exact dot product object on the gfx950 library.
//
// Drop-in for dot_product_gmp<T, T_vec> (source/high_prec/dot_product_gmp.hpp:38-164),
// the reference's exact oracle: set_arrays(N, x, y), dot_exact(), and the error
// measures get_error / get_error_T / get_error_relative / get_error_relative_T
// against the exact value.  The reference sums the products in GMP mpf at
// 512-1024 bits on the host (or in CUMP on the GPU after use_gpu); here the
// device's fixed-point long accumulator (hpb_exact_dot_*) gives the exact value
// as an unevaluated pair hi + lo with hi = round-to-nearest(exact),
// lo = round(exact - hi).  Differences:
//  * use_cpu (the default, as in the reference) takes host arrays and stages them
//    through device memory; use_gpu takes device pointers; both compute on the GPU;
//  * dot_exact() returns round-to-nearest of the exact value; the reference's
//    mpf get_d truncates toward zero;
//  * the error measures subtract hi + lo from the approximation with error-free
//    transformations in double, relative error about 2^-104, where the reference
//    subtracts in mpf;
//  * the precision argument of the constructor is accepted and not needed.
#pragma once
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <type_traits>

#include "../hip_vector_operations.h"

namespace hpb {

template <class T, class T_vec = T*>
class dot_product_exact {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "dot_product_exact: real FP64 / FP32 only");

public:
    explicit dot_product_exact(unsigned int exact_prec_bits = 512, hpb_handle_t handle = nullptr)
        : bits_(exact_prec_bits), h_(handle)
    {
        if (!h_) {
            check(hpb_create(&h_, nullptr), "dot_product_exact");
            own_ = true;
        }
    }
    dot_product_exact(const dot_product_exact&) = delete;
    dot_product_exact& operator=(const dot_product_exact&) = delete;
    ~dot_product_exact()
    {
        release_();
        if (own_) hpb_destroy(h_);
    }

    // arrays given to set_arrays are device pointers after use_gpu, host arrays
    // after use_cpu (dot_product_gmp.hpp:72-100)
    void use_gpu(size_t = 0) { gpu_ = true; }
    void use_cpu() { gpu_ = false; }

    void set_arrays(size_t N, T_vec x, T_vec y) const
    {
        n_ = N;
        x_ = x;
        y_ = y;
        set_ = true;
    }
    // the reference forwards a changed range to its CUMP copy; the arrays are read
    // at dot_exact() time here, so there is nothing to update
    void update_arrays(size_t, size_t, T_vec&, T_vec&) {}

    T dot_exact() const
    {
        if (!set_) throw std::logic_error("dot_product_exact::dot_exact: set_arrays was not called");
        double o[2] = {0.0, 0.0};
        if (n_ > 0) {
            const T* dx = x_;
            const T* dy = y_;
            if (!gpu_) {
                stage_(x_, y_);
                dx = sx_;
                dy = sy_;
            }
            check(abi<T>::exact_dot(h_, n_, dx, dy, o), "dot_product_exact::dot_exact");
        }
        hi_ = o[0];
        lo_ = o[1];
        return T(hi_);
    }
    // round(a + b) (dot_product_gmp.hpp:128-132): one IEEE addition is already that
    T sum_two(const T one, const T two) { return one + two; }

    void print_res() const { std::printf("%.17g + %.17g\n", hi_, lo_); }

    // |approx - exact|
    T get_error(const T& approx) const { return T(std::fabs(diff_(approx))); }
    // |approx - double(exact)|
    T get_error_T(const T& approx) const { return T(std::fabs(double(approx) - hi_)); }
    // |(approx - exact) / exact|
    T get_error_relative(const T& approx) const { return T(std::fabs(diff_(approx) / hi_)); }
    // |(approx - double(exact)) / double(exact)|
    T get_error_relative_T(const T& approx) const { return T(std::fabs((double(approx) - hi_) / hi_)); }

    // the exact value as hi + lo
    double exact_hi() const { return hi_; }
    double exact_lo() const { return lo_; }
    unsigned int precision_bits() const { return bits_; }

private:
    // approx - (hi + lo) with TwoSum on approx - hi
    double diff_(T approx) const
    {
        const double a = double(approx), b = -hi_;
        const double s = a + b;
        const double bs = s - a;
        const double t = (b - bs) + (a - (s - bs));
        return s + (t - lo_);
    }
    void stage_(const T* x, const T* y) const
    {
        if (cap_ < n_) {
            release_();
            check(hpb_malloc((void**)&sx_, n_ * sizeof(T)), "dot_product_exact: staging");
            check(hpb_malloc((void**)&sy_, n_ * sizeof(T)), "dot_product_exact: staging");
            cap_ = n_;
        }
        check(hpb_memcpy_h2d(h_, sx_, x, n_ * sizeof(T)), "dot_product_exact: staging");
        check(hpb_memcpy_h2d(h_, sy_, y, n_ * sizeof(T)), "dot_product_exact: staging");
    }
    void release_() const
    {
        if (sx_) hpb_free(sx_);
        if (sy_) hpb_free(sy_);
        sx_ = sy_ = nullptr;
        cap_ = 0;
    }

    unsigned int bits_;
    hpb_handle_t h_;
    bool own_ = false;
    bool gpu_ = false;
    mutable bool set_ = false;
    mutable size_t n_ = 0;
    mutable T_vec x_ = nullptr;
    mutable T_vec y_ = nullptr;
    mutable T* sx_ = nullptr;
    mutable T* sy_ = nullptr;
    mutable size_t cap_ = 0;
    mutable double hi_ = 0.0, lo_ = 0.0;
};

}  // namespace hpb
