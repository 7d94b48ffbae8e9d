// hip_reduction_ogita.h -- the compensated reduction object on the gfx950 library.
//
// Drop-in for gpu_reduction_ogita<T, T_vec, BLOCK_SIZE, threads_r>
// (source/common/ogita/gpu_reduction_ogita.h:36-157), which the reference's drivers
// construct with the vector length and call as .sum / .asum / .dot / .norm on
// device vectors of that length (source/obtain_condition_precision.cpp:80-81,
// source/test_vector_operations.cpp:60-61).  Same construction, same member
// names and meanings, host scalars returned after the handle's stream has
// finished.  Differences:
//  * one single-launch kernel per call over the handle's O(grid) workspace; the
//    reference allocates 2N device and 2N pinned host words in the constructor
//    (gpu_reduction_ogita.h:44-52) and runs three launches with two blocking
//    copies per call;
//  * asum is the compensated sum of |x_i|; the reference's real-T kernel sums x
//    itself (the absolute value is taken only for complex T);
//  * norm is safe against overflow and underflow (see hpb_nrm2c_* in hpb.h);
//  * BLOCK_SIZE / threads_r are accepted and not used: the launch shape is the
//    handle's tuning (hpb_set_tuning).
// The object creates its own handle (a new HIP stream) unless it is given one.
// Errors throw std::runtime_error with the library's message.
#pragma once
#include <cstddef>
#include <type_traits>

#include "hip_vector_operations.h"

namespace hpb {

template <class T, class T_vec = T*, int BLOCK_SIZE = 1024, int threads_r = 64>
class hip_reduction_ogita {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "hip_reduction_ogita: real FP64 / FP32 only (complex x^H y is hpb_zdot2_*)");
    static_assert(std::is_same<T_vec, T*>::value, "hip_reduction_ogita: vectors are device pointers T*");

public:
    using T_real = T;

    explicit hip_reduction_ogita(size_t vec_size, hpb_handle_t handle = nullptr) : n_(vec_size), h_(handle)
    {
        if (!h_) {
            check(hpb_create(&h_, nullptr), "hip_reduction_ogita");
            own_ = true;
        }
    }
    hip_reduction_ogita(const hip_reduction_ogita&) = delete;
    hip_reduction_ogita& operator=(const hip_reduction_ogita&) = delete;
    ~hip_reduction_ogita()
    {
        if (own_) hpb_destroy(h_);
    }

    // compensated Sum2 of x (gpu_reduction_ogita.h:74-78)
    T sum(const T_vec d_in) { return call1(abi<T>::sum2, d_in, "hip_reduction_ogita::sum")[2]; }
    // compensated sum of |x_i| (gpu_reduction_ogita.h:79-84)
    T_real asum(const T_vec d_in) { return call1(abi<T>::asum2, d_in, "hip_reduction_ogita::asum")[2]; }
    // compensated Dot2 (gpu_reduction_ogita.h:85-89)
    T dot(const T_vec d1_in, const T_vec d2_in)
    {
        check(abi<T>::dot2(h_, n_, d1_in, d2_in, last_), "hip_reduction_ogita::dot");
        return last_[2];
    }
    // sqrt(Dot2(x, x)) (gpu_reduction_ogita.h:91-96)
    T_real norm(const T_vec d_in) { return call1(abi<T>::nrm2c, d_in, "hip_reduction_ogita::norm")[3]; }

    // The last call's unrounded Ogita-Rump-Oishi pair (s, c): the compensated
    // result before its final rounding fl(s + c).
    T last_s() const { return last_[0]; }
    T last_c() const { return last_[1]; }

    size_t get_vector_size() const { return n_; }
    hpb_handle_t handle() const { return h_; }

private:
    template <class F>
    const T* call1(F f, const T_vec x, const char* what)
    {
        check(f(h_, n_, x, last_), what);
        return last_;
    }

    size_t n_;
    hpb_handle_t h_;
    bool own_ = false;
    T last_[4] = {T(0), T(0), T(0), T(0)};
};

}  // namespace hpb
