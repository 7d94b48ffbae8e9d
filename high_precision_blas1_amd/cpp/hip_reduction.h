// hip_reduction.h -- the plain (uncompensated) reduction object on the gfx950 library.
//
// Drop-in for gpu_reduction<T, T_vec, BLOCK_SIZE, threads_r>
// (source/common/gpu_reduction.h:36-100), the "G" column of the reference's
// condition-number driver (source/obtain_condition_precision.cpp:80,152): constructed
// with the vector length; min_max, sum, sum_debug, asum and dot of device vectors
// of that length, returned as host scalars.  Differences:
//  * single-launch kernels over the handle's O(grid) workspace (the reference holds
//    N device + N pinned host words, gpu_reduction.h:43-47);
//  * dot accumulates with fma; sum_debug is sum (the reference's debug variant
//    prints its partial sums);
//  * min_max skips NaN entries and starts from +-infinity; the reference's start
//    value is the integer 0x7f800000 converted to T (gpu_reduction.h:33), so it
//    caps its results at 2139095040 in magnitude;
//  * BLOCK_SIZE / threads_r are accepted and not used.
// The object creates its own handle unless it is given one.
#pragma once
#include <cstddef>
#include <type_traits>
#include <utility>

#include "hip_vector_operations.h"

namespace hpb {

template <class T, class T_vec = T*, int BLOCK_SIZE = 1024, int threads_r = 64>
class hip_reduction {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "hip_reduction: real FP64 / FP32 only");
    static_assert(std::is_same<T_vec, T*>::value, "hip_reduction: vectors are device pointers T*");

public:
    using min_max_t = std::pair<T, T>;

    explicit hip_reduction(size_t vec_size, hpb_handle_t handle = nullptr) : n_(vec_size), h_(handle)
    {
        if (!h_) {
            check(hpb_create(&h_, nullptr), "hip_reduction");
            own_ = true;
        }
    }
    hip_reduction(const hip_reduction&) = delete;
    hip_reduction& operator=(const hip_reduction&) = delete;
    ~hip_reduction()
    {
        if (own_) hpb_destroy(h_);
    }

    min_max_t min_max(const T_vec d_in)
    {
        T o[2];
        check(abi<T>::minmax(h_, n_, d_in, o), "hip_reduction::min_max");
        return min_max_t(o[0], o[1]);
    }
    T sum(const T_vec d_in) { return one(abi<T>::sum, d_in, "hip_reduction::sum"); }
    T sum_debug(const T_vec d_in) { return sum(d_in); }
    T asum(const T_vec d_in) { return one(abi<T>::asum, d_in, "hip_reduction::asum"); }
    T dot(const T_vec d1_in, const T_vec d2_in)
    {
        T o[4];
        check(abi<T>::dot(h_, n_, d1_in, d2_in, o), "hip_reduction::dot");
        return o[2];
    }

    size_t get_vector_size() const { return n_; }
    hpb_handle_t handle() const { return h_; }

private:
    template <class F>
    T one(F f, const T_vec x, const char* what)
    {
        T o[4];
        check(f(h_, n_, x, o), what);
        return o[2];
    }

    size_t n_;
    hpb_handle_t h_;
    bool own_ = false;
};

}  // namespace hpb
