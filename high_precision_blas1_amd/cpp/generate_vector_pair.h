// generate_vector_pair.h -- ill-conditioned vector pairs on the gfx950 library.
//
// Drop-in for generate_vector_pair<VectorOperations, VectorOperationsDouble,
// ExactDot, BLOCK_SIZE> (source/generate_vector_pair.hpp:40-216): constructed from
// the vector-operations objects of the working type and of double and an exact-dot
// object; generate(x, y, cond) fills two device vectors with a GenDot pair of the
// requested condition number and returns the condition estimate
// kappa = sum |x_i y_i| / |x . y| of what it produced (as the reference's
// estimate_condition_reduction, generate_vector_pair.hpp:206-216);
// condition_number_max(x, y) returns that estimate for any pair.
//
// The pair comes from the device generator hpb_gendot_f64 (Ogita-Rump-Oishi 2005
// Alg. 6.1, as the reference): head entries m 2^e with exponents uniform in
// [0, log2(cond)/2], a tail of N_fix entries with exponents descending to 0 whose
// y entries cancel the running dot.  N_fix follows the reference's rule
// min(max(U{10000..20000}, log2(cond)/2), round(N/2)) (generate_vector_pair.hpp:
// 127-133).  Differences:
//  * seeded and reproducible: a Philox stream keyed by seed(), and a seeded engine
//    for N_fix; each generate() call advances the seed, so successive calls give
//    new pairs (the reference seeds from std::random_device);
//  * the exact dot of the generated pair is tracked exactly during generation, so
//    dot_exact() / dot_exact_cuda() only record the request; the estimate uses the
//    exact dots of (x, y) and (|x|, |y|) (the reference uses its compensated dot
//    unless an exact one is requested);
//  * FP32: the pair is generated in FP64 and rounded to FP32; the estimate is that
//    of the rounded pair (rounding changes the cancellation, so it can differ from
//    the request much more than in FP64).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <type_traits>

#include "hip_vector_operations.h"

namespace hpb {

template <class VectorOperations, class VectorOperationsDouble, class ExactDot, int BLOCK_SIZE = 1024>
class generate_vector_pair {
    using T = typename VectorOperations::scalar_type;
    using T_vec = typename VectorOperations::vector_type;
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "generate_vector_pair: real FP64 / FP32 vectors only");

public:
    generate_vector_pair(VectorOperations* vec_ops, VectorOperationsDouble* vec_ops_double, ExactDot* exact_dot)
        : vo_(vec_ops), vod_(vec_ops_double), exact_(exact_dot)
    {
        if (!vo_ || !vod_) throw std::invalid_argument("generate_vector_pair: null vector operations");
        sz_ = vo_->get_vector_size();
        if (vod_->get_vector_size() != sz_)
            throw std::invalid_argument("generate_vector_pair: the two vector operations differ in size");
        vod_->init_vector(xd_);
        vod_->start_use_vector(xd_);
        vod_->init_vector(yd_);
        vod_->start_use_vector(yd_);
    }
    generate_vector_pair(const generate_vector_pair&) = delete;
    generate_vector_pair& operator=(const generate_vector_pair&) = delete;
    ~generate_vector_pair()
    {
        vod_->free_vector(xd_);
        vod_->free_vector(yd_);
    }

    void dot_exact() { use_exact_dot_ = 1; }
    void dot_exact_cuda() { use_exact_dot_ = 2; }

    void set_seed(uint64_t s) { seed_ = s; }
    uint64_t seed() const { return seed_; }

    T generate(T_vec& x, T_vec& y, T condition_number, int use_exact_dot_cuda = 0)
    {
        (void)use_exact_dot_cuda;
        if (!(condition_number >= T(1))) throw std::invalid_argument("generate_vector_pair: condition number < 1");
        const double b = std::log2((double)condition_number);
        for (int attempt = 0; attempt < 16; ++attempt) {
            std::mt19937_64 gen(seed_ ^ 0x9e3779b97f4a7c15ull);
            std::uniform_int_distribution<int> distrib(10000, 20000);
            const long size_exp = std::max<long>(distrib(gen), (long)(0.5 * b));
            const long half = std::lround(0.5 * (double)sz_);
            const size_t n_fix = (size_t)std::min<long>(size_exp, half);
            double* gx = std::is_same<T, double>::value ? reinterpret_cast<double*>(x) : xd_;
            double* gy = std::is_same<T, double>::value ? reinterpret_cast<double*>(y) : yd_;
            double dd[3];
            check(hpb_gendot_f64(vo_->handle(), sz_, b, seed_, n_fix, gx, gy, limbs_, dd), "generate_vector_pair");
            last_seed_ = seed_++;
            last_tail_ = n_fix;
            if constexpr (std::is_same<T, float>::value) {
                check(hpb_convert_f64_to_f32(vo_->handle(), sz_, xd_, reinterpret_cast<float*>(x)),
                      "generate_vector_pair");
                check(hpb_convert_f64_to_f32(vo_->handle(), sz_, yd_, reinterpret_cast<float*>(y)),
                      "generate_vector_pair");
            }
            const T est = condition_number_max(x, y);
            if (std::isfinite((double)est)) return est;
        }
        throw std::runtime_error("generate_vector_pair: no pair with a finite condition estimate in 16 attempts");
    }

    // kappa = sum |x_i y_i| / |x . y| from exact dots
    T condition_number_max(T_vec x1, T_vec x2)
    {
        const hpb_handle_t h = vo_->handle();
        double e[2], a[2];
        check(abi<T>::exact_dot(h, sz_, x1, x2, e), "generate_vector_pair: exact dot");
        T* ax = reinterpret_cast<T*>(xd_);   // |x| and |y| in the FP64 scratch vectors
        T* ay = reinterpret_cast<T*>(yd_);
        check(abi<T>::abs(h, sz_, x1, ax), "generate_vector_pair: |x|");
        check(abi<T>::abs(h, sz_, x2, ay), "generate_vector_pair: |y|");
        check(abi<T>::exact_dot(h, sz_, ax, ay, a), "generate_vector_pair: exact dot");
        last_exact_[0] = e[0];
        last_exact_[1] = e[1];
        return T(a[0] / std::fabs(e[0]));
    }

    // exact x . y of the last pair measured by generate / condition_number_max, as hi + lo
    double exact_hi() const { return last_exact_[0]; }
    double exact_lo() const { return last_exact_[1]; }
    // the generator's own 256-bit record of the exact dot of the last generated
    // pair (two's complement, least significant limb first, LSB 2^-106)
    const uint64_t* exact_limbs() const { return limbs_; }
    // seed and tail length N_fix of the last generated pair
    uint64_t last_seed() const { return last_seed_; }
    size_t last_tail_length() const { return last_tail_; }
    ExactDot* exact_dot_object() const { return exact_; }
    int exact_dot_requested() const { return use_exact_dot_; }

private:
    VectorOperations* vo_;
    VectorOperationsDouble* vod_;
    ExactDot* exact_;
    size_t sz_ = 0;
    double* xd_ = nullptr;
    double* yd_ = nullptr;
    uint64_t seed_ = 0;
    uint64_t last_seed_ = 0;
    size_t last_tail_ = 0;
    char use_exact_dot_ = 0;
    uint64_t limbs_[4] = {0, 0, 0, 0};
    double last_exact_[2] = {0.0, 0.0};
};

}  // namespace hpb
