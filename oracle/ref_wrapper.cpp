// C wrapper around the reference's own CPU code, compiled from the reference
// sources in place (see build_oracle.py).  TEST INFRASTRUCTURE ONLY: it is used
// to pin the oracle's restatement and as the "reference" CPU baseline.
//
// Wrapped:
//   dot_product<T>::dot_ogita / dot_naive     source/common/dot_product.h:96-121
//   threaded_reduction<T,T_vec>::dot/sum/asum source/common/threaded_reduction.h:307-399
//   error_bounds<T>::generate_real_dot/sum    source/high_prec/error_bounds.hpp:65-73
#include <cstddef>
#include <common/dot_product.h>
#include <common/threaded_reduction.h>
#include <high_prec/error_bounds.hpp>

extern "C" {

double ref_dot2(size_t n, const double* x, const double* y)
{
    dot_product<double> d(n, 1);
    return d.dot_ogita(x, y);
}

double ref_dot_naive(size_t n, const double* x, const double* y)
{
    dot_product<double> d(n, 0);
    return d.dot_naive(x, y);
}

float ref_dot2_f32(size_t n, const float* x, const float* y)
{
    dot_product<float> d(n, 1);
    return d.dot_ogita(x, y);
}

// threaded_reduction takes T_vec by reference and indexes it; a raw pointer works.
double ref_dot2_threaded(size_t n, const double* x, const double* y, int parts)
{
    threaded_reduction<double, const double*> r(n, parts, 1);
    return r.dot(x, y);
}

double ref_dot_naive_threaded(size_t n, const double* x, const double* y, int parts)
{
    threaded_reduction<double, const double*> r(n, parts, 0);
    return r.dot(x, y);
}

double ref_sum2_threaded(size_t n, const double* x, int parts)
{
    threaded_reduction<double, const double*> r(n, parts, 1);
    return r.sum(x);
}

double ref_asum2_threaded(size_t n, const double* x, int parts)
{
    threaded_reduction<double, const double*> r(n, parts, 1);
    return r.asum(x);
}

// out6: compensated/naive dot bounds in the order of update_real_dots
// (error_bounds.hpp:288-297): naive sequential, naive block, naive pairwise,
// compensated pairwise, ORO sequential, compensated block.
void ref_dot_bounds(size_t n, double cond, size_t k, double* out6)
{
    error_bounds<double> eb;
    eb.generate_real_dot(n, cond, k);
    out6[0] = (double)eb.dot.real.base.sequential;
    out6[1] = (double)eb.dot.real.base.block_parallel;
    out6[2] = (double)eb.dot.real.base.pairwise_parallel;
    out6[3] = (double)eb.dot.real.compensated.pairwise_parallel;
    out6[4] = (double)eb.dot.real.compensated.sequential;
    out6[5] = (double)eb.dot.real.compensated.block_parallel;
}

void ref_sum_bounds(size_t n, double cond, size_t k, double* out6)
{
    error_bounds<double> eb;
    eb.generate_real_sum(n, cond, k);
    out6[0] = (double)eb.sum.real.base.sequential;
    out6[1] = (double)eb.sum.real.base.block_parallel;
    out6[2] = (double)eb.sum.real.base.pairwise_parallel;
    out6[3] = (double)eb.sum.real.compensated.pairwise_parallel;
    out6[4] = (double)eb.sum.real.compensated.sequential;
    out6[5] = (double)eb.sum.real.compensated.block_parallel;
}

}
