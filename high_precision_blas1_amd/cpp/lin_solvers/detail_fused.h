// detail_fused.h -- optional fused vector operations used by the solvers.
//
// A VectorOperations type may provide
//   void multi_scalar_prod(int k, const vector_type* V, const vector_type& w, scalar_type* out) const;
//   void multi_add_mul(int k, const scalar_type* a, const vector_type* V, vector_type& y) const;  // y += sum a_j V_j
// (one pass over w / y for all k vectors).  When it does not, these helpers fall
// back to k separate scalar_prod / add_mul calls, which is what the reference's
// concept offers (gpu_vector_operations.h:257-280, 459-462).
#pragma once
#include <type_traits>
#include <utility>

namespace numerical_algos {
namespace lin_solvers {
namespace detail {

template <class VO, class = void>
struct has_multi : std::false_type {};
template <class VO>
struct has_multi<VO, std::void_t<decltype(std::declval<const VO&>().multi_scalar_prod(
                         0, (const typename VO::vector_type*)nullptr, std::declval<const typename VO::vector_type&>(),
                         (typename VO::scalar_type*)nullptr)),
                                 decltype(std::declval<const VO&>().multi_add_mul(
                                     0, (const typename VO::scalar_type*)nullptr,
                                     (const typename VO::vector_type*)nullptr,
                                     std::declval<typename VO::vector_type&>()))>> : std::true_type {};

template <class VO>
void multi_dot(const VO& vo, int k, const typename VO::vector_type* V, const typename VO::vector_type& w,
               typename VO::scalar_type* out)
{
    if constexpr (has_multi<VO>::value) {
        vo.multi_scalar_prod(k, V, w, out);
    } else {
        for (int j = 0; j < k; ++j) out[j] = vo.scalar_prod(V[j], w);
    }
}

template <class VO>
void multi_axpy(const VO& vo, int k, const typename VO::scalar_type* a, const typename VO::vector_type* V,
                typename VO::vector_type& y)
{
    if constexpr (has_multi<VO>::value) {
        vo.multi_add_mul(k, a, V, y);
    } else {
        for (int j = 0; j < k; ++j) vo.add_mul(a[j], V[j], y);
    }
}

}  // namespace detail
}  // namespace lin_solvers
}  // namespace numerical_algos
