// hpb_cpp_instances.cpp -- explicit instantiation of the drop-in C++ templates.
//
// Every member of the header-only drop-in classes is compiled for FP64 and FP32
// here, with the library, so that a member that no solver happens to use still
// builds: the VectorOperations concept (hip_vector_operations), the CSR matrix,
// operator and preconditioners, the reduction objects, the exact dot, the
// vector-pair generator, the monitor and the three solvers.
#include "csr/hip_matrix.h"
#include "csr/linear_operator.h"
#include "csr/matrix_market_reader.h"
#include "csr/preconditioner.h"
#include "generate_vector_pair.h"
#include "high_prec/dot_product_exact.h"
#include "hip_reduction.h"
#include "hip_reduction_ogita.h"
#include "hip_vector_operations.h"
#include "lin_solvers/bicgstab.h"
#include "lin_solvers/bicgstabl.h"
#include "lin_solvers/default_monitor.h"
#include "lin_solvers/gmres.h"
#include "utils/hpb_log.h"

template struct hpb::hip_vector_operations<double>;
template struct hpb::hip_vector_operations<float>;
template class csr::hip_matrix<double>;
template class csr::hip_matrix<float>;
template class hpb::hip_reduction_ogita<double, double*>;
template class hpb::hip_reduction_ogita<float, float*>;
template class hpb::hip_reduction<double, double*>;
template class hpb::hip_reduction<float, float*>;
template class hpb::dot_product_exact<double, double*>;
template class hpb::dot_product_exact<float, float*>;
template class csr::matrix_market_reader<double>;
template class csr::matrix_market_reader<float>;

namespace hpb_instances {
template <typename T>
struct types {
    using vo_t = hpb::hip_vector_operations<T>;
    using vod_t = hpb::hip_vector_operations<double>;
    using mat_t = csr::hip_matrix<T>;
    using op_t = csr::linear_operator<vo_t, mat_t>;
    using id_t = numerical_algos::lin_solvers::identity_preconditioner<vo_t, op_t>;
    using prec_t = csr::preconditioner<vo_t, op_t, mat_t>;
    using ilu_t = csr::ilu0_preconditioner<vo_t, op_t, mat_t>;
    using mon_t = numerical_algos::lin_solvers::default_monitor<vo_t, hpb::log_std>;
    using gen_t = hpb::generate_vector_pair<vo_t, vod_t, hpb::dot_product_exact<double, double*>>;
};
}  // namespace hpb_instances

template struct csr::linear_operator<hpb_instances::types<double>::vo_t, hpb_instances::types<double>::mat_t>;
template struct csr::linear_operator<hpb_instances::types<float>::vo_t, hpb_instances::types<float>::mat_t>;
template struct csr::preconditioner<hpb_instances::types<double>::vo_t, hpb_instances::types<double>::op_t,
                                    hpb_instances::types<double>::mat_t>;
template struct csr::preconditioner<hpb_instances::types<float>::vo_t, hpb_instances::types<float>::op_t,
                                    hpb_instances::types<float>::mat_t>;
template struct csr::ilu0_preconditioner<hpb_instances::types<double>::vo_t, hpb_instances::types<double>::op_t,
                                         hpb_instances::types<double>::mat_t>;
template struct csr::ilu0_preconditioner<hpb_instances::types<float>::vo_t, hpb_instances::types<float>::op_t,
                                         hpb_instances::types<float>::mat_t>;
template class numerical_algos::lin_solvers::default_monitor<hpb_instances::types<double>::vo_t, hpb::log_std>;
template class numerical_algos::lin_solvers::default_monitor<hpb_instances::types<float>::vo_t, hpb::log_std>;
template class hpb::generate_vector_pair<hpb_instances::types<double>::vo_t, hpb_instances::types<double>::vod_t,
                                         hpb::dot_product_exact<double, double*>>;
template class hpb::generate_vector_pair<hpb_instances::types<float>::vo_t, hpb_instances::types<float>::vod_t,
                                         hpb::dot_product_exact<double, double*>>;

#define HPB_SOLVERS(T, PREC)                                                                                   \
    template class numerical_algos::lin_solvers::gmres<                                                        \
        hpb_instances::types<T>::op_t, hpb_instances::types<T>::PREC, hpb_instances::types<T>::vo_t,           \
        hpb_instances::types<T>::mon_t, hpb::log_std, 256>;                                                     \
    template class numerical_algos::lin_solvers::bicgstabl<                                                    \
        hpb_instances::types<T>::op_t, hpb_instances::types<T>::PREC, hpb_instances::types<T>::vo_t,           \
        hpb_instances::types<T>::mon_t, hpb::log_std>;                                                          \
    template class numerical_algos::lin_solvers::bicgstab<                                                     \
        hpb_instances::types<T>::op_t, hpb_instances::types<T>::PREC, hpb_instances::types<T>::vo_t,           \
        hpb_instances::types<T>::mon_t, hpb::log_std>;

HPB_SOLVERS(double, id_t)
HPB_SOLVERS(float, id_t)
HPB_SOLVERS(double, ilu_t)
HPB_SOLVERS(float, ilu_t)
