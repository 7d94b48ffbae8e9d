// linear_operator.h -- LinearOperator concept over a CSR matrix.
//
// Same as csr::linear_operator (source/common/csr/linear_operator.h:29-47):
// apply(x, f) computes f := A x through the matrix's axpy with alpha 1, beta 0.
#pragma once

namespace csr {

template <class VectorOperations, class Matrix>
struct linear_operator {
    using T = typename VectorOperations::scalar_type;
    using T_vec = typename VectorOperations::vector_type;

    linear_operator(const VectorOperations* vec_ops, const Matrix* mat) : vec_ops_(vec_ops), mat_(mat) {}
    void apply(const T_vec& x, T_vec& f) const { mat_->axpy(T(1), x, T(0), f); }
    const Matrix* matrix() const { return mat_; }

private:
    const VectorOperations* vec_ops_;
    const Matrix* mat_;
};

}  // namespace csr
