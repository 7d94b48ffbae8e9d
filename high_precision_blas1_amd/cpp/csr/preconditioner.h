// preconditioner.h -- incomplete-factorisation preconditioners for the solvers.
//
// csr::preconditioner: interface and action of the reference's
// csr::preconditioner (source/common/csr/preconditioner.h:30-90): set_operator,
// set_matrix(L, U) with two CSR matrices, apply(x) := U^-1 L^-1 x through
// triangular_solve_lower (unit diagonal) and triangular_solve_upper.
//
// csr::ilu0_preconditioner: the same action with L and U taken from an ILU(0)
// factorisation of the operator's matrix, computed on the device in set_matrix.
// The reference's gpu_ilu0_preconditioner (common/csr/gpu_ilu0_preconditioner.h:
// 30-82) keeps a copy of the matrix and its apply only copies x; this one factors.
#pragma once
#include <stdexcept>

#include "hip_matrix.h"

namespace csr {

template <class VectorOperations, class LinearOperator, class Matrix>
struct preconditioner {
    using T = typename VectorOperations::scalar_type;
    using T_vec = typename VectorOperations::vector_type;

    explicit preconditioner(const VectorOperations* vec_ops) : vec_ops_(vec_ops)
    {
        vec_ops_->init_vector(y_);
        vec_ops_->start_use_vector(y_);
    }
    ~preconditioner()
    {
        vec_ops_->stop_use_vector(y_);
        vec_ops_->free_vector(y_);
    }
    preconditioner(const preconditioner&) = delete;
    preconditioner& operator=(const preconditioner&) = delete;

    void set_operator(const LinearOperator* op) { op_ = op; }
    void set_matrix(const Matrix* L, const Matrix* U)
    {
        L_ = L;
        U_ = U;
    }
    // x := U^-1 L^-1 x; without matrices the identity (as in the reference)
    void apply(T_vec& x) const
    {
        if (!L_ || !U_) return;
        L_->triangular_solve_lower(x, y_);
        U_->triangular_solve_upper(y_, x);
    }

private:
    const VectorOperations* vec_ops_;
    const LinearOperator* op_ = nullptr;
    const Matrix* L_ = nullptr;
    const Matrix* U_ = nullptr;
    mutable T_vec y_ = nullptr;
};

template <class VectorOperations, class LinearOperator, class Matrix>
struct ilu0_preconditioner {
    using T = typename VectorOperations::scalar_type;
    using T_vec = typename VectorOperations::vector_type;

    explicit ilu0_preconditioner(const VectorOperations* vec_ops) : vec_ops_(vec_ops)
    {
        vec_ops_->init_vector(y_);
        vec_ops_->start_use_vector(y_);
    }
    ~ilu0_preconditioner()
    {
        vec_ops_->stop_use_vector(y_);
        vec_ops_->free_vector(y_);
        if (lu_) {
            hpb_free(lu_vals_);
            delete lu_;
        }
    }
    ilu0_preconditioner(const ilu0_preconditioner&) = delete;
    ilu0_preconditioner& operator=(const ilu0_preconditioner&) = delete;

    void set_operator(const LinearOperator* op) { op_ = op; }
    // factor a copy of A's values; the pattern arrays of A are shared and must outlive this object
    void set_matrix(const Matrix& A)
    {
        if (lu_) {
            hpb_free(lu_vals_);
            delete lu_;
            lu_ = nullptr;
        }
        hpb::check(hpb_malloc((void**)&lu_vals_, sizeof(T) * (size_t)A.nnz()), "ilu0_preconditioner::set_matrix");
        hpb::check(hpb_memcpy_d2d(A.handle(), lu_vals_, A.data(), sizeof(T) * (size_t)A.nnz()),
                   "ilu0_preconditioner::set_matrix");
        lu_ = new Matrix(A.handle());
        lu_->adopt_device(A.rows(), A.get_dim().columns, A.nnz(), const_cast<int*>(A.row_ptr()),
                          const_cast<int*>(A.col_ind()), lu_vals_, false);
        lu_->use_compensated_rows(A.compensated_rows());
        lu_->ilu0_in_place();
    }
    void apply(T_vec& x) const
    {
        if (!lu_) return;
        lu_->triangular_solve_lower(x, y_);
        lu_->triangular_solve_upper(y_, x);
    }
    const Matrix* factors() const { return lu_; }

private:
    const VectorOperations* vec_ops_;
    const LinearOperator* op_ = nullptr;
    Matrix* lu_ = nullptr;
    T* lu_vals_ = nullptr;
    mutable T_vec y_ = nullptr;
};

}  // namespace csr
