// bicgstabl.h -- BiCGStab(L) (Sleijpen-Fokkema) over the VectorOperations concept.
//
// Interface of numerical_algos::lin_solvers::bicgstabl
// (source/numerical_algos/lin_solvers/bicgstabl.h:68-275): set_basis_size(L),
// set_use_precond_resid, set_resid_recalc_freq, solve(A, b, x) -> bool.
// Same recurrences as the reference (BiCG part :199-223, MR part by modified
// Gram-Schmidt on r_1..r_L :224-231, gamma / gamma' / gamma'' :232-245, updates
// :250-258, optional true-residual recomputation :247-263).  The MR coefficient
// table tau is (L+1) x (L+1): the reference declares tau[L][L] and indexes it with
// j <= L (bicgstabl.h:180,226,237,243).  A zero rho (r0 orthogonal to r~) stops
// the cycle and is reported by breakdown().
#pragma once
#include <cmath>
#include <stdexcept>
#include <vector>

#include "iter_solver_base.h"

namespace numerical_algos {
namespace lin_solvers {

template <class LinearOperator, class Preconditioner, class VectorOperations, class Monitor, class Log>
class bicgstabl : public iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> {
public:
    typedef typename VectorOperations::scalar_type scalar_type;
    typedef typename VectorOperations::vector_type vector_type;
    typedef LinearOperator linear_operator_type;
    typedef Preconditioner preconditioner_type;
    typedef VectorOperations vector_operations_type;
    typedef Monitor monitor_type;
    typedef Log log_type;

    static const int max_basis_sz_ = 30;  // SCFD_BICGSTABL_MAX_BASIS_SIZE (bicgstabl.h:49-51)

private:
    typedef scalar_type T;
    typedef iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> parent_t;
    typedef hpb::logged_obj_base<Log> logged_obj_t;

    mutable std::vector<vector_type> r_, u_;
    mutable vector_type rtilde_ = nullptr;
    bool use_precond_resid_ = true;
    int resid_recalc_freq_ = 0;
    int basis_sz_ = 2;
    mutable bool breakdown_ = false;

protected:
    using parent_t::monitor_;
    using parent_t::prec_;
    using parent_t::vec_ops_;

public:
    bicgstabl(const vector_operations_type* vec_ops, Log* log = nullptr, int obj_log_lev = 3)
        : parent_t(vec_ops, log, obj_log_lev, "bicgstabl::")
    {
        vec_ops_->init_vector(rtilde_);
        vec_ops_->start_use_vector(rtilde_);
        alloc_(basis_sz_);
    }
    ~bicgstabl() override
    {
        free_();
        vec_ops_->stop_use_vector(rtilde_);
        vec_ops_->free_vector(rtilde_);
    }
    void set_basis_size(int L)
    {
        if (L < 1 || L > max_basis_sz_) throw std::logic_error("bicgstabl::set_basis_size: L out of range");
        free_();
        basis_sz_ = L;
        alloc_(L);
    }
    int basis_size() const { return basis_sz_; }
    void set_use_precond_resid(bool v) { use_precond_resid_ = v; }
    void set_resid_recalc_freq(int v) { resid_recalc_freq_ = v; }
    bool breakdown() const { return breakdown_; }

    bool solve(const linear_operator_type& A, const vector_type& b, vector_type& x) const override
    {
        if (!use_precond_resid_ && prec_ != nullptr)
            throw std::logic_error("bicgstabl::solve: use_precond_resid_ == false with non-empty preconditioner is not supported");
        if (prec_ != nullptr) prec_->set_operator(&A);
        const int L = basis_sz_;
        breakdown_ = false;
        if (use_precond_resid_ && prec_ != nullptr) {
            vec_ops_->assign(b, r_[0]);
            prec_->apply(r_[0]);
            monitor_.start(r_[0]);
        } else {
            monitor_.start(b);
        }
        vec_ops_->assign_scalar(T(0), u_[0]);
        std::vector<T> tau((size_t)(L + 1) * (L + 1), T(0)), sigma(L + 1), gamma(L + 1), gamma_p(L + 1),
            gamma_pp(L + 1);
        auto TAU = [&](int i, int j) -> T& { return tau[(size_t)i * (L + 1) + j]; };

        calc_residual_(A, x, b, r_[0]);
        vec_ops_->assign(r_[0], rtilde_);
        const T rt_norm = vec_ops_->norm(rtilde_);
        if (rt_norm != T(0)) vec_ops_->scale(T(1) / rt_norm, rtilde_);

        T rho = T(1), alpha = T(0), omega = T(1);
        while (!monitor_.check_finished(x, r_[0])) {
            rho = -omega * rho;
            for (int j = 0; j < L; ++j) {
                if (rho == T(0)) {
                    breakdown_ = true;
                    break;
                }
                const T rho1 = vec_ops_->scalar_prod(r_[j], rtilde_);
                const T beta = alpha * rho1 / rho;
                rho = rho1;
                for (int i = 0; i <= j; ++i) vec_ops_->add_mul(T(1), r_[i], -beta, u_[i]);  // u_i = r_i - beta u_i
                A.apply(u_[j], u_[j + 1]);
                if (prec_ != nullptr) prec_->apply(u_[j + 1]);
                alpha = rho / vec_ops_->scalar_prod(u_[j + 1], rtilde_);
                for (int i = 0; i <= j; ++i) vec_ops_->add_mul(-alpha, u_[i + 1], T(1), r_[i]);  // r_i -= alpha u_{i+1}
                A.apply(r_[j], r_[j + 1]);
                if (prec_ != nullptr) prec_->apply(r_[j + 1]);
                vec_ops_->add_mul(alpha, u_[0], T(1), x);  // x += alpha u_0
            }
            if (breakdown_) {
                logged_obj_t::error_f("solve: breakdown, rho = 0");
                break;
            }
            // MR part: modified Gram-Schmidt on r_1..r_L
            for (int j = 1; j <= L; ++j) {
                for (int i = 1; i < j; ++i) {
                    TAU(i, j) = vec_ops_->scalar_prod(r_[j], r_[i]) / sigma[i];
                    vec_ops_->add_mul(-TAU(i, j), r_[i], T(1), r_[j]);
                }
                sigma[j] = vec_ops_->scalar_prod(r_[j], r_[j]);
                gamma_p[j] = vec_ops_->scalar_prod(r_[0], r_[j]) / sigma[j];
            }
            gamma[L] = gamma_p[L];
            omega = gamma[L];
            for (int j = L - 1; j >= 1; --j) {
                gamma[j] = gamma_p[j];
                for (int i = j + 1; i <= L; ++i) gamma[j] -= TAU(j, i) * gamma[i];
            }
            for (int j = 1; j < L; ++j) {
                gamma_pp[j] = gamma[j + 1];
                for (int i = j + 1; i < L; ++i) gamma_pp[j] += TAU(j, i) * gamma[i + 1];
            }
            const bool recalc = !(resid_recalc_freq_ == 0 ||
                                  ((monitor_.iters_performed() / L + 1) % resid_recalc_freq_ != 0));
            vec_ops_->add_mul(gamma[1], r_[0], T(1), x);
            vec_ops_->add_mul(-gamma_p[L], r_[L], T(1), r_[0]);
            vec_ops_->add_mul(-gamma[L], u_[L], T(1), u_[0]);
            for (int j = 1; j < L; ++j) {
                vec_ops_->add_mul(gamma_pp[j], r_[j], T(1), x);
                vec_ops_->add_mul(-gamma_p[j], r_[j], T(1), r_[0]);
                vec_ops_->add_mul(-gamma[j], u_[j], T(1), u_[0]);
            }
            if (recalc) {
                logged_obj_t::info_f("solve: iter = %d recalc real residual", monitor_.iters_performed());
                calc_residual_(A, x, b, r_[0]);
            }
            monitor_ += L;
        }
        const bool res = monitor_.converged() && !breakdown_;
        if (!res) logged_obj_t::error_f("solve: linear solver failed to converge");
        return res;
    }

private:
    void alloc_(int L)
    {
        r_.assign(L + 1, nullptr);
        u_.assign(L + 1, nullptr);
        for (int i = 0; i <= L; ++i) {
            vec_ops_->init_vector(r_[i]);
            vec_ops_->start_use_vector(r_[i]);
            vec_ops_->init_vector(u_[i]);
            vec_ops_->start_use_vector(u_[i]);
        }
    }
    void free_()
    {
        for (auto& v : r_) {
            vec_ops_->stop_use_vector(v);
            vec_ops_->free_vector(v);
        }
        for (auto& v : u_) {
            vec_ops_->stop_use_vector(v);
            vec_ops_->free_vector(v);
        }
        r_.clear();
        u_.clear();
    }
    void calc_residual_(const linear_operator_type& A, const vector_type& x, const vector_type& b,
                        vector_type& r) const
    {
        A.apply(x, r);
        vec_ops_->add_mul(T(1), b, -T(1), r);
        if (prec_ != nullptr) prec_->apply(r);
    }
};

}  // namespace lin_solvers
}  // namespace numerical_algos
