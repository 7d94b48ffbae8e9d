// bicgstabl.h -- BiCGStab(L) over the VectorOperations concept.
//
// Algorithm: G. L. G. Sleijpen and D. R. Fokkema, "BiCGstab(l) for linear equations
// involving unsymmetric matrices with complex spectrum", Electronic Transactions on
// Numerical Analysis 1 (1993), pp. 11-32, Algorithm 3.1 (BiCG part, then the minimal
// residual part by modified Gram-Schmidt on r_1..r_L with the gamma, gamma' and
// gamma'' recurrences; the final update in the paper's order u, x, r).
//
// Drop-in for numerical_algos::lin_solvers::bicgstabl
// (source/numerical_algos/lin_solvers/bicgstabl.h:68-149): set_basis_size(L),
// set_use_precond_resid, set_resid_recalc_freq, monitor(), solve(A, b, x) -> bool.
// The preconditioner is applied on the left (r and the A u_j, A r_j are
// preconditioned), so the monitor tests the preconditioned residual, and a
// preconditioner requires use_precond_resid.  The shadow vector is r_0 / ||r_0||.
// The MR coefficients tau_ij are kept in an (L+1) x (L+1) table indexed 1..L.
// resid_recalc_freq = k > 0 recomputes r_0 = P (b - A x) after every k-th cycle.
// A vanishing denominator (rho or (A u_j, r~) in the BiCG part, (r_j, r_j) in the MR
// part) ends the solve before x changes; breakdown() reports it unless the iterate's
// true residual then meets the tolerance (a lucky breakdown: the residual vanished,
// as on the identity), which counts as convergence.  With
// the monitor's out_min_resid_norm option the iterate of smallest checked residual
// is copied back into x before returning.
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
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

    static const int max_basis_sz_ = 30;  // same cap as the reference (bicgstabl.h:49-51)

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
        if (L < 1 || L > max_basis_sz_) throw std::logic_error("bicgstabl::set_basis_size: L must lie in [1, " + std::to_string(max_basis_sz_) + "]");
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
            throw std::logic_error("bicgstabl::solve: a preconditioner needs the preconditioned residual test "
                                   "(set_use_precond_resid(true))");
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
        auto vanished = [](T d) { return d == T(0) || !(d == d); };
        while (!monitor_.check_finished(x, r_[0])) {
            rho = -omega * rho;
            int steps = 0;   // BiCG steps of this cycle that updated x
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
                const T sg = vec_ops_->scalar_prod(u_[j + 1], rtilde_);
                if (vanished(sg)) {   // stop before x changes
                    breakdown_ = true;
                    break;
                }
                alpha = rho / sg;
                for (int i = 0; i <= j; ++i) vec_ops_->add_mul(-alpha, u_[i + 1], T(1), r_[i]);  // r_i -= alpha u_{i+1}
                A.apply(r_[j], r_[j + 1]);
                if (prec_ != nullptr) prec_->apply(r_[j + 1]);
                vec_ops_->add_mul(alpha, u_[0], T(1), x);  // x += alpha u_0
                steps = j + 1;
            }
            if (breakdown_) {
                logged_obj_t::info_f("solve: the BiCG part stopped at iteration %d: (r_j, r~) or (A u_j, r~) vanished",
                                     monitor_.iters_performed() + steps);
                monitor_ += steps;
                break;
            }
            // MR part: modified Gram-Schmidt on r_1..r_L; a vanishing sigma_j (r_j = 0)
            // ends the solve before the MR update touches x
            for (int j = 1; j <= L && !breakdown_; ++j) {
                for (int i = 1; i < j; ++i) {
                    TAU(i, j) = vec_ops_->scalar_prod(r_[j], r_[i]) / sigma[i];
                    vec_ops_->add_mul(-TAU(i, j), r_[i], T(1), r_[j]);
                }
                sigma[j] = vec_ops_->scalar_prod(r_[j], r_[j]);
                if (vanished(sigma[j])) {
                    breakdown_ = true;
                    break;
                }
                gamma_p[j] = vec_ops_->scalar_prod(r_[0], r_[j]) / sigma[j];
            }
            if (breakdown_) {
                logged_obj_t::info_f("solve: the MR part stopped at iteration %d: (r_j, r_j) vanished",
                                     monitor_.iters_performed() + L);
                monitor_ += L;
                break;
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
            const int cycle = monitor_.iters_performed() / L + 1;   // 1-based index of this cycle
            const bool recalc = resid_recalc_freq_ > 0 && cycle % resid_recalc_freq_ == 0;
            // paper's order: u_0 -= gamma_L u_L, x += gamma_1 r_0, r_0 -= gamma'_L r_L,
            // then for j = 1..L-1 the u, x and r terms of column j
            vec_ops_->add_mul(-gamma[L], u_[L], T(1), u_[0]);
            vec_ops_->add_mul(gamma[1], r_[0], T(1), x);
            vec_ops_->add_mul(-gamma_p[L], r_[L], T(1), r_[0]);
            for (int j = 1; j < L; ++j) {
                vec_ops_->add_mul(-gamma[j], u_[j], T(1), u_[0]);
                vec_ops_->add_mul(gamma_pp[j], r_[j], T(1), x);
                vec_ops_->add_mul(-gamma_p[j], r_[j], T(1), r_[0]);
            }
            if (recalc) {
                logged_obj_t::info_f("solve: cycle ending at iteration %d, residual recomputed from x",
                                     monitor_.iters_performed() + L);
                calc_residual_(A, x, b, r_[0]);
            }
            monitor_ += L;
        }
        if (breakdown_) {
            // a vanishing denominator ended the iteration; when the iterate already solves
            // the system (its residual vanished: a lucky breakdown) that is convergence
            calc_residual_(A, x, b, r_[0]);
            if (use_precond_resid_ && prec_ != nullptr) prec_->apply(r_[0]);
            monitor_.check_finished(x, r_[0]);
            if (monitor_.converged()) breakdown_ = false;
            else logged_obj_t::error_f("solve: breakdown at iteration %d", monitor_.iters_performed());
        }
        if (monitor_.out_min_resid_norm()) vec_ops_->assign(monitor_.min_resid_norm_x(), x);
        const bool ok = monitor_.converged() && !breakdown_;
        if (!ok) logged_obj_t::error_f("solve: tolerance not reached after %d iterations", monitor_.iters_performed());
        return ok;
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
