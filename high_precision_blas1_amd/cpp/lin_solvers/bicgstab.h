// bicgstab.h -- BiCGStab (van der Vorst) over the VectorOperations concept.
//
// Interface of numerical_algos::lin_solvers::bicgstab
// (source/numerical_algos/lin_solvers/bicgstab.h:64-253): template <LinearOperator,
// Preconditioner, VectorOperations, Monitor, Log>, constructor (vec_ops, log,
// obj_log_lev), set_use_precond_resid, set_resid_recalc_freq, solve(A, b, x) -> bool.
//
// Same recurrences as the reference (bicgstab.h:163-240): right-hand side and
// residual optionally preconditioned, r_ = r_0 the shadow residual,
//   rho_i = (r_i, r_), beta = (rho_i / rho_{i-1}) (alpha / omega_{i-1}),
//   p = r - beta omega nu + beta p, nu = P A p, alpha = rho_i / (nu, r_),
//   s = r - alpha nu, t = P A s, omega = (t, s) / (t, t),
//   x += alpha p + omega s, r = s - omega t (or b - A x every recalc_freq steps).
// Difference: a NaN or infinite coefficient stops the iteration and is reported
// by breakdown(); solve() then returns the monitor's decision (false unless the
// tolerance was met).  The reference returns true in that case (bicgstab.h:245-249).
// The reference's unused host buffer some_vec (bicgstab.h:110-113) is not kept.
#pragma once
#include <cmath>
#include <stdexcept>

#include "iter_solver_base.h"

namespace numerical_algos {
namespace lin_solvers {

template <class LinearOperator, class Preconditioner, class VectorOperations, class Monitor, class Log>
class bicgstab : public iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> {
public:
    typedef typename VectorOperations::scalar_type scalar_type;
    typedef typename VectorOperations::vector_type vector_type;
    typedef LinearOperator linear_operator_type;
    typedef Preconditioner preconditioner_type;
    typedef VectorOperations vector_operations_type;
    typedef Monitor monitor_type;
    typedef Log log_type;

private:
    typedef scalar_type T;
    typedef iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> parent_t;
    typedef hpb::logged_obj_base<Log> logged_obj_t;

    // r_i, shadow r_, p_i, s, t, nu_i, unpreconditioned r_i
    mutable vector_type ri_ = nullptr, rs_ = nullptr, pi_ = nullptr, s_ = nullptr, t_ = nullptr, nu_ = nullptr,
                        rnp_ = nullptr;
    bool use_precond_resid_ = true;
    int resid_recalc_freq_ = 0;
    mutable bool breakdown_ = false;

protected:
    using parent_t::monitor_;
    using parent_t::prec_;
    using parent_t::vec_ops_;

public:
    bicgstab(const vector_operations_type* vec_ops, Log* log = nullptr, int obj_log_lev = 3)
        : parent_t(vec_ops, log, obj_log_lev, "bicgstab::")
    {
        for (vector_type* v : bufs_()) {
            vec_ops_->init_vector(*v);
            vec_ops_->start_use_vector(*v);
        }
    }
    ~bicgstab() override
    {
        for (vector_type* v : bufs_()) {
            vec_ops_->stop_use_vector(*v);
            vec_ops_->free_vector(*v);
        }
    }
    void set_use_precond_resid(bool v) { use_precond_resid_ = v; }
    void set_resid_recalc_freq(int v) { resid_recalc_freq_ = v; }
    bool breakdown() const { return breakdown_; }

    bool solve(const linear_operator_type& A, const vector_type& b, vector_type& x) const override
    {
        if (prec_ != nullptr) prec_->set_operator(&A);
        breakdown_ = false;
        if (use_precond_resid_) {
            vec_ops_->assign(b, ri_);
            if (prec_ != nullptr) prec_->apply(ri_);
            monitor_.start(ri_);
        } else {
            monitor_.start(b);
        }
        const vector_type& checked = use_precond_resid_ ? ri_ : rnp_;
        // r_0 = P (b - A x0), shadow residual r_ = r_0
        A.apply(x, ri_);
        vec_ops_->add_mul(T(1), b, -T(1), ri_);
        if (!use_precond_resid_) vec_ops_->assign(ri_, rnp_);
        if (prec_ != nullptr) prec_->apply(ri_);
        vec_ops_->assign(ri_, rs_);
        T rho_prev = T(1), alpha = T(1), omega_prev = T(1);
        vec_ops_->assign_scalar(T(0), nu_);
        vec_ops_->assign_scalar(T(0), pi_);
        while (!monitor_.check_finished(x, checked)) {
            const T rho = vec_ops_->scalar_prod(ri_, rs_);
            const T beta = (rho / rho_prev) * (alpha / omega_prev);
            if (!finite_(beta)) return stop_("beta");
            // p = r - beta omega nu + beta p
            vec_ops_->add_mul(T(1), ri_, -beta * omega_prev, nu_, beta, pi_);
            A.apply(pi_, nu_);
            if (!use_precond_resid_) vec_ops_->assign(nu_, t_);   // t holds the unpreconditioned nu
            if (prec_ != nullptr) prec_->apply(nu_);
            alpha = rho / vec_ops_->scalar_prod(nu_, rs_);
            if (!finite_(alpha)) return stop_("alpha");
            vec_ops_->assign_mul(T(1), ri_, -alpha, nu_, s_);   // s = r - alpha nu
            if (!use_precond_resid_) vec_ops_->add_mul(-alpha, t_, T(1), rnp_);
            A.apply(s_, t_);
            if (!use_precond_resid_) vec_ops_->assign(t_, ri_);   // r_i holds the unpreconditioned t
            if (prec_ != nullptr) prec_->apply(t_);
            const T omega = vec_ops_->scalar_prod(t_, s_) / vec_ops_->scalar_prod(t_, t_);
            if (!finite_(omega)) return stop_("omega");
            vec_ops_->add_mul(alpha, pi_, omega, s_, T(1), x);   // x += alpha p + omega s
            if (resid_recalc_freq_ == 0 || (monitor_.iters_performed() + 1) % resid_recalc_freq_ != 0) {
                if (!use_precond_resid_) vec_ops_->add_mul(-omega, ri_, T(1), rnp_);
                vec_ops_->assign_mul(T(1), s_, -omega, t_, ri_);   // r = s - omega t
            } else {
                logged_obj_t::info_f("solve: iter = %d recalc real residual", monitor_.iters_performed());
                A.apply(x, ri_);
                vec_ops_->add_mul(T(1), b, -T(1), ri_);
                if (!use_precond_resid_) vec_ops_->assign(ri_, rnp_);
                if (prec_ != nullptr) prec_->apply(ri_);
            }
            omega_prev = omega;
            rho_prev = rho;
            ++monitor_;
        }
        const bool res = monitor_.converged();
        if (!res) logged_obj_t::error_f("solve: linear solver failed to converge");
        return res;
    }

private:
    struct buf_list {
        vector_type* p[7];
        vector_type** begin() { return p; }
        vector_type** end() { return p + 7; }
    };
    buf_list bufs_() const { return buf_list{{&ri_, &rs_, &pi_, &s_, &t_, &nu_, &rnp_}}; }
    static bool finite_(T v) { return std::isfinite((double)v); }
    bool stop_(const char* what) const
    {
        breakdown_ = true;
        logged_obj_t::info_f("solve: stop iterations because %s is not a finite number", what);
        return monitor_.converged();
    }
};

}  // namespace lin_solvers
}  // namespace numerical_algos
