// bicgstab.h -- BiCGStab over the VectorOperations concept.
//
// Algorithm: H. A. van der Vorst, "Bi-CGSTAB: a fast and smoothly converging variant
// of Bi-CG for the solution of nonsymmetric linear systems", SIAM J. Sci. Stat.
// Comput. 13(2), 1992, pp. 631-644, with the preconditioner applied on the right
// (Barrett et al., "Templates for the Solution of Linear Systems", SIAM 1994, the
// preconditioned BiCGSTAB of section 2.3.8).
//
// Drop-in for numerical_algos::lin_solvers::bicgstab
// (source/numerical_algos/lin_solvers/bicgstab.h:64-131): the same template
// parameters <LinearOperator, Preconditioner, VectorOperations, Monitor, Log>, the
// constructor (vec_ops, log, obj_log_lev), set_preconditioner, set_use_precond_resid,
// set_resid_recalc_freq, monitor() and solve(A, b, x) -> bool.
//
// Design of this implementation:
//   * Right preconditioning.  The iteration runs on A M y = b with x = M y, where M is
//     the preconditioner's apply().  The recursive vector r is then the residual
//     b - A x of the original system, so no second residual has to be carried.  With
//     use_precond_resid the monitor tests ||M r|| against ||M b|| instead; M r is formed
//     for each check into a work vector (one more preconditioner application per
//     iteration, and none without a preconditioner).
//   * In-place half step: s = r - alpha v overwrites r, and r = s - omega t overwrites
//     it again, so the iteration holds five vectors (r, r~, p, v, t) and three more
//     with a preconditioner (M p, M s and the checked M r).
//   * Lost bi-orthogonality: when (r~, r) or (r~, v) vanishes, or (t, s) does so that
//     the stabilising step makes no progress, the shadow vector is reset to the current
//     residual and the direction to r, and a new Bi-CG sequence begins (at most
//     max_shadow_resets times per solve).  A non-finite coefficient ends the solve;
//     breakdown() reports it and solve() returns whether the monitor's tolerance was
//     met.
//   * resid_recalc_freq = k > 0: every k-th iteration r is recomputed as b - A x.
//   * With the monitor's out_min_resid_norm option the iterate of smallest checked
//     residual is copied back into x before returning.
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

    static const int max_shadow_resets = 8;

private:
    typedef scalar_type T;
    typedef iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> parent_t;
    typedef hpb::logged_obj_base<Log> logged_obj_t;

    enum { R, RT, P, V, TT, MP, MS, MR, NVEC };
    mutable vector_type v_[NVEC] = {};
    mutable bool have_prec_bufs_ = false;
    bool use_precond_resid_ = true;
    int resid_recalc_freq_ = 0;
    mutable bool breakdown_ = false;
    mutable int shadow_resets_ = 0;

protected:
    using parent_t::monitor_;
    using parent_t::prec_;
    using parent_t::vec_ops_;

public:
    bicgstab(const vector_operations_type* vec_ops, Log* log = nullptr, int obj_log_lev = 3)
        : parent_t(vec_ops, log, obj_log_lev, "bicgstab::")
    {
        for (int i = R; i <= TT; ++i) take_(v_[i]);
    }
    ~bicgstab() override
    {
        for (int i = 0; i < NVEC; ++i)
            if (v_[i] != vector_type()) {
                vec_ops_->stop_use_vector(v_[i]);
                vec_ops_->free_vector(v_[i]);
            }
    }
    void set_use_precond_resid(bool v) { use_precond_resid_ = v; }
    void set_resid_recalc_freq(int v) { resid_recalc_freq_ = v; }
    bool breakdown() const { return breakdown_; }
    int shadow_resets() const { return shadow_resets_; }

    bool solve(const linear_operator_type& A, const vector_type& b, vector_type& x) const override
    {
        const bool pc = prec_ != nullptr;
        if (pc) {
            prec_->set_operator(&A);
            if (!have_prec_bufs_) {
                for (int i = MP; i < NVEC; ++i) take_(v_[i]);
                have_prec_bufs_ = true;
            }
        }
        vector_type& r = v_[R];
        vector_type& rt = v_[RT];
        vector_type& p = v_[P];
        vector_type& v = v_[V];
        vector_type& t = v_[TT];
        // the vectors the operator is applied to: M p and M s, or p and s themselves
        vector_type& mp = pc ? v_[MP] : p;
        vector_type& ms = pc ? v_[MS] : r;
        const bool test_mr = pc && use_precond_resid_;
        breakdown_ = false;
        shadow_resets_ = 0;

        if (test_mr) {
            precondition_(b, v_[MR]);
            monitor_.start(v_[MR]);
        } else {
            monitor_.start(b);
        }
        residual_(A, b, x, r);
        restart_shadow_(r, rt, p);
        T rho = vec_ops_->scalar_prod(rt, r);
        // first: p already holds the direction of this step; renew: begin a new Bi-CG
        // sequence from the current residual before the next step
        bool first = true, renew = false;
        T alpha = T(0), omega = T(0);

        while (!monitor_.check_finished(x, checked_(r, test_mr))) {
            if (renew) {
                if (!reset_shadow_(r, rt, p, rho)) break;
                renew = false;
            } else if (!first) {
                const T rho_new = vec_ops_->scalar_prod(rt, r);
                if (rho_new == T(0)) {
                    if (!reset_shadow_(r, rt, p, rho)) break;
                } else {
                    const T beta = (rho_new / rho) * (alpha / omega);
                    if (!finite_(beta)) {
                        stop_("beta");
                        break;
                    }
                    rho = rho_new;
                    vec_ops_->add_mul(T(1), r, -beta * omega, v, beta, p);   // p = r + beta (p - omega v)
                }
            }
            first = false;
            if (pc) precondition_(p, mp);
            A.apply(mp, v);
            const T sigma = vec_ops_->scalar_prod(rt, v);
            if (sigma == T(0)) {
                // (r~, A M p) = 0: no Bi-CG step exists along p
                renew = true;
                continue;
            }
            alpha = rho / sigma;
            if (!finite_(alpha)) {
                stop_("alpha");
                break;
            }
            vec_ops_->add_mul(-alpha, v, T(1), r);   // r <- s = r - alpha v
            if (pc) precondition_(r, ms);
            A.apply(ms, t);
            const T tt = vec_ops_->scalar_prod(t, t);
            const T ts = vec_ops_->scalar_prod(t, r);
            omega = tt != T(0) ? ts / tt : T(0);
            if (!finite_(omega)) {
                stop_("omega");
                break;
            }
            vec_ops_->add_mul(alpha, mp, omega, ms, T(1), x);   // x += alpha M p + omega M s
            ++monitor_;
            const int it = monitor_.iters_performed();
            if (resid_recalc_freq_ > 0 && it % resid_recalc_freq_ == 0) {
                logged_obj_t::info_f("solve: iteration %d, residual recomputed from x", it);
                residual_(A, b, x, r);
            } else {
                vec_ops_->add_mul(-omega, t, T(1), r);   // r = s - omega t
            }
            // omega = 0: the stabilising step made no progress (t orthogonal to s)
            if (omega == T(0)) renew = true;
        }
        if (monitor_.out_min_resid_norm()) vec_ops_->assign(monitor_.min_resid_norm_x(), x);
        const bool ok = monitor_.converged();
        if (!ok) logged_obj_t::error_f("solve: tolerance not reached after %d iterations", monitor_.iters_performed());
        return ok;
    }

private:
    void take_(vector_type& v) const
    {
        vec_ops_->init_vector(v);
        vec_ops_->start_use_vector(v);
    }
    static bool finite_(T v) { return std::isfinite((double)v); }
    void residual_(const linear_operator_type& A, const vector_type& b, const vector_type& x, vector_type& r) const
    {
        A.apply(x, r);
        vec_ops_->add_mul(T(1), b, -T(1), r);   // r = b - A x
    }
    void precondition_(const vector_type& src, vector_type& dst) const
    {
        vec_ops_->assign(src, dst);
        prec_->apply(dst);
    }
    const vector_type& checked_(const vector_type& r, bool test_mr) const
    {
        if (!test_mr) return r;
        precondition_(r, v_[MR]);
        return v_[MR];
    }
    // r~ = r / ||r|| (a scaled copy keeps (r~, .) in range), p = r
    void restart_shadow_(const vector_type& r, vector_type& rt, vector_type& p) const
    {
        vec_ops_->assign(r, rt);
        const T nr = vec_ops_->norm(rt);
        if (nr != T(0) && finite_(nr)) vec_ops_->scale(T(1) / nr, rt);
        vec_ops_->assign(r, p);
    }
    bool reset_shadow_(const vector_type& r, vector_type& rt, vector_type& p, T& rho) const
    {
        if (shadow_resets_ >= max_shadow_resets) {
            stop_("(r~, r) or (r~, v)");
            return false;
        }
        ++shadow_resets_;
        logged_obj_t::info_f("solve: shadow residual reset %d at iteration %d", shadow_resets_,
                             monitor_.iters_performed());
        restart_shadow_(r, rt, p);
        rho = vec_ops_->scalar_prod(rt, r);
        if (!finite_(rho) || rho == T(0)) {
            stop_("(r~, r) after a reset");
            return false;
        }
        return true;
    }
    void stop_(const char* what) const
    {
        breakdown_ = true;
        logged_obj_t::info_f("solve: breakdown at iteration %d: %s is zero or not finite", monitor_.iters_performed(),
                             what);
    }
};

}  // namespace lin_solvers
}  // namespace numerical_algos
