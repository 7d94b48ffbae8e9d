// gmres.h -- restarted GMRES(m) over the VectorOperations concept.
//
// Interface of numerical_algos::lin_solvers::gmres
// (source/numerical_algos/lin_solvers/gmres.h:65-495): template
// <LinearOperator, Preconditioner, VectorOperations, Monitor, Log, MaxBasis>,
// constructor (vec_ops, log, obj_log_lev), set_restarts (throws if >= MaxBasis),
// set_use_precond_resid, set_resid_recalc_freq, set_reorthogonalization,
// solve(A, b, x) -> bool (monitor_.converged()).
//
// Algorithm as in the reference: Arnoldi with Gram-Schmidt, Givens rotations on
// the host Hessenberg (row-major H[j*m + k]), back substitution, x += V s, true
// residual and monitor check at every restart.  Differences, all documented:
//  * orthogonalization: MGS (as the reference, gmres.h:426-452), CGS2 -- two
//    classical passes, each one fused multi-dot (reads w once) plus one fused
//    multi-update, 2 host synchronisations per step instead of i + 2 -- or CGS
//    with the second pass only when the first shrank w by more than 1/sqrt(2);
//  * early_exit: leave the cycle when the Givens estimate |s_{i+1}| is below the
//    tolerance (the reference runs all m steps; its early exit is commented out,
//    gmres.h:459-474); convergence is still decided on the true residual;
//  * the Krylov vector is formed in place in V[i+1] and the restart residual in
//    V[0] (the reference copies through w and r); the arithmetic is the same;
//  * set_reorthogonalization does one extra MGS correction per basis vector
//    (the reference's loop re-tests the old coefficient, gmres.h:435-441);
//  * with the monitor's out_min_resid_norm option the restart iterate of smallest
//    true residual is copied back into x before returning.
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "detail_fused.h"
#include "iter_solver_base.h"

namespace numerical_algos {
namespace lin_solvers {

// mgs: the reference's; cgs2: two classical passes; cgs_dgks: the second pass
// only when the first one shrank w by more than 1/sqrt(2) (Daniel, Gragg,
// Kaufman and Stewart 1976, "twice is enough").
enum class gram_schmidt { mgs, cgs2, cgs_dgks };

template <class LinearOperator, class Preconditioner, class VectorOperations, class Monitor, class Log,
          unsigned int GMRESMaxBasisSize = 100>
class gmres : public iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> {
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
    typedef vector_type T_vec;
    typedef iter_solver_base<LinearOperator, Preconditioner, VectorOperations, Monitor, Log> parent_t;
    typedef hpb::logged_obj_base<Log> logged_obj_t;

    mutable std::vector<T_vec> V_;   // restart_ + 1 basis vectors
    mutable T_vec r_ = nullptr;      // residual / scratch
    mutable std::vector<T> H_, s_, cs_, sn_, h1_, h2_;

    bool use_precond_resid_ = true;
    int resid_recalc_freq_ = 0;
    bool reorthogonalization_ = false;
    int restart_ = 2;
    gram_schmidt gs_ = gram_schmidt::mgs;
    bool early_exit_ = false;

protected:
    using parent_t::monitor_;
    using parent_t::prec_;
    using parent_t::vec_ops_;

public:
    gmres(const vector_operations_type* vec_ops, Log* log = nullptr, int obj_log_lev = 3)
        : parent_t(vec_ops, log, obj_log_lev, "gmres::")
    {
        alloc_(restart_);
    }
    ~gmres() override { free_(); }

    void set_restarts(unsigned int restart)
    {
        if (restart >= GMRESMaxBasisSize)
            throw std::logic_error("gmres::set_restarts: restart length must stay below the MaxBasis template "
                                   "argument (" + std::to_string(GMRESMaxBasisSize) + ")");
        if (restart < 1) throw std::logic_error("gmres::set_restarts: restart must be >= 1");
        free_();
        restart_ = (int)restart;
        alloc_(restart_);
    }
    void set_use_precond_resid(bool v) { use_precond_resid_ = v; }
    void set_resid_recalc_freq(int v) { resid_recalc_freq_ = v; }
    void set_reorthogonalization(bool v) { reorthogonalization_ = v; }
    void set_gram_schmidt(gram_schmidt g) { gs_ = g; }
    void set_early_exit(bool v) { early_exit_ = v; }
    int restarts() const { return restart_; }

    bool solve(const linear_operator_type& A, const T_vec& b, T_vec& x) const override
    {
        if (!use_precond_resid_ && prec_ != nullptr)
            throw std::logic_error("gmres::solve: a preconditioner needs the preconditioned residual test "
                                   "(set_use_precond_resid(true))");
        if (prec_ != nullptr) prec_->set_operator(&A);
        const int m = restart_;
        if (use_precond_resid_ && prec_ != nullptr) {
            vec_ops_->assign(b, r_);
            prec_->apply(r_);
            monitor_.start(r_);
        } else {
            monitor_.start(b);
        }
        calc_residual_(A, x, b, V_[0]);
        bool finished = monitor_.check_finished(x, V_[0]);
        while (!finished) {
            const T beta = monitor_.resid_norm();  // ||r|| just computed by the monitor
            vec_ops_->scale(T(1) / beta, V_[0]);
            zero_host_();
            s_[0] = beta;
            int i = -1;
            do {
                ++i;
                ++monitor_;
                calc_Krylov_vector_(A, V_[i], V_[i + 1]);
                orthogonalize_(i);
                const T h = vec_ops_->norm(V_[i + 1]);
                H_[(i + 1) * m + i] = h;
                if (h != T(0)) vec_ops_->scale(T(1) / h, V_[i + 1]);
                plane_rotation_(i);
                if (h == T(0)) break;  // invariant subspace: the solution is in span V[0..i]
                if (early_exit_ && std::abs(s_[i + 1]) <= monitor_.tol()) break;
            } while (i + 1 < m && monitor_.iters_performed() < monitor_.max_iters_num());
            solve_triangular_(i);
            detail::multi_axpy(*vec_ops_, i + 1, s_.data(), V_.data(), x);
            calc_residual_(A, x, b, V_[0]);
            finished = monitor_.check_finished(x, V_[0]);
        }
        if (monitor_.out_min_resid_norm()) vec_ops_->assign(monitor_.min_resid_norm_x(), x);
        const bool ok = monitor_.converged();
        if (!ok) logged_obj_t::error_f("solve: tolerance not reached after %d iterations", monitor_.iters_performed());
        return ok;
    }

private:
    void alloc_(int m)
    {
        V_.assign(m + 1, nullptr);
        for (auto& v : V_) {
            vec_ops_->init_vector(v);
            vec_ops_->start_use_vector(v);
        }
        vec_ops_->init_vector(r_);
        vec_ops_->start_use_vector(r_);
        H_.assign((size_t)m * (m + 1), T(0));
        s_.assign(m + 1, T(0));
        cs_.assign(m, T(0));
        sn_.assign(m, T(0));
        h1_.assign(m + 1, T(0));
        h2_.assign(m + 1, T(0));
    }
    void free_()
    {
        for (auto& v : V_) {
            vec_ops_->stop_use_vector(v);
            vec_ops_->free_vector(v);
        }
        V_.clear();
        if (r_) {
            vec_ops_->stop_use_vector(r_);
            vec_ops_->free_vector(r_);
        }
    }
    void zero_host_() const
    {
        std::fill(H_.begin(), H_.end(), T(0));
        std::fill(s_.begin(), s_.end(), T(0));
        std::fill(cs_.begin(), cs_.end(), T(0));
        std::fill(sn_.begin(), sn_.end(), T(0));
    }
    // r := b - A x (then P r), gmres.h:122-130
    void calc_residual_(const linear_operator_type& A, const T_vec& x, const T_vec& b, T_vec& r) const
    {
        A.apply(x, r);
        vec_ops_->add_mul(T(1), b, -T(1), r);
        if (prec_ != nullptr) prec_->apply(r);
    }
    void calc_Krylov_vector_(const linear_operator_type& A, const T_vec& v, T_vec& w) const
    {
        A.apply(v, w);
        if (prec_ != nullptr) prec_->apply(w);
    }
    // H(0:i, i) and w = V[i+1] orthogonal to V[0..i]
    void orthogonalize_(int i) const
    {
        const int m = restart_;
        T_vec& w = V_[i + 1];
        if (gs_ == gram_schmidt::mgs) {
            for (int k = 0; k <= i; ++k) {
                T alpha = vec_ops_->scalar_prod(V_[k], w);
                vec_ops_->add_mul(-alpha, V_[k], w);
                if (reorthogonalization_) {
                    T c = vec_ops_->scalar_prod(V_[k], w);
                    vec_ops_->add_mul(-c, V_[k], w);
                    alpha += c;
                }
                H_[k * m + i] = alpha;
            }
        } else {
            const bool selective = gs_ == gram_schmidt::cgs_dgks;
            const T wn = selective ? vec_ops_->norm(w) : T(0);
            detail::multi_dot(*vec_ops_, i + 1, V_.data(), w, h1_.data());
            for (int k = 0; k <= i; ++k) h2_[k] = -h1_[k];
            detail::multi_axpy(*vec_ops_, i + 1, h2_.data(), V_.data(), w);
            if (selective) {
                const T wn1 = vec_ops_->norm(w);
                if (wn1 * wn1 >= T(0.5) * wn * wn) {
                    for (int k = 0; k <= i; ++k) H_[k * m + i] = h1_[k];
                    return;
                }
            }
            detail::multi_dot(*vec_ops_, i + 1, V_.data(), w, h2_.data());
            for (int k = 0; k <= i; ++k) {
                H_[k * m + i] = h1_[k] + h2_[k];
                h2_[k] = -h2_[k];
            }
            detail::multi_axpy(*vec_ops_, i + 1, h2_.data(), V_.data(), w);
        }
    }
    // Givens rotations (gmres.h:224-274)
    static void apply_rot_(T& dx, T& dy, T cs, T sn)
    {
        const T t = cs * dx + sn * dy;
        dy = -sn * dx + cs * dy;
        dx = t;
    }
    static void gen_rot_(T dx, T dy, T& cs, T& sn)
    {
        if (dy == T(0)) {
            cs = T(1);
            sn = T(0);
        } else if (std::abs(dy) > std::abs(dx)) {
            const T t = dx / dy;
            sn = T(1) / std::sqrt(T(1) + t * t);
            cs = t * sn;
        } else {
            const T t = dy / dx;
            cs = T(1) / std::sqrt(T(1) + t * t);
            sn = t * cs;
        }
    }
    void plane_rotation_(int i) const
    {
        const int m = restart_;
        for (int k = 0; k < i; ++k) apply_rot_(H_[k * m + i], H_[(k + 1) * m + i], cs_[k], sn_[k]);
        gen_rot_(H_[i * m + i], H_[(i + 1) * m + i], cs_[i], sn_[i]);
        apply_rot_(H_[i * m + i], H_[(i + 1) * m + i], cs_[i], sn_[i]);
        H_[(i + 1) * m + i] = T(0);
        apply_rot_(s_[i], s_[i + 1], cs_[i], sn_[i]);
    }
    // back substitution, gmres.h:204-215
    void solve_triangular_(int ind) const
    {
        const int m = restart_;
        for (int j = ind; j >= 0; --j) {
            s_[j] /= H_[j * m + j];
            for (int k = j - 1; k >= 0; --k) s_[k] -= H_[k * m + j] * s_[j];
        }
    }
};

}  // namespace lin_solvers
}  // namespace numerical_algos
