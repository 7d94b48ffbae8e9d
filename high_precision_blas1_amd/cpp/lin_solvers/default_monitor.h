// default_monitor.h -- convergence monitor of the Krylov solvers.
//
// Same interface and semantics as numerical_algos::lin_solvers::default_monitor
// (source/numerical_algos/lin_solvers/default_monitor.h:53-227):
//   tol = abs_tol + rel_tol * ||rhs||                            (:156-159)
//   check_finished(x, r): valid(x) && ||r|| <= tol && it >= min, or it >= max (:202-226)
//   optional convergence history of (iteration, ||r|| [/ ||rhs||])          (:215-216)
// The norms go through the VectorOperations given at construction, so they are
// compensated when that object is in high precision.
#pragma once
#include <utility>
#include <vector>

#include "../utils/hpb_log.h"

namespace numerical_algos {
namespace lin_solvers {

template <class VectorOperations, class Log>
class default_monitor : public hpb::logged_obj_base<Log> {
public:
    typedef typename VectorOperations::scalar_type scalar_type;
    typedef typename VectorOperations::vector_type vector_type;
    typedef VectorOperations vector_operations_type;
    typedef scalar_type T_real;

private:
    typedef hpb::logged_obj_base<Log> logged_obj_t;
    const vector_operations_type& vec_ops_;
    T_real rel_tol_ = T_real(0), abs_tol_ = T_real(0), rel_tol_save_ = T_real(0);
    int max_iters_num_ = 100, min_iters_num_ = 0, max_iters_num_save_ = 100;
    bool out_min_resid_norm_ = false;
    bool save_convergence_history_ = false;
    bool divide_out_norms_by_rel_base_ = false;
    int iters_performed_ = 0;
    bool is_valid_number_ = true;
    T_real rhs_norm_ = T_real(0);
    T_real resid_norm_ = T_real(0);
    T_real min_resid_norm_ = T_real(0);
    vector_type min_resid_norm_x_ = nullptr;
    std::vector<std::pair<int, T_real> > convergence_history_;

public:
    default_monitor(const vector_operations_type& vec_ops, Log* log = nullptr, int obj_log_lev = 0)
        : logged_obj_t(log, obj_log_lev, "default_monitor: "), vec_ops_(vec_ops)
    {
    }
    ~default_monitor()
    {
        if (min_resid_norm_x_) vec_ops_.free_vector(min_resid_norm_x_);
    }

    void init(T_real rel_tol, T_real abs_tol = T_real(0), int max_iters_num = 100, int min_iters_num = 0,
              bool out_min_resid_norm = false, bool save_convergence_history = false,
              bool divide_out_norms_by_rel_base = false)
    {
        rel_tol_ = rel_tol;
        abs_tol_ = abs_tol;
        max_iters_num_ = max_iters_num;
        min_iters_num_ = min_iters_num;
        max_iters_num_save_ = max_iters_num;
        out_min_resid_norm_ = out_min_resid_norm;
        if (out_min_resid_norm_ && !min_resid_norm_x_) {
            vec_ops_.init_vector(min_resid_norm_x_);
            vec_ops_.start_use_vector(min_resid_norm_x_);
        }
        save_convergence_history_ = save_convergence_history;
        divide_out_norms_by_rel_base_ = divide_out_norms_by_rel_base;
        rel_tol_save_ = rel_tol_;
    }
    void set_save_convergence_history(bool v) { save_convergence_history_ = v; }
    void set_divide_out_norms_by_rel_base(bool v) { divide_out_norms_by_rel_base_ = v; }
    void set_temp_tolerance(T_real rel_tol)
    {
        rel_tol_save_ = rel_tol_;
        rel_tol_ = rel_tol;
    }
    void restore_tolerance() { rel_tol_ = rel_tol_save_; }
    void set_temp_max_iterations(int m)
    {
        max_iters_num_save_ = max_iters_num_;
        max_iters_num_ = m;
    }
    void restore_max_iterations() { max_iters_num_ = max_iters_num_save_; }

    void start(const vector_type& rhs)
    {
        rhs_norm_ = vec_ops_.norm(rhs);
        iters_performed_ = 0;
        if (save_convergence_history_) convergence_history_.clear();
    }
    // start with a known right-hand-side norm (avoids a reduction)
    void start_with_norm(T_real rhs_norm)
    {
        rhs_norm_ = rhs_norm;
        iters_performed_ = 0;
        if (save_convergence_history_) convergence_history_.clear();
    }
    void stop() {}

    T_real rel_tol() const { return rel_tol_; }
    T_real abs_tol() const { return abs_tol_; }
    T_real rel_tol_base() const { return rhs_norm(); }
    T_real tol() const { return abs_tol() + rel_tol() * rel_tol_base(); }
    int max_iters_num() const { return max_iters_num_; }
    int min_iters_num() const { return min_iters_num_; }
    bool out_min_resid_norm() const { return out_min_resid_norm_; }
    bool save_convergence_history() const { return save_convergence_history_; }
    bool divide_out_norms_by_rel_base() const { return divide_out_norms_by_rel_base_; }
    int iters_performed() const { return iters_performed_; }
    bool is_valid_number() const { return is_valid_number_; }
    T_real rhs_norm() const { return rhs_norm_; }
    T_real resid_norm() const { return resid_norm_; }
    T_real resid_norm_out() const { return divide_out_norms_by_rel_base() ? resid_norm() / rel_tol_base() : resid_norm(); }
    T_real tol_out() const { return divide_out_norms_by_rel_base() ? tol() / rel_tol_base() : tol(); }
    const vector_type& min_resid_norm_x() const { return min_resid_norm_x_; }
    // smallest residual norm checked since start() (meaningful with out_min_resid_norm)
    T_real min_resid_norm() const { return min_resid_norm_; }
    const std::vector<std::pair<int, T_real> >& convergence_history() const { return convergence_history_; }

    default_monitor& operator++()
    {
        ++iters_performed_;
        return *this;
    }
    default_monitor& operator+=(int n)
    {
        iters_performed_ += n;
        return *this;
    }
    bool converged() const { return is_valid_number() && (resid_norm() <= tol()); }

    bool check_finished(const vector_type& x, const vector_type& r)
    {
        logged_obj_t::info_f("iter = %d", iters_performed());
        is_valid_number_ = vec_ops_.check_is_valid_number(x);
        if (!is_valid_number_) {
            logged_obj_t::info_f("solution is not valid number");
            return true;
        }
        return record_norm(vec_ops_.norm(r), &x);
    }
    // Same decision from a residual norm computed elsewhere (x assumed valid when
    // the norm is finite: a non-finite x makes every residual norm non-finite).
    bool check_finished_norm(T_real rnorm, const vector_type* x = nullptr)
    {
        is_valid_number_ = rnorm == rnorm && rnorm - rnorm == T_real(0);
        if (!is_valid_number_) return true;
        return record_norm(rnorm, x);
    }

private:
    bool record_norm(T_real rnorm, const vector_type* x)
    {
        resid_norm_ = rnorm;
        logged_obj_t::info_f("resid norm = %0.6e tol = %0.6e", (double)resid_norm_out(), (double)tol_out());
        if (save_convergence_history_) convergence_history_.push_back(std::make_pair(iters_performed(), resid_norm_out()));
        if (out_min_resid_norm_ && x) {
            if (iters_performed() == 0 || resid_norm() < min_resid_norm_) {
                min_resid_norm_ = resid_norm();
                vec_ops_.assign(*x, min_resid_norm_x_);
            }
        }
        return (converged() && iters_performed() >= min_iters_num()) || iters_performed() >= max_iters_num();
    }
};

}  // namespace lin_solvers
}  // namespace numerical_algos
