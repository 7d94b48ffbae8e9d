// hpb_solvers.cpp -- C ABI of the Krylov solvers on CSR matrices.
//
// Instantiates the solver templates (cpp/lin_solvers: gmres, bicgstabl,
// default_monitor) over hip_vector_operations and csr::hip_matrix, the same
// composition the reference's CSR driver builds from gpu_vector_operations and
// csr::gpu_matrix (source/test_csr_linear_solver.cpp:95-359), for
//   hpb_csr_solve_f64    FP64 vectors, plain or compensated dots / rows;
//   hpb_csr_solve_ir_f64 FP32 vectors and matrix values with compensated FP32
//                        dots in the inner GMRES / BiCGStab(L), wrapped in FP64
//                        iterative refinement with a compensated FP64 residual.
// Every solve reports the TRUE relative residual ||b - A x|| / ||b|| recomputed
// with compensated FP64 rows and norms (test_csr_linear_solver.cpp:308-317).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "csr/hip_matrix.h"
#include "csr/linear_operator.h"
#include "csr/preconditioner.h"
#include "hip_vector_operations.h"
#include "hpb_internal.hpp"
#include "lin_solvers/bicgstab.h"
#include "lin_solvers/bicgstabl.h"
#include "lin_solvers/default_monitor.h"
#include "lin_solvers/gmres.h"

namespace {

using namespace numerical_algos::lin_solvers;

template <typename T, class Prec = void>
struct solver_types {
    using vo_t = hpb::hip_vector_operations<T>;
    using mat_t = csr::hip_matrix<T>;
    using op_t = csr::linear_operator<vo_t, mat_t>;
    using prec_t = std::conditional_t<std::is_void<Prec>::value, identity_preconditioner<vo_t, op_t>, Prec>;
    using ilu_t = csr::ilu0_preconditioner<vo_t, op_t, mat_t>;
    using mon_t = default_monitor<vo_t, hpb::log_std>;
    using gmres_t = gmres<op_t, prec_t, vo_t, mon_t, hpb::log_std, 256>;
    using bicg_t = bicgstabl<op_t, prec_t, vo_t, mon_t, hpb::log_std>;
    using bicg1_t = bicgstab<op_t, prec_t, vo_t, mon_t, hpb::log_std>;
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ||b - A x|| / ||b|| with compensated rows and norms (FP64)
double true_rel_residual(hpb_handle_t h, const csr::hip_matrix<double>& A, const double* b, const double* x,
                         double* work)
{
    hpb::check(hpb_csr_residualc_f64(h, A.rows(), A.rows(), A.nnz(), A.row_ptr(), A.col_ind(), A.data(), b, x, work),
               "true residual");
    double o[4], ob[4];
    hpb::check(hpb_nrm2c_f64(h, (size_t)A.rows(), work, o), "true residual norm");
    hpb::check(hpb_nrm2c_f64(h, (size_t)A.rows(), b, ob), "rhs norm");
    return ob[3] > 0 ? o[3] / ob[3] : o[3];
}

// prec: null for the unpreconditioned solvers, else the preconditioner object
// (applied on the left, gmres.h:396-405; the monitor then sees P r when
// use_precond_resid is set, as in the reference)
template <typename T, class Prec = void>
bool run_solver(const hpb_solver_params& p, hpb::log_std* log, const typename solver_types<T>::vo_t& vo,
                const typename solver_types<T>::op_t& op, const T* b, T* x, hpb_solver_result* res,
                double* history, int history_cap, double rel_tol,
                typename solver_types<T, Prec>::prec_t* prec = nullptr)
{
    using st = solver_types<T, Prec>;
    bool ok;
    auto fill = [&](const typename st::mon_t& mon) {
        res->iters = mon.iters_performed();
        res->resid_norm = (double)mon.resid_norm();
        res->rhs_norm = (double)mon.rhs_norm();
        const auto& hist = mon.convergence_history();
        res->history_len = (int)hist.size();
        if (history)
            for (int i = 0; i < (int)hist.size() && i < history_cap; ++i) {
                history[2 * i] = hist[i].first;
                history[2 * i + 1] = (double)hist[i].second;
            }
    };
    if (p.solver == HPB_SOLVER_GMRES) {
        typename st::gmres_t s(&vo, log, 1);
        if (prec) s.set_preconditioner(prec);
        s.set_use_precond_resid(p.use_precond_resid != 0);
        s.set_restarts((unsigned)p.restart);
        s.set_gram_schmidt(p.gram_schmidt == 1   ? gram_schmidt::cgs2
                           : p.gram_schmidt == 2 ? gram_schmidt::cgs_dgks
                                                 : gram_schmidt::mgs);
        s.set_early_exit(p.early_exit != 0);
        s.set_reorthogonalization(p.reorthogonalization != 0);
        s.monitor().init((T)rel_tol, (T)p.abs_tol, p.max_iters, p.min_iters, false, p.save_history != 0, true);
        ok = s.solve(op, (T*)b, x);
        fill(s.monitor());
        res->breakdown = 0;
    } else if (p.solver == HPB_SOLVER_BICGSTAB) {
        typename st::bicg1_t s(&vo, log, 1);
        if (prec) s.set_preconditioner(prec);
        s.set_use_precond_resid(p.use_precond_resid != 0);
        s.set_resid_recalc_freq(p.resid_recalc_freq);
        s.monitor().init((T)rel_tol, (T)p.abs_tol, p.max_iters, p.min_iters, false, p.save_history != 0, true);
        ok = s.solve(op, (T*)b, x);
        fill(s.monitor());
        res->breakdown = s.breakdown() ? 1 : 0;
    } else {
        typename st::bicg_t s(&vo, log, 1);
        if (prec) s.set_preconditioner(prec);
        s.set_use_precond_resid(p.use_precond_resid != 0);
        s.set_basis_size(p.basis);
        s.set_resid_recalc_freq(p.resid_recalc_freq);
        s.monitor().init((T)rel_tol, (T)p.abs_tol, p.max_iters, p.min_iters, false, p.save_history != 0, true);
        ok = s.solve(op, (T*)b, x);
        fill(s.monitor());
        res->breakdown = s.breakdown() ? 1 : 0;
    }
    return ok;
}

// the device-resident path covers CGS2 GMRES and BiCGStab(L <= 8)
bool use_device(const hpb_solver_params& p)
{
    if (!p.device_resident || p.preconditioner != HPB_PRECOND_NONE) return false;
    if (p.solver == HPB_SOLVER_GMRES) return (p.gram_schmidt == 1 || p.gram_schmidt == 2) && !p.reorthogonalization;
    if (p.solver == HPB_SOLVER_BICGSTAB) return false;
    return p.basis <= 8;
}

int validate(const hpb_solver_params* p)
{
    if (!p) return HPB_ERR_INVALID;
    if (p->solver != HPB_SOLVER_GMRES && p->solver != HPB_SOLVER_BICGSTABL && p->solver != HPB_SOLVER_BICGSTAB)
        return HPB_ERR_INVALID;
    if (p->solver == HPB_SOLVER_GMRES && (p->restart < 1 || p->restart >= 256)) return HPB_ERR_INVALID;
    if (p->gram_schmidt < 0 || p->gram_schmidt > 2) return HPB_ERR_INVALID;
    if (p->solver == HPB_SOLVER_BICGSTABL && (p->basis < 1 || p->basis > 30)) return HPB_ERR_INVALID;
    if (p->max_iters < 1 || !(p->rel_tol >= 0)) return HPB_ERR_INVALID;
    if (p->preconditioner != HPB_PRECOND_NONE && p->preconditioner != HPB_PRECOND_ILU0) return HPB_ERR_INVALID;
    return HPB_OK;
}

}  // namespace

extern "C" {

void hpb_solver_params_default(hpb_solver_params* p)
{
    if (!p) return;
    p->solver = HPB_SOLVER_GMRES;
    p->restart = 50;
    p->basis = 4;
    p->max_iters = 10000;
    p->min_iters = 0;
    p->rel_tol = 1e-12;
    p->abs_tol = 0.0;
    p->high_precision = 1;
    p->compensated_spmv = 0;
    p->gram_schmidt = 1;
    p->early_exit = 1;
    p->reorthogonalization = 0;
    p->resid_recalc_freq = 1;
    p->save_history = 0;
    p->verbose = 0;
    p->device_resident = 1;
    p->preconditioner = HPB_PRECOND_NONE;
    p->use_precond_resid = 1;
    p->refinement = 1;
}

int hpb_csr_solve_f64(hpb_handle_t h, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                      const double* val, const double* b, double* x, const hpb_solver_params* p,
                      hpb_solver_result* res, double* history, int history_cap)
{
    HPB_REQUIRE(h && row_ptr && col_ind && val && b && x && res, HPB_ERR_INVALID, "hpb_csr_solve_f64: null argument");
    HPB_REQUIRE(rows > 0 && nnz > 0, HPB_ERR_INVALID, "hpb_csr_solve_f64: empty matrix");
    HPB_REQUIRE(validate(p) == HPB_OK, HPB_ERR_INVALID, "hpb_csr_solve_f64: invalid solver parameters");
    *res = hpb_solver_result{};
    try {
        using st = solver_types<double>;
        hpb::log_std log(p->verbose);
        st::vo_t vo((size_t)rows, h);
        if (p->high_precision) vo.use_high_precision();
        st::mat_t A(h);
        A.adopt_device(rows, rows, nnz, (int*)row_ptr, (int*)col_ind, (double*)val, false);
        A.use_compensated_rows(p->compensated_spmv != 0);
        st::op_t op(&vo, &A);
        HPB_HIP(hipStreamSynchronize(h->stream));
        const double t0 = now_s();
        bool ok;
        if (use_device(*p)) {
            const int rc = hpb::dsolve<double>(h, rows, nnz, row_ptr, col_ind, val, b, x, *p, res, history, history_cap);
            if (rc != HPB_OK) return rc;
            ok = res->converged != 0;
        } else if (p->preconditioner == HPB_PRECOND_ILU0) {
            typename st::ilu_t ilu(&vo);   // factored inside the timed region
            ilu.set_matrix(A);
            ok = run_solver<double, typename st::ilu_t>(*p, &log, vo, op, b, x, res, history, history_cap,
                                                       p->rel_tol, &ilu);
        } else {
            ok = run_solver<double>(*p, &log, vo, op, b, x, res, history, history_cap, p->rel_tol);
        }
        HPB_HIP(hipStreamSynchronize(h->stream));
        res->seconds = now_s() - t0;
        double* work = nullptr;
        vo.start_use_vector(work);
        res->true_rel_resid = true_rel_residual(h, A, b, x, work);
        vo.free_vector(work);
        // the device-resident solvers stop on a vanishing denominator without touching x
        // again; when x then meets the tolerance on its true residual (the residual
        // vanished: a lucky breakdown, as on the identity) that is convergence, as in
        // the host templates
        if (!ok && res->breakdown && res->rhs_norm > 0.0 &&
            res->true_rel_resid * res->rhs_norm <= p->abs_tol + p->rel_tol * res->rhs_norm) {
            ok = true;
            res->breakdown = 0;
        }
        res->converged = ok ? 1 : 0;
        res->outer_iters = 1;
    } catch (const std::exception& e) {
        hpb::set_error(__FILE__, __LINE__, e.what());
        return HPB_ERR_INVALID;
    }
    return HPB_OK;
}

int hpb_csr_solve_ir_f64(hpb_handle_t h, int rows, int64_t nnz, const int* row_ptr, const int* col_ind,
                         const double* val, const double* b, double* x, const hpb_solver_params* inner,
                         double rel_tol, int max_outer, hpb_solver_result* res)
{
    HPB_REQUIRE(h && row_ptr && col_ind && val && b && x && res, HPB_ERR_INVALID, "hpb_csr_solve_ir_f64: null argument");
    HPB_REQUIRE(rows > 0 && nnz > 0 && max_outer >= 1, HPB_ERR_INVALID, "hpb_csr_solve_ir_f64: bad size");
    HPB_REQUIRE(validate(inner) == HPB_OK, HPB_ERR_INVALID, "hpb_csr_solve_ir_f64: invalid inner parameters");
    *res = hpb_solver_result{};
    try {
        using s64 = solver_types<double>;
        using s32 = solver_types<float>;
        hpb::log_std log(inner->verbose);
        const size_t n = (size_t)rows;
        s64::vo_t vo64(n, h);
        vo64.use_high_precision();
        s32::vo_t vo32(n, h);
        if (inner->high_precision) vo32.use_high_precision();
        s64::mat_t A64(h);
        A64.adopt_device(rows, rows, nnz, (int*)row_ptr, (int*)col_ind, (double*)val, false);
        // the timed region includes building the FP32 copy of the matrix values
        HPB_HIP(hipStreamSynchronize(h->stream));
        const double t0 = now_s();
        float* val32 = nullptr;
        hpb::check(hpb_malloc((void**)&val32, sizeof(float) * (size_t)nnz), "ir: matrix values");
        struct dev_guard {
            void* p;
            ~dev_guard() { hpb_free(p); }
        } val32_guard{val32};
        // FP32 values scaled by a power of two to |a_ij| < 1 (no significand changes):
        // BiCGStab(L)'s A^L r stays far from FP32 overflow on badly scaled operators
        double mscale = 1.0;
        {
            const int rc = hpb::absmax_pow2_scale(h, (size_t)nnz, val, &mscale);
            if (rc != HPB_OK) return rc;
        }
        hpb::check(hpb_scale_convert_f64_to_f32(h, (size_t)nnz, mscale, val, val32), "ir: convert matrix");
        // FP32 values; the index arrays stay the caller's (not owned)
        s32::mat_t A32(h);
        A32.adopt_device(rows, rows, nnz, (int*)row_ptr, (int*)col_ind, val32, false);
        A32.use_compensated_rows(inner->compensated_spmv != 0);
        s32::op_t op32(&vo32, &A32);
        double* r = nullptr;
        float *r32 = nullptr, *d32 = nullptr;
        vo64.start_use_vector(r);
        vo32.start_use_vector(r32);
        vo32.start_use_vector(d32);
        struct vec_guard {
            s64::vo_t& v64;
            s32::vo_t& v32;
            double*& r;
            float*& r32;
            float*& d32;
            ~vec_guard()
            {
                v64.free_vector(r);
                v32.free_vector(r32);
                v32.free_vector(d32);
            }
        } vecs{vo64, vo32, r, r32, d32};
        struct cache_guard {
            hpb::dsolver_base* c = nullptr;
            ~cache_guard() { hpb::dsolve_release(c); }
        } inner_cache;
        std::unique_ptr<s32::ilu_t> ilu32;
        double ob[4];
        hpb::check(hpb_nrm2c_f64(h, n, b, ob), "ir: rhs norm");
        const double bnorm = ob[3];
        int outer = 0, inner_total = 0;
        bool conv = false;
        double rn = 0.0;
        if (inner->solver == HPB_SOLVER_BICGSTABL && inner->refinement == 1 && use_device(*inner)) {
            // reliable updates: one FP32 BiCGStab(L) iteration feeding an FP64 solution
            hpb_solver_result ir{};
            int updates = 0;
            const int rc = hpb::dsolve_reliable(h, rows, nnz, row_ptr, col_ind, val32, mscale, val, b, x, r32, d32,
                                                *inner, rel_tol, &ir, &updates);
            if (rc != HPB_OK) return rc;
            HPB_HIP(hipStreamSynchronize(h->stream));
            res->seconds = now_s() - t0;
            res->converged = ir.converged;
            res->iters = ir.iters;
            res->outer_iters = updates;
            res->breakdown = ir.breakdown;
            res->resid_norm = ir.resid_norm;
            res->rhs_norm = bnorm;
            res->true_rel_resid = true_rel_residual(h, A64, b, x, r);
            return HPB_OK;
        }
        for (;;) {
            hpb::check(hpb_csr_residualc_f64(h, rows, rows, nnz, row_ptr, col_ind, val, b, x, r), "ir: residual");
            double o[4];
            hpb::check(hpb_nrm2c_f64(h, n, r, o), "ir: residual norm");
            rn = o[3];
            if (!(rn == rn) || rn - rn != 0.0) break;
            if (rn <= rel_tol * bnorm) {
                conv = true;
                break;
            }
            if (outer >= max_outer) break;
            ++outer;
            // solve (mscale A) d = r / ||r|| in FP32 to the inner tolerance
            hpb::check(hpb_scale_convert_f64_to_f32(h, n, 1.0 / rn, r, r32), "ir: scale residual");
            vo32.assign_scalar(0.0f, d32);
            hpb_solver_result ir{};
            if (use_device(*inner)) {
                const int rc = hpb::dsolve<float>(h, rows, nnz, row_ptr, col_ind, val32, r32, d32, *inner, &ir,
                                                  nullptr, 0, &inner_cache.c);
                if (rc != HPB_OK) return rc;
            } else if (inner->preconditioner == HPB_PRECOND_ILU0) {
                if (!ilu32) {
                    ilu32.reset(new s32::ilu_t(&vo32));
                    ilu32->set_matrix(A32);
                }
                run_solver<float, s32::ilu_t>(*inner, &log, vo32, op32, r32, d32, &ir, nullptr, 0, inner->rel_tol,
                                              ilu32.get());
            } else {
                run_solver<float>(*inner, &log, vo32, op32, r32, d32, &ir, nullptr, 0, inner->rel_tol);
            }
            inner_total += ir.iters;
            // (mscale A) d = r / ||r||  =>  A (||r|| mscale d) = r
            hpb::check(hpb_axpy_f32_to_f64(h, n, rn * mscale, d32, x), "ir: update");
        }
        HPB_HIP(hipStreamSynchronize(h->stream));
        res->seconds = now_s() - t0;
        res->converged = conv ? 1 : 0;
        res->iters = inner_total;
        res->outer_iters = outer;
        res->resid_norm = rn;
        res->rhs_norm = bnorm;
        res->true_rel_resid = true_rel_residual(h, A64, b, x, r);
    } catch (const std::exception& e) {
        hpb::set_error(__FILE__, __LINE__, e.what());
        return HPB_ERR_INVALID;
    }
    return HPB_OK;
}

}  // extern "C"
