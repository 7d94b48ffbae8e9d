// Host run of the solver templates (GMRES, BiCGStab(L), monitor) on the CPU test
// double.  Problem: the reference's 1-D upwind convection-diffusion operator and
// its diagonal preconditioner (source/linear_operator.hpp:28-90, driver
// test_linear_solver.cpp:65-196).  Prints one line per case:
//   name converged iters true_rel_resid max_err_vs_direct
// plus the failure-handling cases (NaN right-hand side, iteration cap).
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "cpu_vector_operations.h"
#include "lin_solvers/bicgstab.h"
#include "lin_solvers/bicgstabl.h"
#include "lin_solvers/default_monitor.h"
#include "lin_solvers/gmres.h"

using namespace numerical_algos::lin_solvers;

template <class VO>
struct system_operator {
    using T = typename VO::scalar_type;
    using T_vec = typename VO::vector_type;
    int sz;
    T a, re, h;
    system_operator(const VO* vo, T a_, T re_) : sz((int)vo->sz_), a(a_), re(re_), h(T(1) / T(vo->sz_)) {}
    void apply(const T_vec& x, T_vec& f) const
    {
        for (int i = 0; i < sz; ++i) {
            T xm = i > 0 ? x[i - 1] : T(0), xp = i + 1 < sz ? x[i + 1] : T(0);
            f[i] = a * (x[i] - xm) / h - (T(1) / re) * (xp - T(2) * x[i] + xm) / (h * h);
        }
    }
};

// the identity: BiCGStab(L) reaches the exact solution in its first BiCG step, after
// which every denominator vanishes (a lucky breakdown)
template <class VO>
struct identity_operator {
    using T = typename VO::scalar_type;
    using T_vec = typename VO::vector_type;
    int sz;
    explicit identity_operator(const VO* vo) : sz((int)vo->sz_) {}
    void apply(const T_vec& x, T_vec& f) const
    {
        for (int i = 0; i < sz; ++i) f[i] = x[i];
    }
};

template <class VO, class Op>
struct null_prec {   // the solver's preconditioner type; never set
    using T_vec = typename VO::vector_type;
    void set_operator(const Op*) {}
    void apply(T_vec&) const {}
};

template <class VO, class Op>
struct prec_operator {
    using T = typename VO::scalar_type;
    using T_vec = typename VO::vector_type;
    const Op* op = nullptr;
    void set_operator(const Op* o) { op = o; }
    void apply(T_vec& x) const
    {
        for (int i = 0; i < op->sz; ++i) x[i] /= (op->a / op->h - (T(1) / op->re) * (-T(2)) / (op->h * op->h));
    }
};

// direct tridiagonal solve of the same operator (Thomas algorithm, long double)
static std::vector<double> direct(int n, double a, double re, const std::vector<double>& b)
{
    long double h = 1.0L / n, lo = -a / h - 1.0L / (re * h * h), di = a / h + 2.0L / (re * h * h),
                up = -1.0L / (re * h * h);
    std::vector<long double> c(n), d(n);
    c[0] = up / di;
    d[0] = b[0] / di;
    for (int i = 1; i < n; ++i) {
        long double m = di - lo * c[i - 1];
        c[i] = up / m;
        d[i] = (b[i] - lo * d[i - 1]) / m;
    }
    std::vector<double> x(n);
    long double xi = d[n - 1];
    x[n - 1] = (double)xi;
    for (int i = n - 2; i >= 0; --i) {
        xi = d[i] - c[i] * xi;
        x[i] = (double)xi;
    }
    return x;
}

template <class VO>
static void report(const char* name, const VO& vo, const system_operator<VO>& A, bool conv, int iters, double* x,
                   double* b, const std::vector<double>& xd)
{
    double* r = nullptr;
    vo.start_use_vector(r);
    A.apply(x, r);
    double rn = 0, bn = 0, err = 0;
    for (size_t i = 0; i < vo.sz_; ++i) {
        rn += (b[i] - r[i]) * (b[i] - r[i]);
        bn += b[i] * b[i];
        err = std::fmax(err, std::fabs(x[i] - xd[i]));
    }
    std::printf("%s %d %d %.6e %.6e\n", name, conv ? 1 : 0, iters, std::sqrt(rn / bn), err);
    vo.free_vector(r);
}

// smallest checked residual norm: the monitor keeps the iterate, the value is
// recomputed here from the norm of that iterate's residual
template <class Mon>
static double mon_min(const Mon& mon)
{
    return (double)mon.min_resid_norm();
}

template <class VO>
static void run_all(const char* tag, bool high)
{
    typedef system_operator<VO> op_t;
    typedef prec_operator<VO, op_t> prec_t;
    typedef default_monitor<VO, hpb::log_std> mon_t;
    const int n = 200;
    VO vo(n);
    if (high) vo.use_high_precision();
    op_t A(&vo, 1.0, 10.0);
    std::vector<double> bh(n);
    for (int i = 0; i < n; ++i) bh[i] = std::sin(3.0 * i / n) + 0.1;
    std::vector<double> xd = direct(n, 1.0, 10.0, bh);
    double *b = nullptr, *x = nullptr;
    vo.start_use_vector(b);
    vo.start_use_vector(x);
    for (int i = 0; i < n; ++i) b[i] = bh[i];
    char name[128];
    const gram_schmidt variants[3] = {gram_schmidt::mgs, gram_schmidt::cgs2, gram_schmidt::cgs_dgks};
    const char* vname[3] = {"mgs", "cgs2", "dgks"};
    for (int gs = 0; gs < 3; ++gs)
        for (int pc = 0; pc < 2; ++pc) {
            gmres<op_t, prec_t, VO, mon_t, hpb::log_std, 100> s(&vo);
            prec_t P;
            if (pc) s.set_preconditioner(&P);
            s.set_restarts(30);
            s.set_gram_schmidt(variants[gs]);
            s.monitor().init(1e-10, 0.0, 5000);
            vo.assign_scalar(0.0, x);
            bool ok = s.solve(A, b, x);
            std::snprintf(name, sizeof name, "%s_gmres_%s_prec%d", tag, vname[gs], pc);
            report(name, vo, A, ok, s.monitor().iters_performed(), x, b, xd);
        }
    for (int L : {1, 2, 4}) {
        bicgstabl<op_t, prec_t, VO, mon_t, hpb::log_std> s(&vo);
        s.set_basis_size(L);
        s.set_resid_recalc_freq(1);
        s.monitor().init(1e-10, 0.0, 5000);
        vo.assign_scalar(0.0, x);
        bool ok = s.solve(A, b, x);
        std::snprintf(name, sizeof name, "%s_bicgstabl%d", tag, L);
        report(name, vo, A, ok, s.monitor().iters_performed(), x, b, xd);
    }
    for (int pc = 0; pc < 2; ++pc)
        for (int pr = 0; pr < 2; ++pr)
            for (int rf : {0, 5}) {
                bicgstab<op_t, prec_t, VO, mon_t, hpb::log_std> s(&vo);
                prec_t P;
                if (pc) s.set_preconditioner(&P);
                s.set_use_precond_resid(pr != 0);
                s.set_resid_recalc_freq(rf);
                s.monitor().init(1e-10, 0.0, 5000);
                vo.assign_scalar(0.0, x);
                bool ok = s.solve(A, b, x);
                std::snprintf(name, sizeof name, "%s_bicgstab_prec%d_pr%d_rf%d", tag, pc, pr, rf);
                report(name, vo, A, ok && !s.breakdown(), s.monitor().iters_performed(), x, b, xd);
            }
    // lucky breakdown on the identity: x = b exactly, reported as converged
    //   name converged iterations max|x - b| breakdown
    for (int L : {1, 2, 4}) {
        typedef identity_operator<VO> id_t;
        typedef null_prec<VO, id_t> idp_t;
        id_t I(&vo);
        bicgstabl<id_t, idp_t, VO, mon_t, hpb::log_std> s(&vo);
        s.set_basis_size(L);
        s.monitor().init(1e-12, 0.0, 100);
        vo.assign_scalar(0.0, x);
        const bool ok = s.solve(I, b, x);
        double err = 0;
        for (int i = 0; i < n; ++i) err = std::max(err, std::fabs(x[i] - b[i]));
        std::printf("%s_identity_bicgstabl%d %d %d %.6e %d\n", tag, L, ok ? 1 : 0, s.monitor().iters_performed(), err,
                    s.breakdown() ? 1 : 0);
    }
    // failure handling: iteration cap, NaN in the right-hand side
    {
        gmres<op_t, prec_t, VO, mon_t, hpb::log_std, 100> s(&vo);
        s.set_restarts(5);
        s.monitor().init(1e-14, 0.0, 10);
        vo.assign_scalar(0.0, x);
        bool ok = s.solve(A, b, x);
        std::printf("%s_cap %d %d %d\n", tag, ok ? 1 : 0, s.monitor().iters_performed(),
                    s.monitor().is_valid_number() ? 1 : 0);
        b[7] = std::numeric_limits<double>::quiet_NaN();
        vo.assign_scalar(0.0, x);
        ok = s.solve(A, b, x);
        std::printf("%s_nan %d %d %d\n", tag, ok ? 1 : 0, s.monitor().iters_performed(),
                    s.monitor().is_valid_number() ? 1 : 0);
        b[7] = bh[7];
    }
    // out_min_resid_norm: with an iteration cap that stops the solve early, the
    // returned x is the checked iterate of smallest residual; printed as
    //   name min_resid_norm(monitor) ||b - A x||(returned x) last_resid_norm(monitor)
    {
        auto true_resid = [&](double* xx) {
            double* r = nullptr;
            vo.start_use_vector(r);
            A.apply(xx, r);
            double rn = 0;
            for (int i = 0; i < n; ++i) rn += (b[i] - r[i]) * (b[i] - r[i]);
            vo.free_vector(r);
            return std::sqrt(rn);
        };
        auto min_line = [&](const char* what, const mon_t& mon) {
            std::printf("%s_minres_%s %.17e %.17e %.17e\n", tag, what, (double)mon_min(mon), true_resid(x),
                        (double)mon.resid_norm());
        };
        {
            bicgstab<op_t, prec_t, VO, mon_t, hpb::log_std> s(&vo);
            s.monitor().init(1e-14, 0.0, 7, 0, true);
            vo.assign_scalar(0.0, x);
            s.solve(A, b, x);
            min_line("bicgstab", s.monitor());
        }
        {
            bicgstabl<op_t, prec_t, VO, mon_t, hpb::log_std> s(&vo);
            s.set_basis_size(2);
            s.monitor().init(1e-14, 0.0, 6, 0, true);
            vo.assign_scalar(0.0, x);
            s.solve(A, b, x);
            min_line("bicgstabl", s.monitor());
        }
        {
            gmres<op_t, prec_t, VO, mon_t, hpb::log_std, 100> s(&vo);
            s.set_restarts(3);
            s.monitor().init(1e-14, 0.0, 9, 0, true);
            vo.assign_scalar(0.0, x);
            s.solve(A, b, x);
            min_line("gmres", s.monitor());
        }
    }
    vo.free_vector(b);
    vo.free_vector(x);
}

int main()
{
    run_all<cpu_vector_operations<double>>("plain", false);
    run_all<cpu_vector_operations<double>>("high", true);
    run_all<cpu_vector_operations_multi<double>>("multi", true);
    return 0;
}
