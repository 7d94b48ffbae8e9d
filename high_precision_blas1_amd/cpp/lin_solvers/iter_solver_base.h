// iter_solver_base.h -- common base of the Krylov solvers.
//
// Interface of numerical_algos::lin_solvers::iter_solver_base
// (source/numerical_algos/lin_solvers/iter_solver_base.h:50-95): monitor(),
// set_preconditioner(prec, own), virtual solve(A, b, x).
#pragma once
#include <string>

#include "../utils/hpb_log.h"

namespace numerical_algos {
namespace lin_solvers {

template <class LinearOperator, class Preconditioner, class VectorOperations, class Monitor, class Log>
class iter_solver_base : public hpb::logged_obj_base<Log> {
public:
    typedef typename VectorOperations::scalar_type scalar_type;
    typedef typename VectorOperations::vector_type vector_type;
    typedef LinearOperator linear_operator_type;
    typedef Preconditioner preconditioner_type;
    typedef VectorOperations vector_operations_type;
    typedef Monitor monitor_type;
    typedef Log log_type;

protected:
    mutable monitor_type monitor_;
    const vector_operations_type* vec_ops_;
    preconditioner_type* prec_ = nullptr;
    bool own_prec_ = false;

public:
    iter_solver_base(const vector_operations_type* vec_ops, Log* log, int obj_log_lev, const std::string& prefix)
        : hpb::logged_obj_base<Log>(log, obj_log_lev, prefix), monitor_(*vec_ops, log, obj_log_lev), vec_ops_(vec_ops)
    {
        monitor_.set_log_msg_prefix(prefix + monitor_.get_log_msg_prefix());
    }
    virtual ~iter_solver_base()
    {
        if (prec_ && own_prec_) delete prec_;
    }
    Monitor& monitor() { return monitor_; }
    const Monitor& monitor() const { return monitor_; }
    void set_preconditioner(preconditioner_type* prec, bool own_prec = false)
    {
        prec_ = prec;
        own_prec_ = own_prec;
    }
    virtual bool solve(const linear_operator_type& A, const vector_type& b, vector_type& x) const = 0;
};

// Identity preconditioner (the configs are unpreconditioned; the reference's
// csr::preconditioner::apply is a no-op without L/U files, csr/preconditioner.h:63-67).
template <class VectorOperations, class LinearOperator>
struct identity_preconditioner {
    typedef typename VectorOperations::vector_type vector_type;
    void set_operator(const LinearOperator*) {}
    void apply(vector_type&) const {}
};

}  // namespace lin_solvers
}  // namespace numerical_algos
