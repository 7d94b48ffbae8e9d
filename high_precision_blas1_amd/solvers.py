"""ctypes view of the solver and matrix-builder entry points of ``libhpb.so``.

Mirrors ``hpb_solver_params`` / ``hpb_solver_result`` of ``include/hpb.h`` and
offers small helpers to build the benchmark matrices on the device and to call
``hpb_csr_solve_f64`` / ``hpb_csr_solve_ir_f64``.  All work happens in the GPU
library; this module only marshals arguments.
"""
import ctypes

import numpy as np

from . import DeviceArray, call

GMRES = 0
BICGSTABL = 1
BICGSTAB = 2
PRECOND_NONE = 0
PRECOND_ILU0 = 1
GS_MGS = 0       # gram_schmidt: the reference's modified Gram-Schmidt
GS_CGS2 = 1      # two classical passes
GS_DGKS = 2      # classical, second pass only when the first shrank w by more than 1/sqrt(2)


class SolverParams(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("restart", ctypes.c_int), ("basis", ctypes.c_int),
                ("max_iters", ctypes.c_int), ("min_iters", ctypes.c_int), ("rel_tol", ctypes.c_double),
                ("abs_tol", ctypes.c_double), ("high_precision", ctypes.c_int), ("compensated_spmv", ctypes.c_int),
                ("gram_schmidt", ctypes.c_int), ("early_exit", ctypes.c_int), ("reorthogonalization", ctypes.c_int),
                ("resid_recalc_freq", ctypes.c_int), ("save_history", ctypes.c_int), ("verbose", ctypes.c_int),
                ("device_resident", ctypes.c_int), ("preconditioner", ctypes.c_int),
                ("use_precond_resid", ctypes.c_int), ("refinement", ctypes.c_int)]


class SolverResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iters", ctypes.c_int), ("outer_iters", ctypes.c_int),
                ("breakdown", ctypes.c_int), ("history_len", ctypes.c_int), ("resid_norm", ctypes.c_double),
                ("rhs_norm", ctypes.c_double), ("true_rel_resid", ctypes.c_double), ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


def params(**kw):
    from . import lib
    p = SolverParams()
    lib().hpb_solver_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class DeviceCSR:
    """Device CSR arrays (int32 row_ptr / col_ind, float64 values)."""

    def __init__(self, handle, rows, nnz):
        self.handle = handle
        self.rows, self.nnz = int(rows), int(nnz)
        self.rp = DeviceArray(self.rows + 1, np.int32, handle)
        self.col = DeviceArray(self.nnz, np.int32, handle)
        self.val = DeviceArray(self.nnz, np.float64, handle)

    @classmethod
    def stencil7(cls, handle, nx, ny, nz, coef):
        rows, nnz = ctypes.c_int64(), ctypes.c_int64()
        call("hpb_stencil_nnz", nx, ny, nz, ctypes.byref(rows), ctypes.byref(nnz))
        A = cls(handle, rows.value, nnz.value)
        c = np.asarray(coef, dtype=np.float64)
        call("hpb_csr_stencil7_f64", handle.ptr, nx, ny, nz, c.ctypes.data_as(ctypes.c_void_p), A.rp.ptr, A.col.ptr,
             A.val.ptr)
        return A

    @classmethod
    def diffusion2d_lognormal(cls, handle, m, sigma, seed):
        kap = DeviceArray(m * m, np.float64, handle)
        call("hpb_lognormal_field_f64", handle.ptr, m * m, ctypes.c_double(sigma), seed, kap.ptr)
        rows, nnz = ctypes.c_int64(), ctypes.c_int64()
        call("hpb_stencil_nnz", m, m, 1, ctypes.byref(rows), ctypes.byref(nnz))
        A = cls(handle, rows.value, nnz.value)
        call("hpb_csr_diffusion2d_f64", handle.ptr, m, kap.ptr, A.rp.ptr, A.col.ptr, A.val.ptr)
        kap.free()
        return A

    @classmethod
    def from_scipy(cls, handle, M):
        A = cls.__new__(cls)
        A.handle = handle
        A.rows, A.nnz = M.shape[0], M.nnz
        A.rp = DeviceArray.from_numpy(M.indptr.astype(np.int32), handle)
        A.col = DeviceArray.from_numpy(M.indices.astype(np.int32), handle)
        A.val = DeviceArray.from_numpy(M.data.astype(np.float64), handle)
        return A

    def spmv(self, x, y, alpha=1.0, beta=0.0, compensated=False):
        call("hpb_csr_spmvc_f64" if compensated else "hpb_csr_spmv_f64", self.handle.ptr, self.rows, self.rows,
             self.nnz, self.rp.ptr, self.col.ptr, self.val.ptr, ctypes.c_double(alpha), x.ptr, ctypes.c_double(beta),
             y.ptr)

    def free(self):
        for a in (self.rp, self.col, self.val):
            a.free()


# benchmark matrices: coefficients {z-, y-, x-, diagonal, x+, y+, z+}
POISSON7 = (-1.0, -1.0, -1.0, 6.0, -1.0, -1.0, -1.0)
# first-order upwind convection-diffusion, velocity (1, 0.5, 0.25), cell Peclet
# number 100 on the largest component (nu = h |v|_max / 100), scaled by h^2 / nu:
# lower neighbour -1 - Pe_d, upper neighbour -1, diagonal 6 + sum_d Pe_d
CONVDIFF7 = (-26.0, -51.0, -101.0, 181.0, -1.0, -1.0, -1.0)
# anisotropic diffusion, coefficient ratio k_z / k_x = 1e4
ANISO7 = (-1e4, -1.0, -1.0, 4.0 + 2e4, -1.0, -1.0, -1e4)


def solve(handle, A, b, x, p, history_cap=0):
    res = SolverResult()
    hist = np.zeros(2 * max(history_cap, 1))
    call("hpb_csr_solve_f64", handle.ptr, A.rows, A.nnz, A.rp.ptr, A.col.ptr, A.val.ptr, b.ptr, x.ptr,
         ctypes.byref(p), ctypes.byref(res), hist.ctypes.data_as(ctypes.c_void_p) if history_cap else None,
         history_cap)
    out = res.as_dict()
    if history_cap:
        k = min(res.history_len, history_cap)
        out["history"] = hist[:2 * k].reshape(k, 2).tolist()
    return out


def solve_ir(handle, A, b, x, inner, rel_tol=1e-12, max_outer=50):
    res = SolverResult()
    call("hpb_csr_solve_ir_f64", handle.ptr, A.rows, A.nnz, A.rp.ptr, A.col.ptr, A.val.ptr, b.ptr, x.ptr,
         ctypes.byref(inner), ctypes.c_double(rel_tol), max_outer, ctypes.byref(res))
    return res.as_dict()
