#!/usr/bin/env python3
"""Benchmark of the compensated BLAS1 + Krylov path on one MI355X.

Headline (value): FP64 Ogita-Rump-Oishi Dot2 of two device-resident vectors of
N = 2^27 elements (2 GiB read per step, far above the 256 MB MALL).  The inputs
are a GenDot pair built on the device (kappa 1e16, seed 0), whose exact dot is
known.  One "step" = one Dot2 launch over the pair.  The GB/s figure is
algorithmic bytes (16 B per element) divided by the kernel time, measured with
HIP events on the handle's stream over K back-to-back launches.

The same JSON line carries:
  roofline      the Dot2 kernel against the 8 TB/s HBM3E peak;
  cpu_baseline  the reference's own CPU Dot2 (oracle/_ref, compiled from
                source/common/threaded_reduction.h) on all host cores, on a
                2^25-element sample of the same vectors;
  configs       the other benchmark configurations, measured in the same run:
    cpu_plain      N = 2^20, kappa 1e16: reference CPU Dot2 / naive / Sum2, errors and times
    cond_sweep     kappa 1e4..1e32 at N = 2^27: Dot2 GB/s and relative errors (Dot2, naive)
    blas1          Sum2, compensated nrm2, axpy, compensated axpby GB/s at N = 2^27
    spmv           256^3 7-point Poisson, plain and compensated rows, GB/s
    krylov         128^3 upwind convection-diffusion (Peclet 100): GMRES(50), BiCGStab(4),
                   naive vs compensated dots, to 1e-12 true relative residual
    mixed          the same matrix and 256^3 anisotropic diffusion: FP32-inner GMRES(30) /
                   BiCGStab(2) with FP64 refinement vs the same solvers in FP64, time to 1e-12
    multidot       fused compensated multi-dot, k = 1..64, GB/s
    batch          512 2-D diffusion systems (sides 64..1024), batched BiCGStab(2) to 1e-10
    next           MatrixMarket round trip, triangular solves, ILU(0), preconditioned and
                   plain BiCGStab, complex Dot2

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--headline-only]
                       [--suite a,b] [--op dot2|copy|nrm2c|exact]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from fractions import Fraction

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Dot2 GB/s & %HBM peak; Krylov time-to-1e-12, mixed+refined vs FP64 plain"


def log(*a):
    print("bench:", *a, file=sys.stderr, flush=True)


def exact_value(limbs):
    """Exact dot from the generator's 256-bit fixed-point limbs (LSB 2^-106)."""
    v = 0
    for i in range(4):
        v |= int(limbs[i]) << (64 * i)
    if v >> 255:
        v -= 1 << 256
    return Fraction(v, 1 << 106)


def rel_err(res, exact):
    if exact == 0:
        return float("inf") if res != 0 else 0.0
    e = abs(Fraction(float(res)) - exact) / abs(exact)
    return min(float(e), 1.0)


def gendot(H, h, n, cond, seed, x, y):
    import numpy as np
    limbs = np.zeros(4, dtype=np.uint64)
    dd = np.zeros(3)
    H.call("hpb_gendot_f64", h.ptr, n, ctypes.c_double(float(np.log2(cond))), seed, n // 2, x.ptr, y.ptr,
           limbs.ctypes.data_as(ctypes.c_void_p), dd.ctypes.data_as(ctypes.c_void_p))
    return exact_value(limbs), dd


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_reduce_dot2_n2p27_traffic.md")
WARM_S = 1.0   # seconds of naive-dot launches before the timed windows


def profiled_traffic(n, cfg, nontemporal):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a
    copy kernel of known size) for this exact launch, from the committed profile; None
    when the profile does not cover this n and launch shape."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    fields = {}
    for line in text.splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            fields["traffic"] = int(line.split(":")[1])
        elif line.startswith("launch:"):
            fields.update(kv.split("=", 1) for kv in line.split(":", 1)[1].split() if "=" in kv)
    want = {"n": n, "block": cfg[0], "unroll": cfg[1], "grid": cfg[2], "vector_width": cfg[3],
            "nontemporal": int(bool(nontemporal))}
    if "traffic" not in fields or any(int(fields.get(k, -1)) != v for k, v in want.items()):
        return None
    return fields["traffic"]


def timed(h, fn, args, steps, warmup):
    for _ in range(warmup):
        rc = fn(*args)
        if rc:
            from high_precision_blas1_amd import lib
            raise RuntimeError(lib().hpb_last_error().decode())
    h.sync()
    h.timer_start()
    for _ in range(steps):
        fn(*args)
    return h.timer_stop() / steps


def cpu_baseline(H, h, x, y, n, budget_s=12.0):
    """Reference CPU Dot2 on the host cores over a bounded sample of the vectors."""
    import numpy as np
    import oracle as O
    sample = min(n, 1 << 25)
    hx = np.empty(sample)
    hy = np.empty(sample)
    H.call("hpb_memcpy_d2h", h.ptr, hx.ctypes.data_as(ctypes.c_void_p), x.ptr, sample * 8)
    H.call("hpb_memcpy_d2h", h.ptr, hy.ctypes.data_as(ctypes.c_void_p), y.ptr, sample * 8)
    R = O.ref()
    cores = os.cpu_count() or 1
    if R is not None:
        kind = "reference"
        fthr = lambda: R.ref_dot2_threaded(sample, O.ptr(hx), O.ptr(hy), cores)
        fseq = lambda: R.ref_dot2(sample, O.ptr(hx), O.ptr(hy))
    else:
        kind = "port"
        fthr = lambda: O.lib().orc_dot2_threaded(sample, O.ptr(hx), O.ptr(hy), cores)
        fseq = lambda: O.dot2(hx, hy)

    def run(f, budget):
        ts = []
        t_end = time.perf_counter() + budget
        while len(ts) < 4 or (time.perf_counter() < t_end and len(ts) < 200):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        core = ts[1:-1]  # drop min and max (test_vector_operations.cpp:44-49)
        return sum(core) / len(core), len(ts), core[0], core[-1]

    model = "unknown"
    try:
        info = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    info.setdefault(k.strip(), v.strip())
                if not line.strip() and info:
                    break
        model = "%s (vendor %s, family %s, model %s, stepping %s, %s threads)" % (
            info.get("model name", "?"), info.get("vendor_id", "?"), info.get("cpu family", "?"),
            info.get("model", "?"), info.get("stepping", "?"), os.cpu_count())
    except OSError:
        pass
    t_thr, reps_thr, tmin, tmax = run(fthr, budget_s * 0.6)
    t_seq, reps_seq, _, _ = run(fseq, budget_s * 0.4)
    nbytes = 16.0 * sample
    return {
        "value": round(nbytes / t_thr / 1e9, 3),
        "unit": "GB/s",
        "cores": cores,
        "cpu_model": model,
        "kind": kind,
        "sample": "FP64 Dot2 over the first %d elements of the benchmark vectors (%d MiB), threaded_reduction "
                  "with %d threads, mean of %d runs without min/max" % (sample, int(nbytes) >> 20, cores, reps_thr),
        "ms_mean_min_max": [round(1e3 * t_thr, 3), round(1e3 * tmin, 3), round(1e3 * tmax, 3)],
        "single_thread_gbs": round(nbytes / t_seq / 1e9, 3),
        "single_thread_runs": reps_seq,
    }


def cfg_cpu_plain():
    """N = 2^20, kappa 1e16, seed 0 on the host: the reference's CPU Dot2, naive dot,
    threaded Dot2 and Sum2 of the 2N exact product terms, relative errors and times."""
    import numpy as np
    import oracle as O
    n = 1 << 20
    x, y, limbs, dd = O.gendot(n, 1e16, 0)
    ex = exact_value(limbs)
    R = O.ref()
    cores = os.cpu_count() or 1
    # exact split of every product (Dekker's TwoProduct, no FMA needed)
    f = 134217729.0
    def split(v):
        c = f * v
        hi = c - (c - v)
        return hi, v - hi
    p = x * y
    xh, xl = split(x)
    yh, yl = split(y)
    e = xl * yl - (((p - xh * yh) - xl * yh) - xh * yl)
    terms = np.empty(2 * n)
    terms[0::2] = p
    terms[1::2] = e
    if R is not None:
        fns = {"dot2": lambda: R.ref_dot2(n, O.ptr(x), O.ptr(y)),
               "dot_naive": lambda: R.ref_dot_naive(n, O.ptr(x), O.ptr(y)),
               "dot2_threaded": lambda: R.ref_dot2_threaded(n, O.ptr(x), O.ptr(y), cores),
               "sum2_of_products": lambda: R.ref_sum2_threaded(2 * n, O.ptr(terms), 1)}
    else:
        fns = {"dot2": lambda: O.dot2(x, y), "dot_naive": lambda: O.dot_naive(x, y),
               "dot2_threaded": lambda: O.lib().orc_dot2_threaded(n, O.ptr(x), O.ptr(y), cores),
               "sum2_of_products": lambda: O.sum2(terms)}
    out = {"n": n, "cond_target": 1e16, "cond_achieved": float(dd[2]), "kind": "reference" if R else "port",
           "threads": cores}
    for k, fn in fns.items():
        ts = []
        for _ in range(7):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        out[k] = {"rel_err": rel_err(r, ex), "ms": round(1e3 * sum(ts[1:-1]) / 5, 4)}
    return out


def cfg_cond_sweep(H, h, n, log):
    """kappa 1e4..1e32 at N = 2^27 (seed 0): Dot2 GB/s, relative errors of Dot2 and naive dot."""
    import numpy as np
    L = H.lib()
    x = H.DeviceArray(n, np.float64, h)
    y = H.DeviceArray(n, np.float64, h)
    d = H.DeviceArray(4, np.float64, h)
    rows = []
    for cond in (1e4, 1e8, 1e16, 1e24, 1e32):
        log("cond sweep %g" % cond)
        ex, dd = gendot(H, h, n, cond, 0, x, y)
        ms = timed(h, L.hpb_dot2_f64_async, (h.ptr, n, x.ptr, y.ptr, d.ptr), 20, 3)
        g = d.numpy()
        out = np.zeros(4)
        H.call("hpb_dot_f64", h.ptr, n, x.ptr, y.ptr, out.ctypes.data_as(ctypes.c_void_p))
        row = {"cond_target": cond, "cond_achieved": float(dd[2]), "dot2_gbs": round(16.0 * n / ms / 1e6, 1),
               "dot2_rel_err": rel_err(g[2], ex), "naive_rel_err": rel_err(out[2], ex)}
        # device exact dot (long accumulator) against the generator's tracked value
        ms_x = timed(h, L.hpb_exact_dot_f64_async, (h.ptr, n, x.ptr, y.ptr, d.ptr), 20, 5)
        e3 = d.numpy()[:3]
        hi = float(ex)
        row["exact_dot_gbs"] = round(16.0 * n / ms_x / 1e6, 1)
        row["exact_dot_equals_generator"] = bool(e3[2] == 0 and e3[0] == hi and e3[1] == float(ex - Fraction(hi)))
        rows.append(row)
    for a in (x, y, d):
        a.free()
    return rows


def cfg_blas1(H, h, x, y, n):
    import numpy as np
    L = H.lib()
    z = H.DeviceArray(n, np.float64, h)
    d = H.DeviceArray(4, np.float64, h)
    H.call("hpb_copy_f64", h.ptr, n, y.ptr, z.ptr)
    c = ctypes.c_double
    res = {}
    ms = timed(h, L.hpb_nrm2c_f64_async, (h.ptr, n, x.ptr, d.ptr), 20, 3)
    res["nrm2c_gbs"] = round(8.0 * n / ms / 1e6, 1)
    out = np.zeros(4)
    ms = timed(h, L.hpb_sum2_f64, (h.ptr, n, x.ptr, out.ctypes.data_as(ctypes.c_void_p)), 20, 3)
    res["sum2_gbs_with_readback"] = round(8.0 * n / ms / 1e6, 1)
    ms = timed(h, L.hpb_axpy_f64, (h.ptr, n, c(1e-3), x.ptr, z.ptr), 20, 3)
    res["axpy_gbs"] = round(24.0 * n / ms / 1e6, 1)
    ms = timed(h, L.hpb_axpbyc_f64, (h.ptr, n, c(0.5), x.ptr, c(-0.25), y.ptr, z.ptr), 20, 3)
    res["axpbyc_gbs"] = round(24.0 * n / ms / 1e6, 1)
    ms = timed(h, L.hpb_dot_f64_async, (h.ptr, n, x.ptr, y.ptr, d.ptr), 20, 3)
    res["dot_naive_gbs"] = round(16.0 * n / ms / 1e6, 1)
    # end to end: blocking call with the scalar read back to the host, host clock
    for _ in range(3):
        H.call("hpb_dot2_f64", h.ptr, n, x.ptr, y.ptr, out.ctypes.data_as(ctypes.c_void_p))
    ts = []
    for _ in range(20):
        t0 = time.perf_counter()
        L.hpb_dot2_f64(h.ptr, n, x.ptr, y.ptr, out.ctypes.data_as(ctypes.c_void_p))
        ts.append(time.perf_counter() - t0)
    ts.sort()
    res["dot2_end_to_end_ms"] = round(1e3 * sum(ts[1:-1]) / (len(ts) - 2), 4)
    res["dot2_end_to_end_gbs"] = round(16.0 * n / (sum(ts[1:-1]) / (len(ts) - 2)) / 1e9, 1)
    z.free()
    d.free()
    return res


def cfg_spmv(H, h, log):
    import numpy as np
    import high_precision_blas1_amd.solvers as S
    nx = 256
    log("spmv 256^3")
    A = S.DeviceCSR.stencil7(h, nx, nx, nx, S.POISSON7)
    xs = H.DeviceArray(A.rows, np.float64, h)
    ys = H.DeviceArray(A.rows, np.float64, h)
    H.call("hpb_random_uniform_f64", h.ptr, A.rows, 1, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), xs.ptr)
    nbytes = 12.0 * A.nnz + 4.0 * (A.rows + 1) + 16.0 * A.rows
    L = H.lib()
    res = {"rows": A.rows, "nnz": A.nnz, "algorithmic_bytes": int(nbytes)}
    for name, fn in (("plain", L.hpb_csr_spmv_f64), ("compensated", L.hpb_csr_spmvc_f64)):
        ms = timed(h, fn, (h.ptr, A.rows, A.rows, A.nnz, A.rp.ptr, A.col.ptr, A.val.ptr, ctypes.c_double(1.0), xs.ptr,
                           ctypes.c_double(0.0), ys.ptr), 50, 20)
        res[name + "_gbs"] = round(nbytes / ms / 1e6, 1)
        res[name + "_ms"] = round(ms, 4)
        res[name + "_hbm_frac"] = round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)
    A.free()
    xs.free()
    ys.free()
    return res


def _solve_problem(H, h, S, nx, coef, seed):
    """Matrix, b = A x* with x* uniform in [-1, 1] (Philox seed), and x*."""
    import numpy as np
    A = S.DeviceCSR.stencil7(h, nx, nx, nx, coef)
    xs = H.DeviceArray(A.rows, np.float64, h)
    H.call("hpb_random_uniform_f64", h.ptr, A.rows, seed, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), xs.ptr)
    b = H.DeviceArray(A.rows, np.float64, h)
    A.spmv(xs, b, compensated=True)
    return A, b, xs


# ---- algorithmic bytes of the solvers (inputs read once, outputs written once;
# the SpMV reads x once, as the stencil rows of neighbouring threads share it)
def _spmv_bytes(n, nnz, sval, svec, residual=False):
    return (4 + sval) * nnz + 4 * (n + 1) + 2 * svec * n + (svec * n if residual else 0)


def _bicgstabl_bytes(L, n, nnz, sval, s, cycles, recalc=True):
    """Fused device cycle (L <= 4): the dots ride on the SpMV / update passes."""
    S = _spmv_bytes(n, nnz, sval, s)
    b = 0
    for j in range(L):
        b += 3 * (j + 1) * n * s + S + n * s                      # u updates; u_{j+1} = A u_j with (u_{j+1}, r~)
        b += 3 * (j + 1) * n * s + 3 * n * s + S + (n * s if j + 1 < L else 0)   # r, x updates; r_{j+1} = A r_j
    b += (L + 1) * n * s                                          # Gram matrix in one pass
    b += (2 * L + 6) * n * s                                      # MR update of x, r_0, u_0
    if recalc:
        b += _spmv_bytes(n, nnz, sval, s, True) + n * s           # true residual with (r_0, r~), (r_0, r_0)
    else:
        b += n * s                                                # (r_0, r~), (r_0, r_0) on the MR pass
    return cycles * b + 3 * n * s + _spmv_bytes(n, nnz, sval, s, True)


def _gmres_bytes(m, n, nnz, sval, s, iters):
    """Fused device GMRES: per step the SpMV, two projection passes (multi-dot over
    V and w; multi-update reading V and w, writing w), the second with ||w||."""
    S = _spmv_bytes(n, nnz, sval, s)
    def step(j):
        return S + 2 * ((j + 2) * n * s + (j + 3) * n * s)
    cycles, rest = divmod(iters, m)
    restart = _spmv_bytes(n, nnz, sval, s, True) + n * s
    b = cycles * (sum(step(j) for j in range(m)) + restart + (m + 2) * n * s)
    if rest:
        b += sum(step(j) for j in range(rest)) + restart + (rest + 2) * n * s
    return b + restart


def _solve_bytes(p, n, nnz, iters, sval=8, s=8, recalc=True):
    if p.solver == 1:
        return _bicgstabl_bytes(p.basis, n, nnz, sval, s, -(-iters // p.basis), recalc)
    return _gmres_bytes(p.restart, n, nnz, sval, s, iters)


def _rel_diff(H, h, n, a, b):
    """||a - b|| / ||b|| with compensated FP64 norms (device vectors)."""
    import numpy as np
    e = H.DeviceArray(n, np.float64, h)
    H.call("hpb_axpbyc_f64", h.ptr, n, ctypes.c_double(1.0), a.ptr, ctypes.c_double(-1.0), b.ptr, e.ptr)
    o1, o2 = np.zeros(4), np.zeros(4)
    H.call("hpb_nrm2c_f64", h.ptr, n, e.ptr, o1.ctypes.data_as(ctypes.c_void_p))
    H.call("hpb_nrm2c_f64", h.ptr, n, b.ptr, o2.ctypes.data_as(ctypes.c_void_p))
    e.free()
    return float(o1[3] / o2[3])


def _run_solve(H, h, S, A, b, xs, p, ir=False, keep_x=False):
    import numpy as np
    x = H.DeviceArray(A.rows, np.float64, h)
    H.call("hpb_fill_f64", h.ptr, A.rows, ctypes.c_double(0.0), x.ptr)
    r = S.solve_ir(h, A, b, x, p, rel_tol=1e-12, max_outer=40) if ir else S.solve(h, A, b, x, p)
    # relative error against x*: ||x - x*|| / ||x*|| (compensated norms)
    e = H.DeviceArray(A.rows, np.float64, h)
    H.call("hpb_axpbyc_f64", h.ptr, A.rows, ctypes.c_double(1.0), x.ptr, ctypes.c_double(-1.0), xs.ptr, e.ptr)
    o1, o2 = np.zeros(4), np.zeros(4)
    H.call("hpb_nrm2c_f64", h.ptr, A.rows, e.ptr, o1.ctypes.data_as(ctypes.c_void_p))
    H.call("hpb_nrm2c_f64", h.ptr, A.rows, xs.ptr, o2.ctypes.data_as(ctypes.c_void_p))
    if keep_x:
        kept = x
    else:
        x.free()
    e.free()
    out = {k: (round(v, 6) if isinstance(v, float) and k == "seconds" else v)
           for k, v in r.items() if k in ("converged", "iters", "outer_iters", "true_rel_resid", "seconds")}
    if keep_x:
        out["_x"] = kept
    out["rel_err_vs_xstar"] = float(o1[3] / o2[3])
    n, nnz = A.rows, A.nnz
    if ir:
        # inner FP32 solves (FP32 values, int32 indices, FP32 vectors) + per sweep:
        # FP64 compensated residual, its norm, scaling to FP32, FP64 update; plus the
        # FP32 copy of the values
        inner = _solve_bytes(p, n, nnz, r["iters"], sval=4, s=4,
                             recalc=p.refinement != 1 and p.resid_recalc_freq != 0)
        sweep = _spmv_bytes(n, nnz, 8, 8, True) + 8 * n + 12 * n + 20 * n
        out["bytes_moved"] = int(inner + (r["outer_iters"] + 1) * sweep + 12 * nnz)
    else:
        out["bytes_moved"] = int(_solve_bytes(p, n, nnz, r["iters"], recalc=p.resid_recalc_freq != 0))
    return out


def cfg_krylov(H, h, log):
    import high_precision_blas1_amd.solvers as S
    log("krylov 128^3 convection-diffusion")
    A, b, xs = _solve_problem(H, h, S, 128, S.CONVDIFF7, 2)
    res = {"matrix": "128^3 first-order upwind convection-diffusion, v=(1,0.5,0.25), cell Peclet 100",
           "rows": A.rows, "nnz": A.nnz}
    G = dict(solver=S.GMRES, restart=50, rel_tol=1e-12, max_iters=20000)
    B = dict(solver=S.BICGSTABL, basis=4, rel_tol=1e-12, max_iters=20000)
    for name, p in (
            ("gmres50_naive", S.params(high_precision=0, **G)),
            ("gmres50_compensated", S.params(high_precision=1, **G)),
            ("gmres50_compensated_host_driven", S.params(high_precision=1, device_resident=0, **G)),
            ("gmres50_compensated_mgs_reference_schedule", S.params(high_precision=1, gram_schmidt=0, early_exit=0,
                                                                     **G)),
            ("bicgstab4_naive", S.params(high_precision=0, **B)),
            ("bicgstab4_compensated", S.params(high_precision=1, **B)),
            ("bicgstab4_compensated_host_driven", S.params(high_precision=1, device_resident=0, **B))):
        log(name)
        res[name] = _run_solve(H, h, S, A, b, xs, p)
    for a in (A, b, xs):
        a.free()
    return res


def _mixed_on(H, h, S, log, nx, coef, seed, plain, mixed):
    A, b, xs = _solve_problem(H, h, S, nx, coef, seed)
    res = {"rows": A.rows, "nnz": A.nnz}
    for k, p in plain.items():
        log(k)
        res[k] = _run_solve(H, h, S, A, b, xs, p, keep_x=True)
    for k, p in mixed.items():
        log(k)
        res[k] = _run_solve(H, h, S, A, b, xs, p, ir=True, keep_x=True)
    for m in mixed:
        # baselines: the FP64 solves of the same solver that reach 1e-12 true residual;
        # the speedup is priced against the fastest of them
        fam = m.split("_")[0]
        cands = [a for a in plain if a.split("_")[0] == fam]
        # the mixed-precision solution against the first FP64 solve of the same solver
        res[m]["rel_diff_vs_fp64_solve"] = _rel_diff(H, h, A.rows, res[m]["_x"], res[cands[0]]["_x"])
        ok = [a for a in cands if res[a]["converged"] and res[a]["true_rel_resid"] <= 1e-12]
        if ok and res[m]["converged"] and res[m]["true_rel_resid"] <= 1e-12:
            a = min(ok, key=lambda k: res[k]["seconds"])
            res["fp64_baseline_" + m] = a
            res["speedup_" + m] = round(res[a]["seconds"] / res[m]["seconds"], 3)
            res["bytes_ratio_" + m] = round(res[a]["bytes_moved"] / res[m]["bytes_moved"], 3)
    for k in list(plain) + list(mixed):
        res[k].pop("_x").free()
    for a in (A, b, xs):
        a.free()
    return res


def cfg_mixed(H, h, log, problems=("convdiff128", "aniso256")):
    """FP32-storage inner GMRES / BiCGStab(2) with compensated FP32 dots inside FP64
    refinement vs the plain FP64 compensated solvers, time to 1e-12 true residual.
    FP64 BiCGStab(2) runs twice: true residual every cycle (the library default) and
    never (the reference's default resid_recalc_freq 0); each speedup is priced
    against the fastest FP64 solve of the same solver that reaches 1e-12 true residual.
    GMRES: restarted refinement, one FP32 GMRES solve per sweep to the inner
    tolerance.  BiCGStab(2) two ways: reliable updates, one FP32 iteration that is
    never restarted, the FP64 solution and true residual updated each time the FP32
    residual has dropped 10x; and restarted refinement, one FP32 solve per sweep to
    the inner tolerance 1e-2."""
    import high_precision_blas1_amd.solvers as S
    res = {}
    G = dict(solver=S.GMRES, high_precision=1)
    B = dict(solver=S.BICGSTABL, basis=2, high_precision=1)
    if "convdiff128" in problems:
        log("mixed precision 128^3 convection-diffusion")
        res["convdiff128"] = _mixed_on(
            H, h, S, log, 128, S.CONVDIFF7, 2,
            {"gmres30_fp64": S.params(rel_tol=1e-12, max_iters=20000, restart=30, **G),
             "bicgstab2_fp64": S.params(rel_tol=1e-12, max_iters=20000, **B),
             "bicgstab2_fp64_recalc0": S.params(rel_tol=1e-12, max_iters=20000, resid_recalc_freq=0, **B)},
            {"gmres30_fp32_refined": S.params(rel_tol=1e-4, max_iters=5000, restart=30, **G),
             "bicgstab2_fp32_reliable": S.params(rel_tol=0.1, max_iters=20000, refinement=1, **B),
             "bicgstab2_fp32_refined": S.params(rel_tol=1e-2, max_iters=20000, refinement=0, resid_recalc_freq=0,
                                                **B)})
        res["convdiff128"]["matrix"] = "128^3 convection-diffusion (as krylov)"
    if "aniso256" in problems:
        log("mixed precision 256^3 anisotropic diffusion")
        res["aniso256"] = _mixed_on(
            H, h, S, log, 256, S.ANISO7, 2,
            {"gmres30_fp64": S.params(rel_tol=1e-12, max_iters=20000, restart=30, **G),
             "bicgstab2_fp64": S.params(rel_tol=1e-12, max_iters=20000, **B),
             "bicgstab2_fp64_recalc0": S.params(rel_tol=1e-12, max_iters=20000, resid_recalc_freq=0, **B)},
            {"gmres30_fp32_refined": S.params(rel_tol=1e-2, max_iters=5000, restart=30, compensated_spmv=1, **G),
             "bicgstab2_fp32_reliable": S.params(rel_tol=0.1, max_iters=40000, refinement=1, **B),
             "bicgstab2_fp32_refined": S.params(rel_tol=1e-2, max_iters=40000, refinement=0, resid_recalc_freq=0,
                                                **B)})
        res["aniso256"]["matrix"] = "256^3 7-point anisotropic diffusion, k_z / k_x = 1e4, Dirichlet"
    return res


def cfg_multidot(H, h, log, ks=None):
    """Fused compensated multi-dot: k = 1..64 at n = 2^24, and lengths 2^10..2^24 for
    k in {1, 8, 32, 64}.  ks (a profiling aid): only these k at n = 2^24, no sweep."""
    import numpy as np
    log("multidot")
    nmax = 1 << 24
    kmax = 64
    V = [H.DeviceArray(nmax, np.float64, h) for _ in range(kmax)]
    w = H.DeviceArray(nmax, np.float64, h)
    for i, a in enumerate(V + [w]):
        H.call("hpb_random_uniform_f64", h.ptr, nmax, 5, i, ctypes.c_double(-1.0), ctypes.c_double(1.0), a.ptr)
    d = H.DeviceArray(4 * kmax, np.float64, h)
    L = H.lib()
    rows = []
    for k in ks or (1, 2, 4, 8, 16, 32, 50, 64):
        ptrs = (ctypes.c_void_p * k)(*[a.addr for a in V[:k]])
        ms = timed(h, L.hpb_multidot_f64_async, (h.ptr, k, ptrs, nmax, w.ptr, 1, d.ptr), 50, 10)
        nbytes = 8.0 * nmax * (k + 1)
        rows.append({"k": k, "ms": round(ms, 4), "gbs_algorithmic": round(nbytes / ms / 1e6, 1),
                     "separate_dot2_bytes_ratio": round(2.0 * k / (k + 1), 3)})
    sweep = []
    for e in () if ks else (10, 13, 16, 19, 22, 24):
        n = 1 << e
        for k in (1, 8, 32, 64):
            ptrs = (ctypes.c_void_p * k)(*[a.addr for a in V[:k]])
            ms = timed(h, L.hpb_multidot_f64_async, (h.ptr, k, ptrs, n, w.ptr, 1, d.ptr), 20, 3)
            sweep.append({"n": n, "k": k, "us": round(1e3 * ms, 2),
                          "gbs_algorithmic": round(8.0 * n * (k + 1) / ms / 1e6, 1)})
    for a in V + [w, d]:
        a.free()
    return {"n": nmax, "rows": rows, "length_sweep": sweep}


BATCH_LAUNCH = {"graph": 1, "direct": 0, "persistent": 2}


def cfg_batch(H, h, log, nsys=512, recalc=0, use_graph=1):
    """512 independent 2-D 5-point diffusion systems, sides uniform in 64..1024 (numpy
    default_rng(3)), log-normal coefficients (sigma 1, seed 3), x* uniform in [-1, 1]
    (seed 7), BiCGStab(2) with Jacobi right preconditioning to 1e-10 true relative
    residual per system; use_graph 1: one captured hipGraph per cycle over all live
    systems, 0: the same kernels launched directly, 2: every cycle in one persistent
    cooperative launch.
    recalc: the true residual replaces the recursive one every recalc cycles (0:
    never; convergence is confirmed on the true residual either way)."""
    import numpy as np
    r = np.random.default_rng(3)
    sides = [int(v) for v in r.integers(64, 1025, nsys)]
    arr = (ctypes.c_int * nsys)(*sides)
    b = ctypes.c_void_p()
    log("batch %d systems" % nsys)
    t0 = time.perf_counter()
    H.call("hpb_batch_create_diffusion2d", h.ptr, nsys, arr, ctypes.c_double(1.0), 3, 7, ctypes.byref(b))
    t_build = time.perf_counter() - t0
    H.call("hpb_batch_set_jacobi", b, 1)
    H.call("hpb_batch_set_resid_recalc_freq", b, recalc)
    rows, nnz, nt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    H.call("hpb_batch_info", b, ctypes.byref(rows), ctypes.byref(nnz), ctypes.byref(nt))
    it = np.zeros(nsys, dtype=np.int32)
    rr = np.zeros(nsys)
    cv = np.zeros(nsys, dtype=np.int32)
    secs, cyc = ctypes.c_double(), ctypes.c_int()
    H.call("hpb_batch_solve_bicgstab2", b, ctypes.c_double(1e-10), 40000, use_graph, 50,
           it.ctypes.data_as(ctypes.c_void_p),
           rr.ctypes.data_as(ctypes.c_void_p), cv.ctypes.data_as(ctypes.c_void_p), ctypes.byref(secs),
           ctypes.byref(cyc))
    H.call("hpb_batch_destroy", b)
    # lower bound on the bytes moved: each system's iterations priced with the fused
    # single-system BiCGStab(2) cycle model (the batch cycle moves at least that much)
    nbytes = sum(_bicgstabl_bytes(2, m * m, 5 * m * m - 4 * m, 8, 8, (int(k) + 1) // 2, recalc=False)
                 for m, k in zip(sides, it))
    return {"systems": nsys, "sides_min_max": [min(sides), max(sides)], "rows": rows.value, "nnz": nnz.value,
            "bytes_model_lower_bound": int(nbytes), "gbs_lower_bound": round(nbytes / secs.value / 1e9, 1),
            "tiles": nt.value, "preconditioner": "jacobi (right)", "resid_recalc_freq": recalc,
            "build_s": round(t_build, 3),
            "solve_s": round(secs.value, 3), "cycles": cyc.value, "converged": int(cv.sum()),
            "iters_min_mean_max": [int(it.min()), round(float(it.mean()), 1), int(it.max())],
            "max_true_rel_resid": float(rr.max()),
            "launch": {v: k for k, v in BATCH_LAUNCH.items()}[use_graph]}


def cfg_next(H, h, log):
    """The components after the hot path, measured on the 128^3 convection-diffusion
    matrix (2 097 152 rows, 14 581 760 non-zeros) unless stated:
    F1 MatrixMarket write + read of the 64^3 matrix of the same operator (262 144 rows,
       host, a temporary file; the round trip must reproduce every array bit for bit);
    F3 level-scheduled triangular solves (lower with unit diagonal, upper with the stored
       diagonal), the ILU(0) factorisation, and ILU(0)-preconditioned BiCGStab(2) and
       GMRES(30) to 1e-12 against the unpreconditioned device solvers;
    F4 plain BiCGStab to 1e-12, and the complex Dot2 x^H y at n = 2^26 complex numbers
       (2 GiB read per call).
    F2, the exact dot, is timed in the condition sweep."""
    import tempfile
    import numpy as np
    import high_precision_blas1_amd.solvers as S
    L = H.lib()
    out = {}
    # F1
    log("next: MatrixMarket round trip")
    A = S.DeviceCSR.stencil7(h, 64, 64, 64, S.CONVDIFF7)
    n, nnz = A.rows, A.nnz
    rp, ci, va = A.rp.numpy(), A.col.numpy(), A.val.numpy()
    A.free()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "convdiff128.mtx").encode()
        t0 = time.perf_counter()
        H.call("hpb_mm_write_csr", path, n, n, nnz, rp.ctypes.data_as(ctypes.c_void_p),
               ci.ctypes.data_as(ctypes.c_void_p), va.ctypes.data_as(ctypes.c_void_p))
        t_w = time.perf_counter() - t0
        size = os.path.getsize(path)
        rows_, cols_, nnz_ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        prp, pci, pva = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        t0 = time.perf_counter()
        H.call("hpb_mm_read_csr", path, ctypes.byref(rows_), ctypes.byref(cols_), ctypes.byref(nnz_),
               ctypes.byref(prp), ctypes.byref(pci), ctypes.byref(pva))
        t_r = time.perf_counter() - t0
    same = (rows_.value == n and nnz_.value == nnz and
            np.ctypeslib.as_array(ctypes.cast(prp, ctypes.POINTER(ctypes.c_int)), (n + 1,)).tobytes() == rp.tobytes()
            and np.ctypeslib.as_array(ctypes.cast(pci, ctypes.POINTER(ctypes.c_int)), (nnz,)).tobytes() == ci.tobytes()
            and np.ctypeslib.as_array(ctypes.cast(pva, ctypes.POINTER(ctypes.c_double)), (nnz,)).tobytes()
            == va.tobytes())
    for q in (prp, pci, pva):
        L.hpb_mm_free(q)
    out["matrix_market"] = {"rows": n, "nnz": nnz, "file_mb": round(size / 1e6, 1), "write_s": round(t_w, 3),
                            "read_s": round(t_r, 3), "read_mb_per_s": round(size / t_r / 1e6, 1),
                            "round_trip_bit_equal": bool(same)}
    A, b, xs = _solve_problem(H, h, S, 128, S.CONVDIFF7, 2)
    n, nnz = A.rows, A.nnz
    # F3
    log("next: triangular solves and ILU(0)")
    tri = {}
    x1 = H.DeviceArray(n, np.float64, h)
    for uplo, name in ((0, "lower_unit"), (1, "upper")):
        t = ctypes.c_void_p()
        t0 = time.perf_counter()
        H.call("hpb_csr_trsv_analysis", h.ptr, uplo, n, nnz, A.rp.ptr, A.col.ptr, ctypes.byref(t))
        t_an = time.perf_counter() - t0
        nl = ctypes.c_int()
        H.call("hpb_csr_trsv_levels", t, ctypes.byref(nl))
        for comp in (0, 1):
            ms = timed(h, L.hpb_csr_trsv_f64, (h.ptr, t, A.val.ptr, ctypes.c_double(1.0), b.ptr, x1.ptr, comp), 10, 2)
            tri["%s%s_ms" % (name, "_compensated" if comp else "")] = round(ms, 4)
        tri[name + "_levels"] = nl.value
        tri[name + "_analysis_s"] = round(t_an, 3)
        if uplo == 0:
            lu = H.DeviceArray(nnz, np.float64, h)
            H.call("hpb_memcpy_d2d", h.ptr, lu.ptr, A.val.ptr, nnz * 8)
            h.sync()
            h.timer_start()
            H.call("hpb_csr_ilu0_f64", h.ptr, t, lu.ptr)
            tri["ilu0_factor_ms"] = round(h.timer_stop(), 3)
            lu.free()
        H.call("hpb_trsv_destroy", t)
    x1.free()
    out["triangular"] = tri
    sol = {}
    G = dict(solver=S.GMRES, restart=30, high_precision=1)
    B = dict(solver=S.BICGSTABL, basis=2, high_precision=1)
    for k, p in (("bicgstab2", S.params(rel_tol=1e-12, max_iters=20000, **B)),
                 ("bicgstab2_ilu0", S.params(rel_tol=1e-12, max_iters=20000, preconditioner=S.PRECOND_ILU0, **B)),
                 ("gmres30", S.params(rel_tol=1e-12, max_iters=20000, **G)),
                 ("gmres30_ilu0", S.params(rel_tol=1e-12, max_iters=20000, preconditioner=S.PRECOND_ILU0, **G)),
                 ("bicgstab_plain", S.params(rel_tol=1e-12, max_iters=20000, solver=S.BICGSTAB, high_precision=1))):
        log("next: " + k)
        r = _run_solve(H, h, S, A, b, xs, p)
        sol[k] = {q: r[q] for q in ("converged", "iters", "seconds", "true_rel_resid", "rel_err_vs_xstar")}
    out["solvers"] = sol
    for a in (A, b, xs):
        a.free()
    # F4 complex dot
    log("next: complex Dot2")
    nc = 1 << 26
    zx = H.DeviceArray(2 * nc, np.float64, h)
    zy = H.DeviceArray(2 * nc, np.float64, h)
    H.call("hpb_random_uniform_f64", h.ptr, 2 * nc, 11, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), zx.ptr)
    H.call("hpb_random_uniform_f64", h.ptr, 2 * nc, 11, 1, ctypes.c_double(-1.0), ctypes.c_double(1.0), zy.ptr)
    d6 = H.DeviceArray(8, np.float64, h)
    ms = timed(h, L.hpb_zdot2_f64_async, (h.ptr, nc, zx.ptr, zy.ptr, d6.ptr), 50, 20)
    out["complex_dot2"] = {"n_complex": nc, "ms": round(ms, 4), "gbs_algorithmic": round(32.0 * nc / ms / 1e6, 1)}
    for a in (zx, zy, d6):
        a.free()
    return out


SUITES = ("cpu_plain", "blas1", "cond_sweep", "spmv", "krylov", "mixed", "mixed_convdiff128", "mixed_aniso256",
          "multidot", "batch", "next")


def run_suite(H, h, x, y, n, args, log, no_cpu=False, only=None):
    """The configurations after the headline; `only` restricts them to a subset
    (a profiling aid: one solver workload under rocprofv3)."""
    want = set(only) if only else set(SUITES)
    res = {}
    if "cpu_plain" in want and not no_cpu:
        log("cpu plain config")
        res["cpu_plain"] = cfg_cpu_plain()
    if "blas1" in want:
        res["blas1"] = cfg_blas1(H, h, x, y, n)
    if "cond_sweep" in want:
        res["cond_sweep"] = cfg_cond_sweep(H, h, n, log)
    if "spmv" in want:
        res["spmv"] = cfg_spmv(H, h, log)
    if "krylov" in want:
        res["krylov"] = cfg_krylov(H, h, log)
    probs = [p for p in ("convdiff128", "aniso256") if "mixed" in want or "mixed_" + p in want]
    if probs:
        res["mixed"] = cfg_mixed(H, h, log, probs)
    if "multidot" in want:
        ks = [int(v) for v in args.multidot_k.split(",")] if args.multidot_k else None
        res["multidot"] = cfg_multidot(H, h, log, ks)
    if "batch" in want:
        mode = "direct" if args.batch_direct else args.batch_launch
        res["batch"] = cfg_batch(H, h, log, use_graph=BATCH_LAUNCH[mode])
    if "next" in want:
        res["next"] = cfg_next(H, h, log)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1 << 27)
    ap.add_argument("--cond", type=float, default=1e16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--op", default="dot2", choices=("dot2", "copy", "nrm2c", "exact"),
                    help="kernel timed for the headline (copy: the plain copy used to calibrate HBM counters)")
    ap.add_argument("--batch-direct", action="store_true",
                    help="batch suite with directly launched cycles instead of graph replays (profiling aid)")
    ap.add_argument("--multidot-k", default=None,
                    help="multidot suite: only these comma-separated k at n = 2^24 (profiling aid)")
    ap.add_argument("--batch-launch", default="graph", choices=tuple(BATCH_LAUNCH),
                    help="how the batch suite runs its cycles: graph replays, direct launches, or one persistent "
                         "cooperative launch")
    ap.add_argument("--suite", default=None,
                    help="comma-separated subset of the configurations after the headline: " + ",".join(SUITES))
    args = ap.parse_args()
    only = None
    if args.suite:
        only = [v for v in args.suite.split(",") if v]
        bad = [v for v in only if v not in SUITES]
        if bad:
            print("bench.py: unknown --suite entries %s (choose from %s)" % (bad, ",".join(SUITES)), file=sys.stderr)
            sys.exit(2)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the hot path is measured on one MI355X; "
              "--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import high_precision_blas1_amd as H
    h = H.Handle()
    L = H.lib()
    n = args.n
    log("generating GenDot pair n=%d kappa=%g" % (n, args.cond))
    x = H.DeviceArray(n, np.float64, h)
    y = H.DeviceArray(n, np.float64, h)
    t_gen = time.perf_counter()
    exact, dd = gendot(H, h, n, args.cond, args.seed, x, y)
    t_gen = time.perf_counter() - t_gen
    dout = H.DeviceArray(4, np.float64, h)
    if args.op != "dot2":
        # profiling aid: time another kernel over the same vectors and stop
        if args.op == "copy":
            ms = timed(h, L.hpb_copy_f64, (h.ptr, n, x.ptr, y.ptr), args.steps, args.warmup)
            nbytes = 16.0 * n
        elif args.op == "exact":
            ms = timed(h, L.hpb_exact_dot_f64_async, (h.ptr, n, x.ptr, y.ptr, dout.ptr), args.steps, args.warmup)
            nbytes = 16.0 * n
        else:
            ms = timed(h, L.hpb_nrm2c_f64_async, (h.ptr, n, x.ptr, dout.ptr), args.steps, args.warmup)
            nbytes = 8.0 * n
        print(json.dumps({"op": args.op, "n": n, "ms": round(ms, 5), "gbs": round(nbytes / ms / 1e6, 2),
                          "algorithmic_bytes": int(nbytes)}), flush=True)
        return
    # measured ceiling, same process, same vectors: the naive dot (same launch shape and
    # single-launch reduction as Dot2, one fma per element: the read stream alone) and
    # a copy; the naive dot first runs for WARM_S seconds so that the device is at its
    # sustained clocks when the timed windows start (a cold device times its ramp)
    log("warming the device for %.1f s with the naive dot" % WARM_S)
    t_end = time.perf_counter() + WARM_S
    nwarm = 0
    while time.perf_counter() < t_end:
        for _ in range(50):
            L.hpb_dot_f64_async(h.ptr, n, x.ptr, y.ptr, dout.ptr)
        h.sync()
        nwarm += 50
    ms_naive = timed(h, L.hpb_dot_f64_async, (h.ptr, n, x.ptr, y.ptr, dout.ptr), args.steps, args.warmup)
    # in place (x -> x): the benchmark vectors keep their values
    ms_copy = timed(h, L.hpb_copy_f64, (h.ptr, n, x.ptr, x.ptr), args.steps, args.warmup)
    log("timing %d Dot2 steps after %d warmup" % (args.steps, args.warmup))
    ms = timed(h, L.hpb_dot2_f64_async, (h.ptr, n, x.ptr, y.ptr, dout.ptr), args.steps, args.warmup)
    res = dout.numpy()
    nbytes = 16.0 * n
    gbs = nbytes / (ms * 1e-3) / 1e9
    cfg = h.reduce_config("float64", n, True)
    traffic = profiled_traffic(n, cfg, h.get_tuning(0, np.float64)[3] & 1)
    out = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device GenDot pair, kappa %g, seed %d)" % (args.cond, args.seed),
        "config": {"workload": "FP64 compensated Dot2 (Ogita-Rump-Oishi), N=2^%d per vector, device-resident"
                               % (n.bit_length() - 1), "n": n,
                   "launch": {"block": cfg[0], "unroll": cfg[1], "grid": cfg[2], "vector_width": cfg[3]}},
        "hbm_peak_frac": round(gbs / HBM_PEAK_GBS, 4),
        "ceiling": {"dot_naive_gbs": round(16.0 * n / (ms_naive * 1e-3) / 1e9, 2),
                    "copy_gbs": round(16.0 * n / (ms_copy * 1e-3) / 1e9, 2),
                    "dot2_frac_of_dot_naive": round(ms_naive / ms, 4),
                    "warmup": "naive dot for %.1f s (%d launches) before the timed windows" % (WARM_S, nwarm),
                    "note": "naive dot: same launch shape and reduction as Dot2, one fma per element (the read "
                            "stream); copy: reads and rewrites x in place (16 B per element)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                     "kernel": "hpb::k_reduce<double, DOT2> (hpb_dot2_f64)",
                     "algorithmic_bytes_per_launch": int(nbytes)},
        "result": {"dot2": float(res[2]), "exact": float(exact), "rel_err": rel_err(res[2], exact),
                   "cond_achieved": float(dd[2]), "generator_s": round(t_gen, 3)},
    }
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(H, h, x, y, n)
    if not args.headline_only:
        out["configs"] = run_suite(H, h, x, y, n, args, log, no_cpu=args.no_cpu, only=only)
    x.free()
    y.free()
    dout.free()
    h.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
