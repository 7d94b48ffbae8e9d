#!/usr/bin/env python3
"""Benchmark of the compensated BLAS1 hot path on one MI355X.

Default workload (one "step" = one pass of the hot path over one batch):
  FP64 Ogita-Rump-Oishi Dot2 of two resident vectors of N = 2^27 elements
  (2 GiB read per step, far above the 256 MB MALL), GenDot-style ill-conditioned
  pair generated on the device (kappa 1e16, seed 0).

Printed: one JSON line with the whole-job throughput (GB/s of algorithmic bytes,
16 B per element), the roofline block of the Dot2 kernel (achieved vs the
8 TB/s HBM3E peak, timed with HIP events on the handle's stream), and the CPU
baseline: the reference's own CPU Dot2 (oracle/_ref, compiled from
source/common/threaded_reduction.h and dot_product.h) timed on this host on a
bounded sample of the same vectors.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Dot2 GB/s & %HBM peak; Krylov time-to-1e-12, mixed+refined vs FP64 plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(H, h, n, cond, seed):
    """Device vectors x, y of length n; GenDot pair when the generator exists."""
    x = H.DeviceArray(n, "float64", h)
    y = H.DeviceArray(n, "float64", h)
    lib = H.lib()
    if hasattr(lib, "hpb_gendot_f64"):
        import numpy as np
        limbs = np.zeros(4, dtype=np.uint64)
        dd = np.zeros(3)
        H.call("hpb_gendot_f64", h.ptr, n, ctypes.c_double(float(np.log2(cond))), seed, n // 2, x.ptr, y.ptr,
               limbs.ctypes.data_as(ctypes.c_void_p), dd.ctypes.data_as(ctypes.c_void_p))
        return x, y, {"generator": "gendot", "cond": cond, "achieved_cond": float(dd[2]), "exact_dot": float(dd[0])}
    H.call("hpb_random_uniform_f64", h.ptr, n, seed, 0, ctypes.c_double(-1.0), ctypes.c_double(1.0), x.ptr)
    H.call("hpb_random_uniform_f64", h.ptr, n, seed, 1, ctypes.c_double(-1.0), ctypes.c_double(1.0), y.ptr)
    return x, y, {"generator": "uniform(-1,1)"}


def time_dot2(H, h, x, y, n, steps, warmup):
    dout = H.DeviceArray(4, "float64", h)
    fn = H.lib().hpb_dot2_f64_async
    for _ in range(warmup):
        rc = fn(h.ptr, n, x.ptr, y.ptr, dout.ptr)
        if rc:
            raise RuntimeError(H.lib().hpb_last_error().decode())
    h.sync()
    h.timer_start()
    for _ in range(steps):
        fn(h.ptr, n, x.ptr, y.ptr, dout.ptr)
    ms = h.timer_stop()
    res = dout.numpy()
    dout.free()
    return ms, res


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

    def timed(f, budget):
        ts = []
        t_end = time.perf_counter() + budget
        while len(ts) < 4 or (time.perf_counter() < t_end and len(ts) < 200):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        core = ts[1:-1]  # drop min and max (test_vector_operations.cpp:44-49)
        return sum(core) / len(core), len(ts)

    t_thr, reps_thr = timed(fthr, budget_s * 0.6)
    t_seq, reps_seq = timed(fseq, budget_s * 0.4)
    nbytes = 16.0 * sample
    return {
        "value": round(nbytes / t_thr / 1e9, 3),
        "unit": "GB/s",
        "cores": cores,
        "kind": kind,
        "sample": "FP64 Dot2 over the first %d elements of the benchmark vectors (%d MiB), threaded_reduction "
                  "with %d threads, mean of %d runs without min/max" % (sample, int(nbytes) >> 20, cores, reps_thr),
        "single_thread_gbs": round(nbytes / t_seq / 1e9, 3),
        "single_thread_runs": reps_seq,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1 << 27)
    ap.add_argument("--cond", type=float, default=1e16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the hot path is measured on one MI355X; "
              "--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import high_precision_blas1_amd as H
    h = H.Handle()
    n = args.n
    log("bench: generating inputs n=%d" % n)
    x, y, gen = make_inputs(H, h, n, args.cond, args.seed)
    h.sync()
    log("bench: timing %d steps after %d warmup" % (args.steps, args.warmup))
    ms, res = time_dot2(H, h, x, y, n, args.steps, args.warmup)
    per = ms / args.steps
    nbytes = 16.0 * n
    gbs = nbytes / (per * 1e-3) / 1e9
    cfg = h.reduce_config("float64", n, True)
    out = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FP64 compensated Dot2 (Ogita-Rump-Oishi), N=2^%d per vector, device-resident"
                               % (n.bit_length() - 1), "n": n, "input": gen,
                   "launch": {"block": cfg[0], "unroll": cfg[1], "grid": cfg[2], "vector_width": cfg[3]}},
        "hbm_peak_frac": round(gbs / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_reduce<double,DOT2> (hpb_dot2_f64)",
                     "algorithmic_bytes_per_launch": int(nbytes)},
        "result": {"s": float(res[0]), "c": float(res[1]), "dot": float(res[2])},
    }
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(H, h, x, y, n)
    x.free()
    y.free()
    h.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
