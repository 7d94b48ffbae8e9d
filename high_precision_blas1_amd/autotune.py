"""Autotuner for the launch shapes of the reduction and elementwise kernels.

The reference fixes BLOCK_SIZE = 1024 and two elements per thread at compile
time (source/common/ogita/gpu_reduction_ogita.h:36, gpu_reduction_ogita_impl.cuh:683-694).
Here the shape is a runtime property of the handle (hpb_set_tuning): threads per
block, 16-byte vectors per thread per tile, the cap on the grid, and optional
nontemporal loads (and stores, for the elementwise kernels).  For the fused
multi-dot of Gram-Schmidt (kind "multi") the second field is the number of basis
columns per chunk (4 or 8) and the grid is counted over all chunks; it is timed
with 16 basis vectors of length min(n, 2^24).  This module times candidate shapes with HIP events on the
handle's stream, keeps the fastest per (kind, dtype) and stores the table as
JSON; ``apply`` loads a table into a handle.

    python -m high_precision_blas1_amd.autotune --n 134217728 --out tuning.json

The table shipped next to this file (tuning_gfx950.json) holds the shapes
measured on MI355X; handles use it when ``apply`` is called without a path.
Note that the summation order of a reduction depends on its launch shape, so
results are reproducible for a fixed table (the table is part of the
reproducibility contract).
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

from . import DeviceArray, Handle, call, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning_gfx950.json")

KINDS = {"reduce": 0, "map": 1, "reduce1": 2, "multi": 3}
MULTI_K = 16              # basis vectors of the multi-dot timed by the tuner
MULTI_N_MAX = 1 << 24     # its vector length (at most)
DTYPES = {"f64": np.float64, "f32": np.float32}


MAX_GRID = 4096   # hpb_set_tuning's limit on grid_max


def candidates(kind, num_cu):
    """Launch shapes (block, unroll, grid_max, flags) tried for a kernel family."""
    if kind == "multi":
        # block 256; unroll = basis columns per chunk; grid_max = blocks over all chunks
        return [(256, kc, min(MAX_GRID, max(1, int(num_cu * m))), f)
                for kc in (4, 8) for m in (1, 2, 4, 8, 16) for f in (0, 1)]
    mults = (0.5, 0.75, 1, 1.5, 2, 4) if kind in ("reduce", "reduce1") else (1, 2, 4, 8, 16)
    out = []
    for block in (256, 512):
        for unroll in ((2, 4, 8) if kind in ("reduce", "reduce1") else (1, 2, 4)):
            if kind == "map" and block == 512 and unroll == 1:
                continue
            for mult in mults:
                for flags in (0, 1):
                    cand = (block, unroll, min(MAX_GRID, max(1, int(num_cu * mult))), flags)
                    if cand not in out:
                        out.append(cand)
    return out


def _time(h, fn, args, reps):
    for _ in range(3):
        fn(*args)
    h.sync()
    h.timer_start()
    for _ in range(reps):
        fn(*args)
    return h.timer_stop() / reps


def tune(h, n, reps=20, dtypes=("f64", "f32"), verbose=True):
    """Return {"reduce_f64": [block, unroll, grid_max, flags, GB/s], ...}."""
    L = lib()
    ncu = h.num_cu()
    table = {}
    for dname in dtypes:
        dt = DTYPES[dname]
        isz = np.dtype(dt).itemsize
        x = DeviceArray(n, dt, h)
        y = DeviceArray(n, dt, h)
        z = DeviceArray(n, dt, h)
        c = ctypes.c_double if dt == np.float64 else ctypes.c_float
        for a, sid in ((x, 0), (y, 1), (z, 2)):
            call("hpb_random_uniform_%s" % dname, h.ptr, n, 7, sid, c(-1), c(1), a.ptr)
        dout = DeviceArray(4 * MULTI_K, dt, h)
        nm = min(n, MULTI_N_MAX)
        basis = [DeviceArray(nm, dt, h) for _ in range(MULTI_K)]
        for i, a in enumerate(basis):
            call("hpb_random_uniform_%s" % dname, h.ptr, nm, 8, i, c(-1), c(1), a.ptr)
        ptrs = (ctypes.c_void_p * MULTI_K)(*[a.addr for a in basis])
        saved = {k: h.get_tuning(v, dt) for k, v in KINDS.items()}
        for kind in ("reduce", "reduce1", "map", "multi"):
            best = None
            for cand in candidates(kind, ncu):
                h.set_tuning(KINDS[kind], dt, *cand)
                if kind == "reduce":
                    fn = getattr(L, "hpb_dot2_%s_async" % dname)
                    ms = _time(h, fn, (h.ptr, n, x.ptr, y.ptr, dout.ptr), reps)
                    gbs = 2 * isz * n / ms / 1e6
                elif kind == "reduce1":
                    fn = getattr(L, "hpb_nrm2c_%s_async" % dname)
                    ms = _time(h, fn, (h.ptr, n, x.ptr, dout.ptr), reps)
                    gbs = isz * n / ms / 1e6
                elif kind == "multi":
                    if dname == "f64":
                        ms = _time(h, L.hpb_multidot_f64_async, (h.ptr, MULTI_K, ptrs, nm, x.ptr, 1, dout.ptr), reps)
                    else:
                        outh = np.zeros(4 * MULTI_K, dtype=np.float32)
                        ms = _time(h, L.hpb_multidot_f32, (h.ptr, MULTI_K, ptrs, nm, x.ptr, 1,
                                                           outh.ctypes.data_as(ctypes.c_void_p)), reps)
                    gbs = (MULTI_K + 1) * isz * nm / ms / 1e6
                else:
                    fn = getattr(L, "hpb_axpy_%s" % dname)
                    ms = _time(h, fn, (h.ptr, n, c(1e-3), x.ptr, z.ptr), reps)
                    gbs = 3 * isz * n / ms / 1e6
                if verbose:
                    print("tune", kind, dname, cand, "%.1f GB/s" % gbs, flush=True)
                if best is None or gbs > best[-1]:
                    best = list(cand) + [round(gbs, 1)]
            table["%s_%s" % (kind, dname)] = best
            h.set_tuning(KINDS[kind], dt, *saved[kind])
        for a in [x, y, z, dout] + basis:
            a.free()
    table["n"] = n
    table["num_cu"] = ncu
    return table


def apply(h, path=None):
    """Load a tuning table into handle h (returns the table, or None if absent)."""
    path = path or DEFAULT_TABLE
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    if table.get("num_cu") not in (None, h.num_cu()):
        return None
    for key, val in table.items():
        if "_" not in key or not isinstance(val, list):
            continue
        kind, dname = key.rsplit("_", 1)
        h.set_tuning(KINDS[kind], DTYPES[dname], int(val[0]), int(val[1]), int(val[2]), int(val[3]))
    return table


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 27)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    with Handle() as h:
        table = tune(h, args.n, args.reps)
    print(json.dumps(table), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
