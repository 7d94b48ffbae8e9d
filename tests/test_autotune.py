"""The autotuner (high_precision_blas1_amd/autotune.py): its candidate shapes are
all accepted by hpb_set_tuning's rules, the shipped MI355X table is complete and
valid, `apply` loads it into a handle (and refuses a table measured on another
CU count); on the GPU a short tuning run produces a table whose shapes give the
library-order results bit for bit."""
import json

import numpy as np
import pytest

import high_precision_blas1_amd as H
import high_precision_blas1_amd.autotune as AT
import oracle as O
from conftest import say

KEYS = ["%s_%s" % (k, d) for k in ("reduce", "reduce1", "map", "multi") for d in ("f64", "f32")]


def _valid(block, unroll, grid, flags, kind=None):
    # the rules of hpb_set_tuning (hpb_capi.cpp)
    if kind == "multi" and not (block == 256 and unroll in (4, 8)):
        return False
    return block in (256, 512) and unroll in (1, 2, 4, 8) and 1 <= grid <= 4096 and flags in (0, 1)


@pytest.mark.parametrize("kind", ["reduce", "reduce1", "map", "multi"])
@pytest.mark.parametrize("ncu", [1, 80, 256, 304])
def test_candidates_are_valid(kind, ncu):
    c = AT.candidates(kind, ncu)
    assert len(c) == len(set(c)) > 0
    for cand in c:
        assert _valid(*cand, kind=kind), cand


def test_shipped_table():
    with open(AT.DEFAULT_TABLE) as f:
        t = json.load(f)
    assert t["num_cu"] == 256 and t["n"] == 1 << 27
    for k in KEYS:
        block, unroll, grid, flags, gbs = t[k]
        assert _valid(block, unroll, grid, flags, kind=k.rsplit("_", 1)[0]), k
        assert 0 < gbs < 8000.0, k   # below the HBM peak


class _FakeHandle:
    def __init__(self, ncu):
        self.ncu = ncu
        self.calls = []

    def num_cu(self):
        return self.ncu

    def set_tuning(self, kind, dtype, block, unroll, grid, flags=0):
        self.calls.append((kind, np.dtype(dtype).name, block, unroll, grid, flags))


def test_apply_loads_every_entry(tmp_path):
    h = _FakeHandle(256)
    t = AT.apply(h)
    assert t is not None and len(h.calls) == len(KEYS)
    with open(AT.DEFAULT_TABLE) as f:
        ref = json.load(f)
    for kind, dname, block, unroll, grid, flags in h.calls:
        name = {0: "reduce", 1: "map", 2: "reduce1", 3: "multi"}[kind] + "_" + {"float64": "f64", "float32": "f32"}[dname]
        assert [block, unroll, grid, flags] == ref[name][:4]
    # a table measured on another CU count is not applied
    other = dict(ref, num_cu=304)
    p = tmp_path / "t.json"
    p.write_text(json.dumps(other))
    h2 = _FakeHandle(256)
    assert AT.apply(h2, str(p)) is None and h2.calls == []
    assert AT.apply(h2, str(tmp_path / "missing.json")) is None


@pytest.mark.gpu
def test_tune_then_parity(handle):
    n = (1 << 16) + 3
    say("autotune n=%d" % n)
    table = AT.tune(handle, n, reps=2, verbose=False)   # restores the handle's shapes
    for k in KEYS:
        assert _valid(*table[k][:4], kind=k.rsplit("_", 1)[0]) and table[k][4] > 0
    h2 = H.Handle()
    for k in KEYS:
        kind, dname = k.rsplit("_", 1)
        h2.set_tuning(AT.KINDS[kind], AT.DTYPES[dname], *table[k][:4])
    r = np.random.default_rng(5)
    for dt in (np.float64, np.float32):
        x = r.standard_normal(n).astype(dt)
        y = r.standard_normal(n).astype(dt)
        dx = H.DeviceArray.from_numpy(x, h2)
        dy = H.DeviceArray.from_numpy(y, h2)
        for op, binary in (("dot2", True), ("nrm2c", False)):
            got = H.reduce(op, h2, dx, dy if binary else None)
            cfg = h2.reduce_config(dt, n, True, binary)
            emu = O.hpb_order_reduce(op, x, y if binary else None, cfg)
            assert got[2].tobytes() == emu[2].tobytes(), (op, np.dtype(dt).name, cfg)
        dx.free()
        dy.free()
    h2.close()
