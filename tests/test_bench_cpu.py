"""bench.py pieces that need no GPU: the committed Dot2 traffic profile parses for the
launch shape the tuned handle uses, so the bench line carries a measured traffic."""
import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_profile_matches_tuned_dot2_launch():
    bench = _bench()
    tune = json.load(open(os.path.join(ROOT, "high_precision_blas1_amd", "tuning_gfx950.json")))
    block, unroll, grid, flags = tune["reduce_f64"][:4]
    n = tune["n"]
    # the Dot2 launch reads 16-byte vectors of two doubles
    traffic = bench.profiled_traffic(n, (block, unroll, grid, 2), flags & 1)
    assert isinstance(traffic, int)
    algorithmic = 2 * 8 * n
    assert algorithmic <= traffic <= 1.01 * algorithmic


def test_traffic_profile_rejects_other_shapes():
    bench = _bench()
    assert bench.profiled_traffic(1 << 20, (512, 2, 192, 2), 1) is None
    assert bench.profiled_traffic(1 << 27, (256, 2, 192, 2), 1) is None
