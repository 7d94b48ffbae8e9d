"""hpb_multidot_f64_noalloc: the fused multi-dot whose per-block partials live in a
file-scope device array instead of a handle's workspace, for callers that cannot
allocate.  Called through the C ABI on the handle's stream and compared bit for bit
with hpb_multidot_f64_async, which runs the same kernels with the same launch shape
on the handle's workspace: two calls in a row, the second smaller, compensated and
plain; k beyond the array's 64 columns is rejected."""
import ctypes

import numpy as np
import pytest

import high_precision_blas1_amd as H
from conftest import say

pytestmark = pytest.mark.gpu


def _cols(handle, k, n, seed):
    r = np.random.default_rng(seed)
    V = [H.DeviceArray.from_numpy(r.standard_normal(n) * np.exp2(r.integers(-8, 8, n)), handle) for _ in range(k)]
    w = H.DeviceArray.from_numpy(r.standard_normal(n), handle)
    return V, w


def _both(handle, V, w, n, comp):
    k = len(V)
    ptrs = (ctypes.c_void_p * k)(*[v.addr for v in V])
    a = H.DeviceArray(4 * k, np.float64, handle)
    b = H.DeviceArray(4 * k, np.float64, handle)
    H.call("hpb_multidot_f64_async", handle.ptr, k, ptrs, n, w.ptr, comp, a.ptr)
    stream = ctypes.c_void_p()
    H.call("hpb_get_stream", handle.ptr, ctypes.byref(stream))
    H.call("hpb_multidot_f64_noalloc", stream, k, ptrs, n, w.ptr, comp, b.ptr)
    ra, rb = a.numpy(), b.numpy()
    a.free()
    b.free()
    return ra, rb


@pytest.mark.parametrize("comp", [1, 0])
def test_noalloc_matches_handle_entry(handle, comp):
    for k, n, seed in ((64, (1 << 20) + 5, 1), (9, 70001, 2)):   # the second call smaller
        say("multidot noalloc", comp, k, n)
        V, w = _cols(handle, k, n, seed)
        ra, rb = _both(handle, V, w, n, comp)
        assert ra.tobytes() == rb.tobytes(), (k, n, comp)
        for v in V + [w]:
            v.free()


def test_noalloc_rejects_too_many_columns(handle):
    n = 1000
    V, w = _cols(handle, 2, n, 3)
    ptrs = (ctypes.c_void_p * 65)(*([V[0].addr] * 65))
    out = H.DeviceArray(4 * 65, np.float64, handle)
    stream = ctypes.c_void_p()
    H.call("hpb_get_stream", handle.ptr, ctypes.byref(stream))
    with pytest.raises(H.HPBError, match="64"):
        H.call("hpb_multidot_f64_noalloc", stream, 65, ptrs, n, w.ptr, 1, out.ptr)
    for v in V + [w, out]:
        v.free()
