"""The allocation-free fold (msl_fold_sum_f32_noalloc) against the allocating
msl_fold_sum_f32: same kernels and launch shape, so the results must be
identical bit for bit, on consecutive calls (the second on a smaller input,
over the scratch contents the first one left) and under HIP graph capture,
where nothing may be allocated.  Both are also within the fold tolerance of
the oracle's FP64 sum."""
import ctypes as C
import sys

import numpy as np
import pytest

from muesli4_amd import hip
from muesli4_amd.hip import DeviceArray, call
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

FOLD_REL_TOL = 1e-6


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    hip.init(0)
    yield
    hip.synchronize()


def _fold_both(d, n):
    res = DeviceArray(2, np.float64)
    call("msl_fill_f32", res.ptr, 4, float("nan"))  # both slots start as NaN
    call("msl_fold_sum_f32", d.ptr, n, res.ptr)
    call("msl_fold_sum_f32_noalloc", d.ptr, n, res.ptr + 8)
    return res.to_host()


def test_noalloc_equals_allocating_fold_on_consecutive_calls():
    for n, seed in ((8192 * 8192 + 3, 301), (123457, 302), (5, 303)):
        note(f"fold noalloc n={n}")
        d = DeviceArray(n, np.float32)
        call("msl_fill_uniform_f32", d.ptr, n, seed, float(R.SCALE_255))
        got = _fold_both(d, n)
        assert got.view(np.uint64)[0] == got.view(np.uint64)[1], got
        ref = R.fold_sum_f64(R.fill_uniform_f32(n, seed))
        assert abs(got[1] - ref) <= FOLD_REL_TOL * abs(ref), (got, ref)


def test_noalloc_empty_input_and_errors():
    note("fold noalloc empty / null result")
    d = DeviceArray(4, np.float32)
    got = _fold_both(d, 0)
    assert got[0] == 0.0 and got[1] == 0.0
    with pytest.raises(hip.MslError):
        call("msl_fold_sum_f32_noalloc", d.ptr, 4, None)


def test_noalloc_under_graph_capture():
    n = 1000003
    note(f"fold noalloc captured n={n}")
    d = DeviceArray(n, np.float32)
    call("msl_fill_uniform_f32", d.ptr, n, 304, float(R.SCALE_255))
    res = DeviceArray(1, np.float64)
    g = C.c_void_p()
    call("msl_graph_begin")
    call("msl_fold_sum_f32_noalloc", d.ptr, n, res.ptr)
    call("msl_graph_end", C.byref(g))
    eager = DeviceArray(1, np.float64)
    call("msl_fold_sum_f32", d.ptr, n, eager.ptr)
    expect = eager.to_host()[0]
    small = DeviceArray(999, np.float32)
    call("msl_fill_f32", small.ptr, 999, 1.0)
    for rep in range(4):
        # eager work between replays, over the same scratch array
        call("msl_fill_f32", res.ptr, 2, -1.0)
        call("msl_fold_sum_f32_noalloc", small.ptr, 999, eager.ptr)
        assert eager.to_host()[0] == 999.0
        call("msl_graph_launch", g)
        assert res.to_host()[0] == expect, f"replay {rep}"
    call("msl_graph_destroy", g)
