"""The allocation-free DOT entry point against the scratch-taking one.

rpx_dot_f64_noalloc runs the kernels of rpx_dot_f64 on a library-owned
device array, so for the same (n, tuning) both must give the same bits.
Two calls in a row, the second on a smaller input, check that the array
is left ready (tickets back at zero) and that a smaller layout does not
read stale partials from the larger one.  Integer-valued inputs make the
dot products exact, so both are also checked against the int64 sum.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(rpx, x):
    return rpx.DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float64))


@pytest.mark.parametrize("tune_items", [0, 8])
def test_dot_noalloc_matches_dot(rpx, tune_items):
    L = rpx.lib()
    rng = np.random.default_rng(31)
    t = rpx.tuning(0, tune_items, 0, 0) if tune_items else None
    for n in (16_777_219, 1_000_003):  # second call smaller than the first
        print(f"noalloc DOT n={n} items={tune_items}", flush=True)
        a, b = rng.random(n), rng.random(n)
        da, db = _dev(rpx, a), _dev(rpx, b)
        sb = L.rpx_reduce_scratch_bytes(n, t)
        assert 0 < sb <= L.rpx_dot_noalloc_max_scratch_bytes()
        scr = rpx.DeviceBuffer(sb)
        L.rpx_reduce_scratch_init(scr.ptr, sb, None)
        res = rpx.DeviceBuffer(16)
        L.rpx_fill_const_f64(res.ptr, 2, 0.5, None)
        L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 0, scr.ptr, sb, t, None)
        L.rpx_dot_f64_noalloc(da.ptr, db.ptr, n, res.ptr + 8, 0, t, None)
        got = res.to_array()
        assert got[0].tobytes() == got[1].tobytes(), (n, got)
        # accumulate form: *result += dot, on both entry points
        L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 1, scr.ptr, sb, t, None)
        L.rpx_dot_f64_noalloc(da.ptr, db.ptr, n, res.ptr + 8, 1, t, None)
        got = res.to_array()
        assert got[0].tobytes() == got[1].tobytes(), (n, got)
        ref = float(np.dot(a, b))
        assert abs(got[1] - 2 * ref) <= 1e-12 * abs(ref)


def test_dot_noalloc_exact_on_integers_twice(rpx):
    L = rpx.lib()
    rng = np.random.default_rng(32)
    for n in (5_000_011, 333_333):
        print(f"noalloc DOT integer n={n}", flush=True)
        a = rng.integers(-1000, 1000, n).astype(np.float64)
        b = rng.integers(-1000, 1000, n).astype(np.float64)
        da, db = _dev(rpx, a), _dev(rpx, b)
        res = rpx.DeviceBuffer(8)
        L.rpx_dot_f64_noalloc(da.ptr, db.ptr, n, res.ptr, 0, None, None)
        assert res.to_array()[0] == float(np.dot(a.astype(np.int64), b.astype(np.int64)))


def test_dot_noalloc_rejects_oversize_and_bad_arguments(rpx):
    L = rpx.lib()
    res = rpx.DeviceBuffer(8)
    raw = L.raw("rpx_dot_f64_noalloc")
    # no launch happens for either: the sizes are checked first
    big = 1 << 34
    assert L.rpx_reduce_scratch_bytes(big, None) > L.rpx_dot_noalloc_max_scratch_bytes()
    st = raw(C.c_void_p(res.ptr), C.c_void_p(res.ptr), C.c_int64(big), C.c_void_p(res.ptr), 0, None, None)
    assert st == rpx.RPX_ERR_SCRATCH_TOO_SMALL
    st = raw(None, None, C.c_int64(10), C.c_void_p(res.ptr), 0, None, None)
    assert st == rpx.RPX_ERR_INVALID_ARGUMENT
    st = raw(C.c_void_p(res.ptr), C.c_void_p(res.ptr), C.c_int64(-1), C.c_void_p(res.ptr), 0, None, None)
    assert st == rpx.RPX_ERR_INVALID_ARGUMENT
