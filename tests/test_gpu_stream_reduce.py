"""GPU parity: input generators, COPY/MUL/ADD/TRIAD, DOT, REDUCE_SUM through
the C ABI against the oracle.  Element-wise kernels are bit-exact; the
reductions are compared within a bound derived from the summed magnitudes
and are exact on integer-valued inputs."""

import numpy as np
import pytest

from oracle import SEQ

pytestmark = pytest.mark.gpu


def dev(rpx, a):
    return rpx.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64))


def test_seeded_fill_matches_oracle(rpx, orc):
    n = 1_000_003
    for sid, lo, hi in [(0, 0.0, 1.0), (5, -1.0, 1.0), (9, 0.5, 2.0)]:
        d = rpx.DeviceBuffer(8 * n)
        rpx.lib().rpx_fill_uniform_f64(d.ptr, n, 42, sid, lo, hi, None)
        assert np.array_equal(d.to_array(), orc.fill_uniform(n, 42, sid, lo, hi))


def test_formula_fill_matches_suite_initialiser(rpx, orc):
    n = 1_000_001
    d = rpx.DeviceBuffer(8 * n)
    for factor, count in [(0.2, 0), (0.1, 1)]:
        rpx.lib().rpx_fill_suite_formula_f64(d.ptr, n, factor, None)
        orc.reset_init_count()
        for _ in range(count):
            orc.init_data_const(0, 0.0)
        assert np.array_equal(d.to_array(), orc.init_data(n))


def test_grid_and_tensor_fill(rpx, orc):
    N = 130
    d = rpx.DeviceBuffer(8 * N * N)
    rpx.lib().rpx_fill_grid_f64(d.ptr, N, 42, 3, 0.0, 1.0, 0.0, None)
    assert np.array_equal(d.to_array(), orc.fill_grid(N, 42, 3))
    ne, q3 = 37, 125
    t = rpx.DeviceBuffer(8 * ne * q3 * 6)
    rpx.lib().rpx_fill_sym_tensor_f64(t.ptr, ne, q3, 6, 42, 2, None)
    assert np.array_equal(t.to_array(), orc.fill_sym_tensor(ne, q3, 6, 42, 2))


SIZES = [0, 1, 2, 3, 255, 4096, 4097, 1_000_000, 1_000_003]


@pytest.mark.parametrize("n", SIZES)
def test_stream_bit_exact(rpx, orc, rng, n):
    a, b = rng.random(n), rng.random(n)
    alpha = 0.75
    da, db = dev(rpx, a), dev(rpx, b)
    out = rpx.DeviceBuffer(8 * max(n, 1))
    L = rpx.lib()
    L.rpx_copy_f64(out.ptr, da.ptr, n, None, None)
    assert np.array_equal(out.to_array(count=n), orc.copy(a))
    L.rpx_mul_f64(out.ptr, da.ptr, alpha, n, None, None)
    assert np.array_equal(out.to_array(count=n), orc.mul(a, alpha))
    L.rpx_add_f64(out.ptr, da.ptr, db.ptr, n, None, None)
    assert np.array_equal(out.to_array(count=n), orc.add(a, b))
    L.rpx_triad_f64(out.ptr, da.ptr, db.ptr, alpha, n, None, None)
    assert np.array_equal(out.to_array(count=n), orc.triad(a, b, alpha))


@pytest.mark.parametrize("bs,items,flags", [(64, 2, 0), (128, 4, 1), (256, 8, 0), (256, 16, 1),
                                             (512, 8, 0), (1024, 4, 0)])
def test_triad_tunings_and_unaligned(rpx, orc, rng, bs, items, flags):
    n = 300_001
    b, c = rng.random(n + 1), rng.random(n + 1)
    db, dc = dev(rpx, b), dev(rpx, c)
    out = rpx.DeviceBuffer(8 * (n + 1))
    t = rpx.tuning(bs, items, 0, flags)
    rpx.lib().rpx_triad_f64(out.ptr, db.ptr, dc.ptr, 0.75, n, t, None)
    assert np.array_equal(out.to_array(count=n), orc.triad(b[:n], c[:n], 0.75))
    # operands offset by one double: the 8-byte-aligned path
    rpx.lib().rpx_triad_f64(out.ptr + 8, db.ptr + 8, dc.ptr + 8, 0.75, n, t, None)
    assert np.array_equal(out.to_array(count=n, offset_bytes=8), orc.triad(b[1:], c[1:], 0.75))


def test_stream_rejects_bad_arguments(rpx):
    from rajaperf_amd.abi import RpxError
    with pytest.raises(RpxError):
        rpx.lib().rpx_triad_f64(None, None, None, 1.0, 10, None, None)
    with pytest.raises(RpxError):
        rpx.lib().rpx_copy_f64(None, None, -1, None, None)
    a = rpx.DeviceBuffer(64)
    c = rpx.DeviceBuffer(64)
    with pytest.raises(RpxError):  # unsupported block size is refused before any launch
        rpx.lib().rpx_copy_f64(c.ptr, a.ptr, 4, rpx.tuning(96, 8), None)


def _reduce(rpx, fn, *args, n, tune=None):
    sb = rpx.lib().rpx_reduce_scratch_bytes(n, tune)
    scratch = rpx.DeviceBuffer(sb)
    rpx.lib().rpx_reduce_scratch_init(scratch.ptr, sb, None)
    res = rpx.DeviceBuffer(8)
    return scratch, sb, res


@pytest.mark.parametrize("n", [0, 1, 7, 4096, 1_000_000, 1_000_003, 16_777_217])
def test_dot_and_reduce_sum(rpx, orc, rng, n):
    a, b = rng.random(n), rng.random(n)
    x = rng.uniform(-1, 1, n)
    da, db, dx = dev(rpx, a), dev(rpx, b), dev(rpx, x)
    scratch, sb, res = _reduce(rpx, None, n=n)
    L = rpx.lib()
    L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 0, scratch.ptr, sb, None, None)
    got = res.to_array()[0]
    ref = orc.dot(a, b)
    assert abs(got - ref) <= 4e-16 * max(1, n) ** 0.5 * np.sum(np.abs(a * b)) + 1e-300
    L.rpx_reduce_sum_f64(dx.ptr, n, res.ptr, scratch.ptr, sb, None, None)
    got = res.to_array()[0]
    ref = orc.reduce_sum(x)
    assert abs(got - ref) <= 4e-16 * max(1, n) ** 0.5 * np.sum(np.abs(x)) + 1e-300


@pytest.mark.parametrize("n,items,flags", [
    (64 * 8192, 32, 0),          # exactly one full group of 64 workgroups
    (3 * 64 * 8192, 32, 0),      # three full groups
    (3 * 64 * 8192 + 3, 32, 1),  # a fourth group with one workgroup, odd tail
    (1_000_001, 2, 0), (1_000_001, 4, 1), (1_000_001, 8, 0), (1_000_001, 64, 0),
])
def test_reduction_ticket_levels_exact(rpx, rng, n, items, flags):
    a = rng.integers(-1000, 1000, n).astype(np.float64)
    b = rng.integers(-1000, 1000, n).astype(np.float64)
    da, db = dev(rpx, a), dev(rpx, b)
    t = rpx.tuning(0, items, 0, flags)
    scratch, sb, res = _reduce(rpx, None, n=n, tune=t)
    L = rpx.lib()
    for _ in range(2):  # tickets of both levels are reset for the next call
        L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 0, scratch.ptr, sb, t, None)
        assert res.to_array()[0] == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
        L.rpx_reduce_sum_f64(db.ptr, n, res.ptr, scratch.ptr, sb, t, None)
        assert res.to_array()[0] == float(b.astype(np.int64).sum())


def test_reduction_scratch_reused_across_layouts(rpx, rng):
    """One scratch, initialised once, serves calls of different n and
    tuning.  A layout with more groups than the call before it has ticket
    slots where that call kept its partials, so each call must leave every
    partial it consumed at zero (integer data: every sum is exact)."""
    L = rpx.lib()
    plan = [(16_777_219, 32), (16_777_219, 8), (1_000_003, 2), (5_000_011, 64), (16_777_219, 2),
            (300, 8), (16_777_219, 8)]
    big = max(L.rpx_reduce_scratch_bytes(n, rpx.tuning(0, it)) for n, it in plan)
    scratch = rpx.DeviceBuffer(big)
    L.rpx_reduce_scratch_init(scratch.ptr, big, None)
    a = rng.integers(-1000, 1000, 16_777_219).astype(np.float64)
    b = rng.integers(-1000, 1000, 16_777_219).astype(np.float64)
    da, db = dev(rpx, a), dev(rpx, b)
    res = rpx.DeviceBuffer(8)
    for n, items in plan:
        t = rpx.tuning(0, items)
        L.rpx_fill_const_f64(res.ptr, 1, 0.5, None)
        L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 0, scratch.ptr, big, t, None)
        assert res.to_array()[0] == float(np.dot(a[:n].astype(np.int64), b[:n].astype(np.int64))), (n, items)
        L.rpx_reduce_sum_f64(da.ptr, n, res.ptr, scratch.ptr, big, t, None)
        assert res.to_array()[0] == float(a[:n].astype(np.int64).sum()), (n, items)


def test_reductions_exact_on_integers_and_deterministic(rpx, rng):
    n = 5_000_011
    a = rng.integers(-1000, 1000, n).astype(np.float64)
    b = rng.integers(-1000, 1000, n).astype(np.float64)
    da, db = dev(rpx, a), dev(rpx, b)
    scratch, sb, res = _reduce(rpx, None, n=n)
    L = rpx.lib()
    L.rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 0, scratch.ptr, sb, None, None)
    assert res.to_array()[0] == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    L.rpx_reduce_sum_f64(da.ptr, n, res.ptr, scratch.ptr, sb, None, None)
    assert res.to_array()[0] == float(a.astype(np.int64).sum())
    # run-to-run bit identical on non-integer data
    x = rng.uniform(-1, 1, n)
    dx = dev(rpx, x)
    vals = set()
    for _ in range(4):
        L.rpx_reduce_sum_f64(dx.ptr, n, res.ptr, scratch.ptr, sb, None, None)
        vals.add(res.to_array()[0].tobytes())
    assert len(vals) == 1


def test_dot_accumulates_like_the_suite(rpx, orc, rng):
    # DOT-Seq.cpp:39-45: m_dot += dot every rep
    n = 100_000
    a, b = rng.random(n), rng.random(n)
    da, db = dev(rpx, a), dev(rpx, b)
    scratch, sb, res = _reduce(rpx, None, n=n)
    res.upload(np.zeros(1))
    for _ in range(5):
        rpx.lib().rpx_dot_f64(da.ptr, db.ptr, n, res.ptr, 1, scratch.ptr, sb, None, None)
    ref = 0.0
    for _ in range(5):
        ref += orc.dot(a, b, SEQ)
    assert abs(res.to_array()[0] - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("bs,items,grid", [(128, 2, 0), (256, 4, 100), (512, 16, 3000), (1024, 8, 7)])
def test_reduce_tunings(rpx, orc, rng, bs, items, grid):
    n = 2_000_003
    x = rng.uniform(-1, 1, n)
    dx = dev(rpx, x)
    t = rpx.tuning(bs, items, grid, 0)
    sb = rpx.lib().rpx_reduce_scratch_bytes(n, t)
    scratch = rpx.DeviceBuffer(sb)
    rpx.lib().rpx_reduce_scratch_init(scratch.ptr, sb, None)
    res = rpx.DeviceBuffer(8)
    rpx.lib().rpx_reduce_sum_f64(dx.ptr, n, res.ptr, scratch.ptr, sb, t, None)
    assert abs(res.to_array()[0] - orc.reduce_sum(x)) <= 1e-12 * np.sum(np.abs(x))
    from rajaperf_amd.abi import RpxError
    with pytest.raises(RpxError):  # scratch too small is an error, not a fault
        rpx.lib().rpx_reduce_sum_f64(dx.ptr, n, res.ptr, scratch.ptr, 64, t, None)
