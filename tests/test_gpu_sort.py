"""GPU parity: segmented SORT against std::sort (the oracle) and numpy.
Value-exact: the sorted values equal std::sort's; the bit patterns are a
permutation of the input's; -0.0 orders before +0.0 (std::sort leaves the
relative order of equal zeros unspecified)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_sort(rpx, x, offsets):
    d = rpx.DeviceBuffer.from_array(x)
    o = rpx.offsets_array(offsets)
    nseg = len(offsets) - 1
    sb = rpx.lib().rpx_sort_segmented_scratch_bytes(o, nseg)
    s = rpx.DeviceBuffer(max(sb, 8))
    rpx.lib().rpx_sort_segmented_f64(d.ptr, o, nseg, s.ptr, sb, None)
    return d.to_array()


def check_segment(got, x, orc):
    ref = orc.sort(x)
    assert np.array_equal(got, ref)  # value-exact (+0.0 == -0.0)
    assert np.array_equal(np.sort(got.view(np.uint64)), np.sort(x.view(np.uint64)))
    zeros = np.flatnonzero(got == 0.0)
    if zeros.size:
        signs = np.signbit(got[zeros])
        assert np.all(signs[:np.count_nonzero(signs)])  # all -0.0 before +0.0


@pytest.mark.parametrize("n", [1, 2, 3, 100, 4095, 4096, 4097, 65537, 1_000_000])
def test_sort_single_segment(rpx, orc, rng, n):
    x = rng.uniform(-1, 1, n)
    got = run_sort(rpx, x, [0, n])
    check_segment(got, x, orc)


def test_sort_special_values(rpx, orc, rng):
    n = 50_000
    x = rng.uniform(-1, 1, n)
    specials = [0.0, -0.0, np.inf, -np.inf, 5e-324, -5e-324, 2.2250738585072014e-308, 1.0, -1.0, 1e308, -1e308]
    x[rng.integers(0, n, 2000)] = rng.choice(specials, 2000)
    x[:100] = 0.5  # long run of duplicates
    got = run_sort(rpx, x, [0, n])
    check_segment(got, x, orc)


def test_sort_reference_layout(rpx, orc, rng):
    # SORT.hpp:21-25: rep irep sorts x[irep*N, (irep+1)*N)
    N, reps = 10_000, 5
    x = rng.random(N * reps)
    got = run_sort(rpx, x, [i * N for i in range(reps + 1)])
    for r in range(reps):
        check_segment(got[r * N:(r + 1) * N], x[r * N:(r + 1) * N], orc)


def test_sort_ragged_segments_and_gaps(rpx, orc, rng):
    n = 30_000
    x = rng.uniform(-1, 1, n)
    offs = [10, 20, 20, 5000, 9001, 29_000]  # starts at 10, an empty segment, a tail gap
    got = run_sort(rpx, x, offs)
    assert np.array_equal(got[:10], x[:10]) and np.array_equal(got[29_000:], x[29_000:])
    for b, e in zip(offs[:-1], offs[1:]):
        check_segment(got[b:e], x[b:e], orc)


def test_sort_log_uniform_segments(rpx, orc):
    # 256 segments with lengths log-uniform in [2^10, 2^20]
    rng = np.random.default_rng(42)
    lens = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 256)).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)])
    x = rng.uniform(-1, 1, int(offs[-1]))
    got = run_sort(rpx, x, offs.tolist())
    for b, e in zip(offs[:-1], offs[1:]):
        assert np.array_equal(got[b:e], np.sort(x[b:e]))


def test_sort_empty_and_errors(rpx):
    from rajaperf_amd.abi import RpxError
    assert run_sort(rpx, np.zeros(4), [0, 0, 0]).size == 4
    o = rpx.offsets_array([0, 100])
    d = rpx.DeviceBuffer(800)
    with pytest.raises(RpxError):  # scratch too small is refused before any launch
        rpx.lib().rpx_sort_segmented_f64(d.ptr, o, 1, d.ptr, 16, None)
    with pytest.raises(RpxError):
        rpx.lib().rpx_sort_segmented_f64(d.ptr, rpx.offsets_array([5, 2]), 1, d.ptr, 800, None)
