"""GPU parity: SCAN (decoupled look-back) and JACOBI_2D through the C ABI
against the oracle.  Scans are exact on integer-valued inputs and within a
magnitude-derived bound otherwise; Jacobi sweeps are bit-exact."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(rpx, a):
    return rpx.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64))


def scan_scratch(rpx, n):
    sb = rpx.lib().rpx_scan_scratch_bytes(n, None)
    s = rpx.DeviceBuffer(sb)
    rpx.lib().rpx_scan_scratch_init(s.ptr, sb, None)
    return s, sb


@pytest.mark.parametrize("bs", [256, 512])
@pytest.mark.parametrize("ipt", [16, 32, 64])
@pytest.mark.parametrize("n", [0, 1, 2, 17, 4095, 4096, 4097, 8192, 12289, 1_000_000, 3_000_017])
def test_scan_matches_oracle(rpx, orc, rng, n, ipt, bs):
    x = rng.uniform(-1, 1, n)
    dx = dev(rpx, x)
    y = rpx.DeviceBuffer(8 * max(n, 1))
    s, sb = scan_scratch(rpx, n)
    rpx.lib().rpx_scan_exclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(bs, ipt), None)
    got = y.to_array(count=n)
    ref = orc.scan(x)
    if n:
        assert got[0] == 0.0
        bound = 1e-15 * np.cumsum(np.abs(x)) + 1e-300
        assert np.all(np.abs(got - ref)[1:] <= 8 * np.sqrt(np.arange(1, n)) * bound[:-1] + 1e-14)
    assert rpx.lib().rpx_scan_scratch_status(s.ptr) == 0


@pytest.mark.parametrize("bs,ipt", [(256, 16), (256, 64), (512, 32)])
@pytest.mark.parametrize("n", [1, 2, 4097, 12289, 1_000_003])
def test_scan_inclusive(rpx, orc, rng, n, ipt, bs):
    """The inclusive form y[i] = x[0] + ... + x[i].  Exact on integer data
    (against numpy's cumsum) and, on uniform data, within the exclusive
    scan's bound of the oracle's exclusive prefix plus x[i]."""
    s, sb = scan_scratch(rpx, n)
    y = rpx.DeviceBuffer(8 * n)
    xi = rng.integers(-50, 50, n).astype(np.float64)
    dx = dev(rpx, xi)
    rpx.lib().rpx_scan_inclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(bs, ipt), None)
    assert np.array_equal(y.to_array(count=n), np.cumsum(xi))
    x = rng.uniform(-1, 1, n)
    dx = dev(rpx, x)
    rpx.lib().rpx_scan_inclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(bs, ipt), None)
    got = y.to_array(count=n)
    ref = orc.scan(x) + x
    bound = 1e-15 * np.cumsum(np.abs(x))
    assert np.all(np.abs(got - ref) <= 8 * np.sqrt(np.arange(1, n + 1)) * bound + 1e-14)
    # the exclusive scan after an inclusive one on the same scratch
    rpx.lib().rpx_scan_exclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(bs, ipt), None)
    assert y.to_array(count=n)[0] == 0.0
    assert rpx.lib().rpx_scan_scratch_status(s.ptr) == 0


def test_scan_exact_on_integers_large(rpx, rng):
    n = 1 << 26
    x = rng.integers(-50, 50, n).astype(np.float64)
    dx = dev(rpx, x)
    y = rpx.DeviceBuffer(8 * n)
    s, sb = scan_scratch(rpx, n)
    for ipt in (16, 32, 64, 16):  # the scratch is reusable without re-initialisation
        rpx.lib().rpx_scan_exclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(0, ipt), None)
        got = y.to_array()
        assert got[0] == 0.0
        assert np.array_equal(got[1:], np.cumsum(x)[:-1])
    assert rpx.lib().rpx_scan_scratch_status(s.ptr) == 0


def test_scan_unaligned_and_bad_tuning(rpx, orc, rng):
    n = 20_001
    x = rng.integers(-9, 9, n + 1).astype(np.float64)
    dx = dev(rpx, x)
    y = rpx.DeviceBuffer(8 * (n + 1))
    s, sb = scan_scratch(rpx, n)
    # one-double offsets: the 16-byte row loads are not legal, rows go scalar
    rpx.lib().rpx_scan_exclusive_f64(dx.ptr + 8, y.ptr + 8, n, s.ptr, sb, None, None)
    assert np.array_equal(y.to_array(count=n, offset_bytes=8), orc.scan(x[1:]))
    with pytest.raises(rpx.RpxError):
        rpx.lib().rpx_scan_exclusive_f64(dx.ptr, y.ptr, n, s.ptr, sb, rpx.tuning(0, 24), None)
    assert rpx.lib().rpx_scan_scratch_status(s.ptr) == 0


def test_scan_in_place_and_segmented(rpx, orc, rng):
    n = 100_003
    x = rng.integers(-9, 9, n).astype(np.float64)
    d = dev(rpx, x)
    s, sb = scan_scratch(rpx, n)
    rpx.lib().rpx_scan_exclusive_f64(d.ptr, d.ptr, n, s.ptr, sb, None, None)
    assert np.array_equal(d.to_array(), orc.scan(x))
    offs = [0, 5, 5, 4096, 4100, 50_000, n]
    d2 = dev(rpx, x)
    y = rpx.DeviceBuffer(8 * n)
    o = rpx.offsets_array(offs)
    sb2 = rpx.lib().rpx_scan_segmented_scratch_bytes(o, len(offs) - 1)
    s2 = rpx.DeviceBuffer(sb2)
    rpx.lib().rpx_scan_exclusive_segmented_f64(d2.ptr, y.ptr, o, len(offs) - 1, s2.ptr, sb2, None)
    got = y.to_array()
    for b, e in zip(offs[:-1], offs[1:]):
        assert np.array_equal(got[b:e], orc.scan(x[b:e]))


def test_scan_segmented_then_plain_share_scratch(rpx, orc, rng):
    """A single-array scan on a freshly initialised scratch that a segmented
    call has just written (its flags carry tag 0) must not accept those flags,
    and the segmented form must still be right afterwards."""
    L = rpx.lib()
    n = 600_011
    x = rng.integers(-20, 20, n).astype(np.float64)
    offs = [0, 7, 70_000, 70_001, 333_333, n]
    o = rpx.offsets_array(offs)
    sb = max(L.rpx_scan_segmented_scratch_bytes(o, len(offs) - 1), L.rpx_scan_scratch_bytes(n, None))
    scr = rpx.DeviceBuffer(sb)
    L.rpx_scan_scratch_init(scr.ptr, sb, None)
    dx, y = dev(rpx, x), rpx.DeviceBuffer(8 * n)
    seg_ref = np.concatenate([orc.scan(x[b:e]) for b, e in zip(offs[:-1], offs[1:])])
    for _ in range(2):
        L.rpx_scan_exclusive_segmented_f64(dx.ptr, y.ptr, o, len(offs) - 1, scr.ptr, sb, None)
        assert np.array_equal(y.to_array(), seg_ref)
        L.rpx_scan_exclusive_f64(dx.ptr, y.ptr, n, scr.ptr, sb, None, None)
        assert np.array_equal(y.to_array(), orc.scan(x))
    assert L.rpx_scan_scratch_status(scr.ptr) == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_scan_segmented_ragged_one_launch(rpx, orc, seed):
    """Ragged segments (empty, odd offsets, multi-tile) in one launch, twice on
    the same scratch; exact on integer-valued inputs."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(0, np.log(70_000), 97)).astype(np.int64)
    lens[::13] = 0
    offs = np.concatenate([[0], np.cumsum(lens)]).tolist()
    n = offs[-1]
    x = rng.integers(-50, 50, n).astype(np.float64)
    dx, y = dev(rpx, x), rpx.DeviceBuffer(8 * n)
    o = rpx.offsets_array(offs)
    L = rpx.lib()
    sb = L.rpx_scan_segmented_scratch_bytes(o, len(offs) - 1)
    scr = rpx.DeviceBuffer(sb)
    for _ in range(2):
        L.rpx_scan_exclusive_segmented_f64(dx.ptr, y.ptr, o, len(offs) - 1, scr.ptr, sb, None)
        got = y.to_array()
        for b, e in zip(offs[:-1], offs[1:]):
            assert np.array_equal(got[b:e], orc.scan(x[b:e]))
    assert L.rpx_scan_scratch_status(scr.ptr) == 0
    with pytest.raises(rpx.RpxError):  # scratch too small is an error, not a fault
        L.rpx_scan_exclusive_segmented_f64(dx.ptr, y.ptr, o, len(offs) - 1, scr.ptr, sb - 256, None)


@pytest.mark.parametrize("start", [1, 2])
def test_scan_segmented_tile_edges(rpx, orc, rng, start):
    """Segment lengths around the 16384-element tile at odd and even starts
    (a segment's tiles are counted from its even-rounded start), with one
    element before the first segment and one after the last that the call
    must leave untouched."""
    T = 16384
    lens = [1, T - 1, T, T + 1, 2 * T - 1, 2 * T, 2 * T + 1, 1, 2, 0, T - 2, 3, T + 2]
    offs = np.concatenate([[start], start + np.cumsum(lens)]).tolist()
    n = offs[-1] + 1
    x = rng.integers(-50, 50, n).astype(np.float64)
    dx, y = dev(rpx, x), dev(rpx, np.full(n, 7.5))
    o = rpx.offsets_array(offs)
    L = rpx.lib()
    sb = L.rpx_scan_segmented_scratch_bytes(o, len(offs) - 1)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_scan_exclusive_segmented_f64(dx.ptr, y.ptr, o, len(offs) - 1, scr.ptr, sb, None)
    got = y.to_array()
    for b, e in zip(offs[:-1], offs[1:]):
        assert np.array_equal(got[b:e], orc.scan(x[b:e])), (b, e)
    assert np.all(got[:start] == 7.5) and got[-1] == 7.5
    assert L.rpx_scan_scratch_status(scr.ptr) == 0


@pytest.mark.parametrize("N", [3, 4, 5, 8, 33, 34, 130, 514, 1026, 2050])
@pytest.mark.parametrize("tsteps", [1, 3])
def test_jacobi_bit_exact(rpx, orc, rng, N, tsteps):
    A0 = rng.random(N * N)
    B0 = rng.random(N * N)
    dA, dB = dev(rpx, A0), dev(rpx, B0)
    rpx.lib().rpx_jacobi2d_f64(dA.ptr, dB.ptr, N, tsteps, None, None)
    A, B = orc.jacobi2d(A0, B0, N, tsteps)
    assert np.array_equal(dA.to_array(), A)
    assert np.array_equal(dB.to_array(), B)


@pytest.mark.parametrize("N", [3, 4, 6, 8, 10, 130, 250, 251, 1026, 2050])
@pytest.mark.parametrize("tsteps", [1, 2, 3, 5, 6])
@pytest.mark.parametrize("rows,depth", [(4, 1), (8, 1), (16, 1), (32, 1), (4, 2), (8, 2), (12, 2), (16, 2),
                                        (4, 3), (8, 3), (12, 3), (16, 3),
                                        (48, 1), (54, 2), (54, 3), (99, 3), (207, 3), (4096, 3),
                                        (54, 4), (144, 4)])
def test_jacobi_fused_bit_exact(rpx, orc, rng, N, tsteps, rows, depth):
    """Fused launches against the two-sweep oracle, bit for bit.  Rows of 48
    and more run the streaming kernel (a wavefront walks a tall chunk)."""
    A0 = rng.random(N * N)
    B0 = rng.random(N * N)
    dA, dB = dev(rpx, A0), dev(rpx, B0)
    sb = rpx.lib().rpx_jacobi2d_fused_scratch_bytes(N)
    scr = rpx.DeviceBuffer(sb)
    rpx.lib().rpx_jacobi2d_fused_f64(dA.ptr, dB.ptr, N, tsteps, scr.ptr, sb, rpx.tuning(0, rows, depth=depth), None)
    A, B = orc.jacobi2d(A0, B0, N, tsteps)
    assert np.array_equal(dA.to_array(), A)
    assert np.array_equal(dB.to_array(), B)


@pytest.mark.parametrize("rows,depth", [(32, 2), (32, 3), (6, 1), (8, 4), (16, 4), (8, 0), (47, 3), (4097, 3),
                                        (64, 5)])
def test_jacobi_fused_unsupported_shapes(rpx, rows, depth):
    """Tile shapes without a kernel are refused before anything is launched."""
    N = 130
    A, B, Q = (rpx.DeviceBuffer(8 * N * N) for _ in range(3))
    L = rpx.lib()
    with pytest.raises(rpx.RpxError):
        L.rpx_jacobi2d_fused_launch_f64(Q.ptr, B.ptr, A.ptr, N, depth, rpx.tuning(0, rows), None)
    if depth > 0:
        with pytest.raises(rpx.RpxError):
            L.rpx_jacobi2d_fused_f64(A.ptr, B.ptr, N, 5, Q.ptr, 8 * N * N, rpx.tuning(0, rows, depth=depth), None)


@pytest.mark.parametrize("rows", [4, 8, 16, 32])
def test_jacobi_tunings(rpx, orc, rng, rows):
    N = 1030
    A0 = rng.random(N * N)
    dA, dB = dev(rpx, A0), rpx.DeviceBuffer(8 * N * N)
    dB.upload(A0)
    rpx.lib().rpx_jacobi2d_sweep_f64(dB.ptr, dA.ptr, N, rpx.tuning(0, rows), None)
    # one sweep A -> B equals the B of the first half-step
    Bref = A0.reshape(N, N).copy()
    An = A0.reshape(N, N)
    Bref[1:-1, 1:-1] = 0.2 * ((((An[1:-1, 1:-1] + An[1:-1, :-2]) + An[1:-1, 2:]) + An[2:, 1:-1]) + An[:-2, 1:-1])
    assert np.array_equal(dB.to_array(), Bref.reshape(-1))


@pytest.mark.parametrize("N", [6, 130, 1026])
@pytest.mark.parametrize("depth,rows", [(d, r) for d in (1, 2, 3) for r in (4, 8, 16, 54, 207)] +
                         [(4, 54), (4, 144)])
def test_jacobi_launch_sequence_skip_b(rpx, orc, rng, N, depth, rows):
    """A run of fused launches where only the last stores B (RPX_TUNE_JACOBI_SKIP_B
    on the others) ends in the same A and B as the two-sweep oracle."""
    launches = 3
    A0, B0 = rng.random(N * N), rng.random(N * N)
    grids = [dev(rpx, A0), rpx.DeviceBuffer(8 * N * N)]
    dB = dev(rpx, B0)
    L = rpx.lib()
    L.rpx_jacobi2d_copy_boundary_f64(grids[1].ptr, grids[0].ptr, N, None)
    for k in range(launches):
        flags = rpx.RPX_TUNE_JACOBI_SKIP_B if k < launches - 1 else 0
        L.rpx_jacobi2d_fused_launch_f64(grids[1].ptr, dB.ptr, grids[0].ptr, N, depth,
                                        rpx.tuning(0, rows, 0, flags), None)
        grids.reverse()
    A, B = orc.jacobi2d(A0, B0, N, launches * depth)
    assert np.array_equal(grids[0].to_array(), A)
    assert np.array_equal(dB.to_array(), B)
