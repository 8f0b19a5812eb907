"""GPU parity at the full workload sizes the benchmark and the roofline runs use.

* Streams: TRIAD and COPY at N = 2^28 (2 GiB per array) and at 2^28 + 5,
  whose last elements sit past byte offset 2^31; every element compared with
  the oracle.
* DOT, REDUCE_SUM and SCAN at N = 2^28 on integer-valued data, where every
  FP64 partial sum is exact, so the results must equal the int64 sums; SCAN
  also on uniform [-1,1) data against the long-double prefix bound.
* JACOBI_2D on the 8192 x 8192 grid, 50 time steps: the benchmark's launch
  plan (16 launches of 3 time steps that leave B unwritten, one of 2 that
  stores it) and the fused entry point, each bit-exact against the two-sweep
  path and against the oracle's Base_OpenMP.
* MASS3DPA and DIFFUSION3DPA at NE = 262,144, (D1D, Q1D) = (4, 5), with the
  Gauss-Lobatto basis, bit-exact against the oracle's Base_OpenMP.
* GEMM 4096^3 (alpha 1.5): FP64 MFMA and the 13-modulus mixed path, on sampled rows
  against 80-bit host arithmetic.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 42
N28 = 1 << 28


def _fill(L, buf, n, stream_id, lo, hi):
    L.rpx_fill_uniform_f64(buf.ptr, n, SEED, stream_id, lo, hi, None)


@pytest.mark.parametrize("n", [N28, N28 + 5])
def test_triad_copy_full_size(rpx, orc, n):
    L = rpx.lib()
    b, c, a = rpx.DeviceBuffer(8 * n), rpx.DeviceBuffer(8 * n), rpx.DeviceBuffer(8 * n)
    _fill(L, b, n, 1, 0.0, 1.0)
    _fill(L, c, n, 2, 0.0, 1.0)
    hb = orc.fill_uniform(n, SEED, 1, 0.0, 1.0)
    hc = orc.fill_uniform(n, SEED, 2, 0.0, 1.0)
    assert np.array_equal(b.to_array(), hb)
    L.rpx_triad_f64(a.ptr, b.ptr, c.ptr, 0.75, n, None, None)
    got = a.to_array()
    ref = orc.triad(hb, hc, 0.75, orc.OMP)
    assert np.array_equal(got, ref)
    # the last elements, past byte offset 2^31 when n > 2^28
    assert np.array_equal(a.to_array(count=5, offset_bytes=8 * (n - 5)), ref[-5:])
    del got, ref
    L.rpx_copy_f64(a.ptr, c.ptr, n, None, None)
    assert np.array_equal(a.to_array(), hc)


def _int_valued(rng, n, k):
    return rng.integers(-k, k + 1, size=n).astype(np.float64)


def _reduce_scratch(rpx, n):
    L = rpx.lib()
    sb = L.rpx_reduce_scratch_bytes(n, None)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_reduce_scratch_init(scr.ptr, sb, None)
    return scr, sb


def test_dot_reduce_sum_full_size_exact(rpx):
    L = rpx.lib()
    rng = np.random.default_rng(11)
    n = N28
    ha = _int_valued(rng, n, 3)
    hb = _int_valued(rng, n, 3)
    a, b = rpx.DeviceBuffer.from_array(ha), rpx.DeviceBuffer.from_array(hb)
    scr, sb = _reduce_scratch(rpx, n)
    res = rpx.DeviceBuffer(16)
    L.rpx_fill_const_f64(res.ptr, 2, 0.0, None)
    # two reps accumulate (m_dot += dot, DOT-Seq.cpp:45); REDUCE_SUM overwrites
    L.rpx_dot_f64(a.ptr, b.ptr, n, res.ptr, 1, scr.ptr, sb, None, None)
    L.rpx_dot_f64(a.ptr, b.ptr, n, res.ptr, 1, scr.ptr, sb, None, None)
    L.rpx_reduce_sum_f64(a.ptr, n, res.ptr + 8, scr.ptr, sb, None, None)
    L.rpx_reduce_sum_f64(a.ptr, n, res.ptr + 8, scr.ptr, sb, None, None)
    got = res.to_array()
    dot_ref = int(np.dot(ha.astype(np.int64), hb.astype(np.int64)))
    sum_ref = int(ha.astype(np.int64).sum())
    assert got[0] == 2 * dot_ref
    assert got[1] == sum_ref


def _scan_scratch(rpx, n):
    L = rpx.lib()
    sb = L.rpx_scan_scratch_bytes(n, None)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_scan_scratch_init(scr.ptr, sb, None)
    return scr, sb


def test_scan_full_size_integers_exact(rpx):
    L = rpx.lib()
    rng = np.random.default_rng(12)
    n = N28
    hx = _int_valued(rng, n, 7)
    x, y = rpx.DeviceBuffer.from_array(hx), rpx.DeviceBuffer(8 * n)
    scr, sb = _scan_scratch(rpx, n)
    for _ in range(2):  # the second call reuses the scratch (next epoch)
        L.rpx_scan_exclusive_f64(x.ptr, y.ptr, n, scr.ptr, sb, None, None)
        assert L.rpx_scan_scratch_status(scr.ptr) == 0
        got = y.to_array()
        ref = np.empty(n, dtype=np.int64)
        ref[0] = 0
        np.cumsum(hx[:-1].astype(np.int64), out=ref[1:])
        assert np.array_equal(got, ref.astype(np.float64))
        del got, ref


def test_scan_full_size_uniform_bound(rpx, orc):
    """U[-1,1) at 2^28 against a long-double prefix.  Each device prefix is a
    chain of at most n/4096 tile aggregates (4096-element tiles at the
    smallest) plus an in-tile tree of fewer than 256 additions, so its error is
    below (n/4096 + 256) eps times the prefix of |x|; Base_Seq's sequential
    sum is below (i + 1) eps times the same prefix."""
    L = rpx.lib()
    n = N28
    x, y = rpx.DeviceBuffer(8 * n), rpx.DeviceBuffer(8 * n)
    _fill(L, x, n, 0, -1.0, 1.0)
    scr, sb = _scan_scratch(rpx, n)
    L.rpx_scan_exclusive_f64(x.ptr, y.ptr, n, scr.ptr, sb, None, None)
    assert L.rpx_scan_scratch_status(scr.ptr) == 0
    hx = orc.fill_uniform(n, SEED, 0, -1.0, 1.0)
    got = y.to_array()
    exact = np.zeros(n, dtype=np.longdouble)
    np.cumsum(hx[:-1].astype(np.longdouble), out=exact[1:])
    absx = np.zeros(n)
    np.cumsum(np.abs(hx[:-1]), out=absx[1:])
    eps = np.finfo(np.float64).eps
    assert got[0] == 0.0
    dev_err = np.abs(got - exact).astype(np.float64)
    assert np.all(dev_err <= (n / 4096 + 256) * eps * absx + 1e-300)
    seq = orc.scan(hx)
    seq_err = np.abs(seq - exact).astype(np.float64)
    assert np.all(seq_err <= (np.arange(n) + 1.0) * eps * absx + 1e-300)


def _jacobi_grids(rpx, orc, N):
    L = rpx.lib()
    gA, gB = rpx.DeviceBuffer(8 * N * N), rpx.DeviceBuffer(8 * N * N)
    L.rpx_fill_grid_f64(gA.ptr, N, SEED, 0, 0.0, 1.0, 0.0, None)
    L.rpx_fill_grid_f64(gB.ptr, N, SEED, 1, 0.0, 1.0, 0.0, None)
    return gA, gB


@pytest.fixture(scope="module")
def jacobi_8192_ref(rpx, orc):
    """The two-sweep path and the oracle's Base_OpenMP after 50 time steps."""
    L = rpx.lib()
    N, ts = 8192, 50
    A0 = orc.fill_grid(N, SEED, 0)
    B0 = orc.fill_grid(N, SEED, 1)
    gA, gB = _jacobi_grids(rpx, orc, N)
    assert np.array_equal(gA.to_array(), A0) and np.array_equal(gB.to_array(), B0)
    L.rpx_jacobi2d_f64(gA.ptr, gB.ptr, N, ts, None, None)
    A2, B2 = gA.to_array(), gB.to_array()
    Ao, Bo = orc.jacobi2d(A0, B0, N, ts, orc.OMP)
    assert np.array_equal(A2, Ao), "two-sweep A differs from Base_OpenMP at 8192^2 x 50"
    assert np.array_equal(B2, Bo), "two-sweep B differs from Base_OpenMP at 8192^2 x 50"
    return Ao, Bo


def test_jacobi_bench_plan_8192_50(rpx, orc, jacobi_8192_ref):
    L = rpx.lib()
    N, depth, rows = 8192, 3, 16
    launches, last = 50 // depth, 50 - (50 // depth) * depth
    gA, gB = _jacobi_grids(rpx, orc, N)
    gQ = rpx.DeviceBuffer(8 * N * N)
    L.rpx_jacobi2d_copy_boundary_f64(gQ.ptr, gA.ptr, N, None)
    grids = [gA, gQ]
    skip_b = rpx.tuning(0, rows, 0, rpx.RPX_TUNE_JACOBI_SKIP_B)
    for _ in range(launches):
        L.rpx_jacobi2d_fused_launch_f64(grids[1].ptr, gB.ptr, grids[0].ptr, N, depth, skip_b, None)
        grids.reverse()
    L.rpx_jacobi2d_fused_launch_f64(grids[1].ptr, gB.ptr, grids[0].ptr, N, last, rpx.tuning(0, rows), None)
    grids.reverse()
    Ao, Bo = jacobi_8192_ref
    assert np.array_equal(grids[0].to_array(), Ao)
    assert np.array_equal(gB.to_array(), Bo)


def test_jacobi_fused_entry_8192_50(rpx, orc, jacobi_8192_ref):
    L = rpx.lib()
    N = 8192
    gA, gB = _jacobi_grids(rpx, orc, N)
    sb = L.rpx_jacobi2d_fused_scratch_bytes(N)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_jacobi2d_fused_f64(gA.ptr, gB.ptr, N, 50, scr.ptr, sb, None, None)
    Ao, Bo = jacobi_8192_ref
    assert np.array_equal(gA.to_array(), Ao)
    assert np.array_equal(gB.to_array(), Bo)


NE3, D3, Q3 = 262144, 4, 5


def test_mass3dpa_config3_bit_exact(rpx, orc):
    L = rpx.lib()
    B, _ = orc.fem_basis(D3, Q3)
    Bt = B.copy()  # read through the suite's literal Bt[q + D1D*d] indexing
    D = orc.fill_uniform(NE3 * Q3 ** 3, SEED, 2, 0.5, 2.0)
    X = orc.fill_uniform(NE3 * D3 ** 3, SEED, 3, -1.0, 1.0)
    Y0 = np.zeros(NE3 * D3 ** 3)
    bufs = [rpx.DeviceBuffer.from_array(a) for a in (B, Bt, D, X, Y0)]
    for _ in range(2):  # Y += per rep (MASS3DPA.hpp:338)
        L.rpx_mass3dpa_f64(*(b.ptr for b in bufs), NE3, D3, Q3, None, None)
    ref = orc.mass3dpa(B, Bt, D, X, Y0, NE3, D3, Q3, orc.OMP)
    ref = orc.mass3dpa(B, Bt, D, X, ref, NE3, D3, Q3, orc.OMP)
    assert np.array_equal(bufs[4].to_array(), ref)


def test_diffusion3dpa_config3_bit_exact(rpx, orc):
    L = rpx.lib()
    B, G = orc.fem_basis(D3, Q3)
    D = orc.fill_sym_tensor(NE3, Q3 ** 3, 6, SEED, 2)
    X = orc.fill_uniform(NE3 * D3 ** 3, SEED, 3, -1.0, 1.0)
    Y0 = np.zeros(NE3 * D3 ** 3)
    bufs = [rpx.DeviceBuffer.from_array(a) for a in (B, G, D, X, Y0)]
    dD = rpx.DeviceBuffer(8 * D.size)
    L.rpx_fill_sym_tensor_f64(dD.ptr, NE3, Q3 ** 3, 6, SEED, 2, None)
    assert np.array_equal(dD.to_array(), D)
    L.rpx_diffusion3dpa_f64(bufs[0].ptr, bufs[1].ptr, dD.ptr, bufs[3].ptr, bufs[4].ptr, NE3, D3, Q3, 1, None,
                            None)
    ref = orc.diffusion3dpa(B, G, D, X, Y0, NE3, D3, Q3, True, orc.OMP)
    assert np.array_equal(bufs[4].to_array(), ref)


def test_gemm_4096_sampled_rows(rpx):
    """FP64 MFMA within its forward-error bound and the mixed path (13
    moduli, b = 44 bits) within |alpha| 2^(e_i+f_j-b) nk (1 + 2^-b), where
    2^e_i and 2^f_j are the power-of-two scales of row i of A and column j
    of B, both against 80-bit host arithmetic on sampled rows (including the
    first and last)."""
    L = rpx.lib()
    n, alpha = 4096, 1.5
    dA, dB = rpx.DeviceBuffer(8 * n * n), rpx.DeviceBuffer(8 * n * n)
    _fill(L, dA, n * n, 0, -1.0, 1.0)
    _fill(L, dB, n * n, 1, -1.0, 1.0)
    Cf, Cm = rpx.DeviceBuffer(8 * n * n), rpx.DeviceBuffer(8 * n * n)
    L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, None, None)
    sb = L.rpx_gemm_mixed_scratch_bytes(n, n, n, 0)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_gemm_mixed_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, 0, scr.ptr, sb, None)
    b = L.rpx_gemm_mixed_bits(n, 0)
    assert b == 44
    A = dA.to_array().reshape(n, n)
    Bh = dB.to_array().reshape(n, n)
    rows = np.sort(np.random.default_rng(3).choice(n, 8, replace=False))
    rows = np.unique(np.concatenate([rows, [0, n - 1]]))
    ref = (A[rows].astype(np.longdouble) @ Bh.astype(np.longdouble)) * np.longdouble(alpha)
    scale = alpha * (np.abs(A[rows]) @ np.abs(Bh))
    eps = np.finfo(np.float64).eps
    cf = Cf.to_array().reshape(n, n)[rows]
    assert np.all(np.abs(cf - ref) <= 2 * (n + 2) * eps * scale)
    cm = Cm.to_array().reshape(n, n)[rows]
    _, e = np.frexp(np.abs(A[rows]).max(axis=1, keepdims=True))
    _, f = np.frexp(np.abs(Bh).max(axis=0, keepdims=True))
    bound = abs(alpha) * np.ldexp(1.0, e + f - b) * n * (1 + 2.0 ** -b) + 4 * eps * np.abs(ref.astype(np.float64))
    assert np.all(np.abs(cm - ref) <= bound)
    # and the per-entry relative error the bench reports
    assert float(np.max(np.abs(cm - ref) / scale)) <= 1e-12
