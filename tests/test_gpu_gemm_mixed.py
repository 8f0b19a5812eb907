"""GPU: mixed-precision GEMM and MASS3DPA (exact int8 slice products on
MFMA, FP64 recombination).  GEMM error is measured per element against an
80-bit (long double) reference, relative to |alpha| (|A||B|)_ij, the
conditioning of each entry; MASS3DPA error against the FP64 oracle relative
to the operator applied to absolute values.  The stated tolerance is 1e-12
at the default 7 slices."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(rpx, a):
    return rpx.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64))


def run_mixed(rpx, A, B, alpha, ni, nj, nk, s=0):
    L = rpx.lib()
    sb = L.rpx_gemm_mixed_scratch_bytes(ni, nj, nk, s)
    scr = rpx.DeviceBuffer(sb)
    dA, dB, dC = dev(rpx, A), dev(rpx, B), rpx.DeviceBuffer(8 * ni * nj)
    L.rpx_gemm_mixed_f64(dC.ptr, dA.ptr, dB.ptr, alpha, 1.2, ni, nj, nk, s, scr.ptr, sb, None)
    return dC.to_array().reshape(ni, nj)


def rel_err(C, A, B, alpha, ni, nj, nk):
    Al = A.reshape(ni, nk).astype(np.longdouble)
    Bl = B.reshape(nk, nj).astype(np.longdouble)
    ref = (Al @ Bl) * np.longdouble(alpha)
    scale = abs(alpha) * (np.abs(A.reshape(ni, nk)) @ np.abs(B.reshape(nk, nj)))
    return np.max(np.abs(C.astype(np.longdouble) - ref) / np.maximum(scale, 1e-300))


@pytest.mark.parametrize("ni,nj,nk", [(64, 64, 64), (130, 70, 200), (257, 129, 1000), (512, 512, 512)])
def test_mixed_gemm_meets_1e12(rpx, rng, ni, nj, nk):
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    C = run_mixed(rpx, A, B, 1.5, ni, nj, nk)
    assert rel_err(C, A, B, 1.5, ni, nj, nk) <= 1e-12


def test_mixed_gemm_wide_dynamic_range(rpx, rng):
    ni, nj, nk = 96, 80, 300
    A = rng.uniform(-1, 1, ni * nk) * np.exp2(rng.integers(-30, 30, ni * nk))
    B = rng.uniform(-1, 1, nk * nj) * np.exp2(rng.integers(-30, 30, nk * nj))
    A[:nk] = 0.0  # an all-zero row
    C = run_mixed(rpx, A, B, -0.62, ni, nj, nk)
    assert np.all(C[0] == 0.0)
    # entries far below their row/column maxima carry less absolute precision;
    # the bound is relative to max_k|A[i,k]| max_k|B[k,j]| * nk
    rowmax = np.abs(A.reshape(ni, nk)).max(axis=1)
    colmax = np.abs(B.reshape(nk, nj)).max(axis=0)
    ref = (A.reshape(ni, nk).astype(np.longdouble) @ B.reshape(nk, nj).astype(np.longdouble)) * np.longdouble(-0.62)
    bound = 0.62 * nk * np.outer(rowmax, colmax) * 1e-13
    assert np.all(np.abs(C - ref) <= bound)


def test_mixed_gemm_slices_trade_accuracy(rpx, rng):
    ni = nj = nk = 128
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    errs = [rel_err(run_mixed(rpx, A, B, 1.0, ni, nj, nk, s), A, B, 1.0, ni, nj, nk) for s in (2, 3, 4, 5, 6, 7, 8)]
    assert all(errs[i + 1] < errs[i] for i in range(4))  # each slice gains ~7 bits until FP64 rounding
    for s, e in zip((2, 3, 4, 5, 6, 7, 8), errs):
        assert e <= 4 * (s + 3) * 2.0 ** (-7 * s) + 1e-15


def test_mixed_gemm_exact_on_small_integers(rpx, rng):
    ni, nj, nk = 64, 48, 96
    A = rng.integers(-8, 8, ni * nk).astype(np.float64)
    B = rng.integers(-8, 8, nk * nj).astype(np.float64)
    C = run_mixed(rpx, A, B, 1.0, ni, nj, nk)
    assert np.array_equal(C, A.reshape(ni, nk) @ B.reshape(nk, nj))


# ---- MASS3DPA on int8 matrix cores -------------------------------------------

MIXED_MASS_ORDERS = [(d, q) for d in range(1, 5) for q in (d, d + 1, d + 2) if q <= 5]


def _mass_mixed(rpx, bufs, NE, d1d, q1d, nslices=7):
    L = rpx.lib()
    sb = L.rpx_mass3dpa_mixed_scratch_bytes(d1d, q1d, nslices)
    scr = rpx.DeviceBuffer(sb)
    L.rpx_mass3dpa_mixed_f64(*[b.ptr for b in bufs], NE, d1d, q1d, nslices, scr.ptr, sb, None)


@pytest.mark.parametrize("d1d,q1d", MIXED_MASS_ORDERS)
def test_mass3dpa_mixed_meets_1e12(rpx, orc, rng, d1d, q1d):
    """Error relative to |Bt'|^3 (|D| . |B|^3 |X|), the magnitude bound of
    every output, against the FP64 Base_Seq oracle; NE = 67 leaves a partial
    workgroup of elements."""
    NE = 67
    B, Bt = rng.uniform(-1, 1, q1d * d1d), rng.uniform(-1, 1, q1d * d1d)
    D = rng.uniform(0.5, 2, NE * q1d ** 3)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    bufs = [dev(rpx, a) for a in (B, Bt, D, X, np.zeros(NE * d1d ** 3))]
    _mass_mixed(rpx, bufs, NE, d1d, q1d)
    zeros = np.zeros(NE * d1d ** 3)
    ref = orc.mass3dpa(B, Bt, D, X, zeros, NE, d1d, q1d)
    mag = orc.mass3dpa(np.abs(B), np.abs(Bt), D, np.abs(X), zeros, NE, d1d, q1d)
    err = np.max(np.abs(bufs[4].to_array() - ref) / mag)
    assert err <= 1e-12, err


def test_mass3dpa_mixed_accumulates_and_slices_trade_accuracy(rpx, orc, rng):
    NE, d1d, q1d = 100, 4, 5
    B, Bt = rng.uniform(-1, 1, 20), rng.uniform(-1, 1, 20)
    D = rng.uniform(0.5, 2, NE * 125)
    X = rng.uniform(-1, 1, NE * 64)
    Y0 = rng.uniform(-1, 1, NE * 64)
    zeros = np.zeros(NE * 64)
    mag = orc.mass3dpa(np.abs(B), np.abs(Bt), D, np.abs(X), zeros, NE, d1d, q1d)
    ref = orc.mass3dpa(B, Bt, D, X, Y0, NE, d1d, q1d)
    errs = {}
    for s in (3, 7):
        bufs = [dev(rpx, a) for a in (B, Bt, D, X, Y0)]
        _mass_mixed(rpx, bufs, NE, d1d, q1d, s)
        # Y += keeps Y0; the final FP64 add may round once more than the oracle's
        errs[s] = np.max(np.abs(bufs[4].to_array() - ref) / (mag + np.spacing(np.abs(ref))))
    assert errs[7] <= 1e-12
    assert 1e-9 < errs[3] <= 1e-5  # three 7-bit digits: about 6 * 2^-21


def test_mass3dpa_mixed_exact_on_all_ones(rpx):
    # The suite's reference inputs (MASS3DPA.cpp:76-84): every product is exact.
    NE, d1d, q1d = 1000, 4, 5
    bufs = [dev(rpx, a) for a in (np.ones(20), np.ones(20), np.ones(NE * 125), np.ones(NE * 64),
                                  np.zeros(NE * 64))]
    _mass_mixed(rpx, bufs, NE, d1d, q1d)
    assert np.all(bufs[4].to_array() == 8000.0)


def test_mass3dpa_mixed_rejects_orders_and_small_scratch(rpx):
    L = rpx.lib()
    assert L.rpx_mass3dpa_mixed_scratch_bytes(5, 6, 7) == 0
    bufs = [rpx.DeviceBuffer(8 * 1000) for _ in range(5)]
    sb = L.rpx_mass3dpa_mixed_scratch_bytes(4, 5, 7)
    scr = rpx.DeviceBuffer(sb)
    with pytest.raises(rpx.RpxError):
        L.rpx_mass3dpa_mixed_f64(*[b.ptr for b in bufs], 1, 5, 6, 7, scr.ptr, sb, None)
    with pytest.raises(rpx.RpxError):
        L.rpx_mass3dpa_mixed_f64(*[b.ptr for b in bufs], 1, 4, 5, 7, scr.ptr, sb - 256, None)
