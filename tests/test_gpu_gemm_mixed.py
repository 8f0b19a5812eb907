"""GPU: mixed-precision GEMM (exact int8 slice products on MFMA, FP64
recombination).  Error is measured per element against an 80-bit
(long double) reference, relative to |alpha| (|A||B|)_ij, the conditioning
of each entry; the stated tolerance is 1e-12 at the default 7 slices."""

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
