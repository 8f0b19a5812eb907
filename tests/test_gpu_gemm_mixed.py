"""GPU: mixed-precision GEMM, MASS3DPA and DIFFUSION3DPA on int8 matrix
cores with FP64 results.  GEMM (the modular scheme: one int8 GEMM per
modulus, FP64 reconstruction) is measured per element against an 80-bit
(long double) reference, both relative to |alpha| (|A||B|)_ij, the
conditioning of each entry, and against the stated bound
|alpha| 2^(e_i+f_j-b) nk (1 + 2^-b) from the row/column scales and the bit
count b the library reports; MASS3DPA and DIFFUSION3DPA against the FP64
oracle relative to the operator applied to absolute values.  The stated
tolerance is 1e-12 at the default settings."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def dev(rpx, a):
    return rpx.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64))


def run_mixed(rpx, A, B, alpha, ni, nj, nk, t=0):
    L = rpx.lib()
    sb = L.rpx_gemm_mixed_scratch_bytes(ni, nj, nk, t)
    assert sb > 0
    scr = rpx.DeviceBuffer(sb)
    dA, dB, dC = dev(rpx, A), dev(rpx, B), rpx.DeviceBuffer(8 * ni * nj)
    print(f"mixed GEMM {ni}x{nj}x{nk} moduli={t}", flush=True)
    L.rpx_gemm_mixed_f64(dC.ptr, dA.ptr, dB.ptr, alpha, 1.2, ni, nj, nk, t, scr.ptr, sb, None)
    return dC.to_array().reshape(ni, nj)


def exact(A, B, alpha, ni, nj, nk):
    return (A.reshape(ni, nk).astype(np.longdouble) @ B.reshape(nk, nj).astype(np.longdouble)) * np.longdouble(alpha)


def rel_err(C, A, B, alpha, ni, nj, nk):
    ref = exact(A, B, alpha, ni, nj, nk)
    scale = abs(alpha) * (np.abs(A.reshape(ni, nk)) @ np.abs(B.reshape(nk, nj)))
    return np.max(np.abs(C.astype(np.longdouble) - ref) / np.maximum(scale, 1e-300))


def stated_bound(rpx, A, B, alpha, ni, nj, nk, t=0):
    """|alpha| 2^(e_i+f_j-b) nk (1 + 2^-b), with 2^(e-1) <= max|row| < 2^e,
    plus the FP64 rounding of the result."""
    b = rpx.lib().rpx_gemm_mixed_bits(nk, t)
    assert b > 0
    _, e = np.frexp(np.abs(A.reshape(ni, nk)).max(axis=1))
    _, f = np.frexp(np.abs(B.reshape(nk, nj)).max(axis=0))
    scales = np.ldexp(1.0, np.add.outer(e, f) - b)
    ref = exact(A, B, alpha, ni, nj, nk)
    return abs(alpha) * scales * nk * (1 + 2.0 ** -b) + 4 * EPS * np.abs(ref.astype(np.float64)), b


@pytest.mark.parametrize("ni,nj,nk", [(64, 64, 64), (130, 70, 200), (257, 129, 1000), (512, 512, 512),
                                      (300, 520, 1)])
def test_mixed_gemm_meets_1e12(rpx, rng, ni, nj, nk):
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    C = run_mixed(rpx, A, B, 1.5, ni, nj, nk)
    assert rel_err(C, A, B, 1.5, ni, nj, nk) <= 1e-12
    bound, b = stated_bound(rpx, A, B, 1.5, ni, nj, nk)
    assert b >= 42
    assert np.all(np.abs(C - exact(A, B, 1.5, ni, nj, nk)) <= bound)


@pytest.mark.parametrize("ni,nj,nk", [(64, 96, 200), (300, 257, 513)])
def test_mixed_gemm_textbook_beta(rpx, rng, ni, nj, nk):
    """rpx_gemm_mixed_tuned_f64 with RPX_TUNE_GEMM_BETA: C = alpha*A*B +
    beta*C0 within the product's stated bound plus the FP64 rounding of the
    sum; without the flag (or through rpx_gemm_mixed_f64) C0 is ignored and
    both entry points agree bit for bit."""
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    C0 = rng.uniform(-1, 1, ni * nj)
    alpha, beta = 1.5, 1.2
    L = rpx.lib()
    sb = L.rpx_gemm_mixed_scratch_bytes(ni, nj, nk, 0)
    scr = rpx.DeviceBuffer(sb)
    dA, dB = dev(rpx, A), dev(rpx, B)
    print(f"mixed textbook GEMM {ni}x{nj}x{nk}", flush=True)

    def run(tune):
        dC = dev(rpx, C0)
        L.rpx_gemm_mixed_tuned_f64(dC.ptr, dA.ptr, dB.ptr, alpha, beta, ni, nj, nk, 0, tune, scr.ptr, sb, None)
        return dC.to_array().reshape(ni, nj)

    got = run(rpx.tuning(flags=rpx.RPX_TUNE_GEMM_BETA))
    c0 = C0.reshape(ni, nj).astype(np.longdouble)
    ref = exact(A, B, alpha, ni, nj, nk) + np.longdouble(beta) * c0
    bound, _ = stated_bound(rpx, A, B, alpha, ni, nj, nk)
    bound = bound + 4 * EPS * (abs(beta) * np.abs(C0.reshape(ni, nj)) + np.abs(ref.astype(np.float64)))
    assert np.all(np.abs(got.astype(np.longdouble) - ref) <= bound)
    plain = run(None)
    assert np.array_equal(plain, run_mixed(rpx, A, B, alpha, ni, nj, nk))
    assert np.all(np.abs(plain.astype(np.longdouble) - exact(A, B, alpha, ni, nj, nk))
                  <= stated_bound(rpx, A, B, alpha, ni, nj, nk)[0])


def test_mixed_gemm_wide_dynamic_range(rpx, rng):
    ni, nj, nk = 96, 80, 300
    A = rng.uniform(-1, 1, ni * nk) * np.exp2(rng.integers(-30, 30, ni * nk))
    B = rng.uniform(-1, 1, nk * nj) * np.exp2(rng.integers(-30, 30, nk * nj))
    A[:nk] = 0.0  # an all-zero row
    B[5::nj] = 0.0  # an all-zero column
    C = run_mixed(rpx, A, B, -0.62, ni, nj, nk)
    assert np.all(C[0] == 0.0) and np.all(C[:, 5] == 0.0)
    # entries far below their row/column maxima carry less absolute precision:
    # the bound is relative to the row and column scales, not to (|A||B|)_ij
    bound, _ = stated_bound(rpx, A, B, -0.62, ni, nj, nk)
    assert np.all(np.abs(C - exact(A, B, -0.62, ni, nj, nk)) <= bound)


def test_mixed_gemm_moduli_trade_accuracy(rpx, rng):
    ni = nj = nk = 128
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    ts = list(range(2, 17))
    bits = [rpx.lib().rpx_gemm_mixed_bits(nk, t) for t in ts]
    assert all(bits[i + 1] >= bits[i] for i in range(len(bits) - 1)) and bits[-1] == 46
    errs = []
    for t in ts:
        C = run_mixed(rpx, A, B, 1.0, ni, nj, nk, t)
        bound, _ = stated_bound(rpx, A, B, 1.0, ni, nj, nk, t)
        assert np.all(np.abs(C - exact(A, B, 1.0, ni, nj, nk)) <= bound), t
        errs.append(rel_err(C, A, B, 1.0, ni, nj, nk))
    # each modulus adds about 4 bits to A and to B up to the 46-bit cap
    assert all(errs[i + 1] < errs[i] for i in range(6))
    assert errs[-1] <= 1e-14


def test_mixed_gemm_nk_limit_with_extreme_residues(rpx):
    """A = B = 1 - 2^(7-b) makes every scaled entry 2^b - 128, whose residue
    modulo 256 is -128 (for B, times an odd y_0: still -128), so the first
    modulus's int32 sums reach 128 * 128 * nk.  At the largest accepted nk,
    131071, that is 2^31 - 16384; the result matches the 80-bit product, and
    one more nk is refused."""
    L = rpx.lib()
    ni = nj = 16
    nk = 131071
    assert 128 * 128 * nk < 2 ** 31 <= 128 * 128 * (nk + 1)
    b = L.rpx_gemm_mixed_bits(nk, 0)
    v = 1.0 - 2.0 ** (7 - b)
    A = np.full(ni * nk, v)
    B = np.full(nk * nj, v)
    C = run_mixed(rpx, A, B, 1.0, ni, nj, nk)
    ref = np.longdouble(nk) * np.longdouble(v) * np.longdouble(v)
    assert np.all(np.abs(C - ref) <= 4 * EPS * float(ref))
    assert L.rpx_gemm_mixed_scratch_bytes(ni, nj, nk + 1, 0) == 0
    dA, dB, dC = rpx.DeviceBuffer(8 * ni * (nk + 1)), rpx.DeviceBuffer(8 * nj * (nk + 1)), \
        rpx.DeviceBuffer(8 * ni * nj)
    scr = rpx.DeviceBuffer(1 << 20)
    with pytest.raises(rpx.RpxError) as ei:
        L.rpx_gemm_mixed_f64(dC.ptr, dA.ptr, dB.ptr, 1.0, 0.0, ni, nj, nk + 1, 0, scr.ptr, 1 << 20, None)
    assert ei.value.status == rpx.RPX_ERR_UNSUPPORTED


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


# ---- DIFFUSION3DPA on int8 MFMA ---------------------------------------------------

from test_oracle_kernels import _diffusion_dense, _diffusion_dense_full, _effective_tables  # noqa: E402


def _diff_mixed(rpx, bufs, NE, d1d, symmetric=True, nslices=7):
    L = rpx.lib()
    sb = L.rpx_diffusion3dpa_mixed_scratch_bytes(d1d, d1d + 1, nslices)
    assert sb > 0
    scr = rpx.DeviceBuffer(sb)
    L.rpx_diffusion3dpa_mixed_f64(*[b.ptr for b in bufs], NE, d1d, d1d + 1, 1 if symmetric else 0, nslices,
                                  scr.ptr, sb, None)


def _diff_magnitude(b, g, D, X, NE, d1d, symmetric):
    """sum_c |K_c|^T (|O| (|K| |X|))_c with the effective half-storage tables
    (Base_Seq write order), the magnitude bound of every output."""
    q1d = d1d + 1
    Bf, Gf, Bb, Gb = (np.abs(t) for t in _effective_tables(b, g, d1d, q1d))
    if symmetric:
        return _diffusion_dense(Bf, Gf, Bb, Gb, np.abs(D), np.abs(X), NE, d1d, q1d)
    assert np.array_equal(Bf, Bb) and np.array_equal(Gf, Gb)
    return _diffusion_dense_full(Bf, Gf, np.abs(D), np.abs(X), NE, d1d, q1d)


@pytest.mark.parametrize("d1d", [1, 2, 3, 4])
@pytest.mark.parametrize("bases", ["gll", "random"])
def test_diffusion3dpa_mixed_meets_1e12(rpx, orc, rng, d1d, bases):
    """Against the FP64 Base_Seq oracle, relative to the magnitude operator;
    NE = 37 leaves a partial workgroup of elements.  Random bases break the
    half-storage symmetry, so only the Base_Seq write order defines the
    operator; Y starts non-zero (Y +=)."""
    q1d = d1d + 1
    NE = 37
    if bases == "gll":
        b, g = orc.fem_basis(d1d, q1d)
    else:
        b, g = rng.uniform(-1, 1, q1d * d1d), rng.uniform(-1, 1, q1d * d1d)
    D = orc.fill_sym_tensor(NE, q1d ** 3, 6, 42, 5)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    Y0 = rng.uniform(-1, 1, NE * d1d ** 3)
    bufs = [dev(rpx, a) for a in (b, g, D, X, Y0)]
    _diff_mixed(rpx, bufs, NE, d1d)
    ref = orc.diffusion3dpa(b, g, D, X, Y0, NE, d1d, q1d)
    mag = _diff_magnitude(b, g, D, X, NE, d1d, True)
    err = np.max(np.abs(bufs[4].to_array() - ref) / (mag + np.spacing(np.abs(ref))))
    assert err <= 1e-12, err


@pytest.mark.parametrize("d1d", [3, 4])
def test_diffusion3dpa_mixed_full_tensor(rpx, orc, rng, d1d):
    # nine independent components per point (symmetric = 0)
    q1d = d1d + 1
    NE = 50
    b, g = orc.fem_basis(d1d, q1d)
    D9 = rng.uniform(-1, 1, NE * 9 * q1d ** 3)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    bufs = [dev(rpx, a) for a in (b, g, D9, X, np.zeros(NE * d1d ** 3))]
    _diff_mixed(rpx, bufs, NE, d1d, symmetric=False)
    ref = orc.diffusion3dpa(b, g, D9, X, np.zeros_like(X), NE, d1d, q1d, symmetric=False)
    mag = _diff_magnitude(b, g, D9, X, NE, d1d, False)
    err = np.max(np.abs(bufs[4].to_array() - ref) / mag)
    assert err <= 1e-12, err


def test_diffusion3dpa_mixed_config3_size_and_slices(rpx, orc):
    """NE = 262,144 at (4,5) with the seeded inputs: within 1e-12 of Base_OpenMP
    relative to the magnitude operator (taken on every 64th element); five
    slices trade accuracy for speed (about 8 * 2^-35)."""
    NE, d1d, q1d = 262144, 4, 5
    b, g = orc.fem_basis(d1d, q1d)
    D = orc.fill_sym_tensor(NE, q1d ** 3, 6, 42, 2)
    X = orc.fill_uniform(NE * d1d ** 3, 42, 3, -1.0, 1.0)
    zeros = np.zeros(NE * d1d ** 3)
    ref = orc.diffusion3dpa(b, g, D, X, zeros, NE, d1d, q1d, True, orc.OMP)
    sel = np.arange(0, NE, 64)
    Dsel = D.reshape(NE, -1)[sel].reshape(-1)
    Xsel = X.reshape(NE, -1)[sel].reshape(-1)
    mag = _diff_magnitude(b, g, Dsel, Xsel, sel.size, d1d, True).reshape(sel.size, -1)
    errs = {}
    for s in (7, 5):
        bufs = [dev(rpx, a) for a in (b, g, D, X, zeros)]
        _diff_mixed(rpx, bufs, NE, d1d, nslices=s)
        got = bufs[4].to_array().reshape(NE, -1)
        errs[s] = np.max(np.abs(got[sel] - ref.reshape(NE, -1)[sel]) / mag)
        assert np.all(np.isfinite(got))
    assert errs[7] <= 1e-12, errs
    assert errs[5] <= 1e-8, errs


def test_diffusion3dpa_mixed_rejects_bad_arguments(rpx):
    L = rpx.lib()
    assert L.rpx_diffusion3dpa_mixed_scratch_bytes(5, 6, 7) == 0
    assert L.rpx_diffusion3dpa_mixed_scratch_bytes(4, 4, 7) == 0
    assert L.rpx_diffusion3dpa_mixed_scratch_bytes(4, 5, 4) == 0
    bufs = [rpx.DeviceBuffer(8 * 6000) for _ in range(5)]
    sb = L.rpx_diffusion3dpa_mixed_scratch_bytes(4, 5, 7)
    scr = rpx.DeviceBuffer(sb)
    with pytest.raises(rpx.RpxError):
        L.rpx_diffusion3dpa_mixed_f64(*[b.ptr for b in bufs], 4, 4, 5, 1, 7, scr.ptr, sb - 256, None)
    with pytest.raises(rpx.RpxError):
        L.rpx_diffusion3dpa_mixed_f64(*[b.ptr for b in bufs], 4, 5, 6, 1, 7, scr.ptr, sb, None)
