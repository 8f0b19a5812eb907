"""Pins the oracle's restatement of the suite's data utilities
(src/common/DataUtils.cpp:477-653) and of the seeded generator against
independent sources: published SplitMix64 vectors, the C library's own
rand(), closed forms and exact rational arithmetic."""

import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest


def test_splitmix64_published_vectors(orc):
    # Vigna's splitmix64.c; first outputs for seed 1234567 (as tabulated by
    # the Rosetta Code SplitMix64 task) and for seed 0.
    assert orc.splitmix64(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]
    assert orc.splitmix64(0, 1) == [0xE220A8397B1DCDAF]


def test_uniform_ranges_and_determinism(orc):
    u = orc.fill_uniform(100000, 42, 3, 0.0, 1.0)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    # 53-bit grid: every value is an integer multiple of 2^-53
    assert np.all(np.ldexp(u, 53) == np.floor(np.ldexp(u, 53)))
    v = orc.fill_uniform(100000, 42, 3, -1.0, 1.0)
    assert np.array_equal(v, 2.0 * u - 1.0)  # exact affine map for [-1,1)
    assert not np.array_equal(u, orc.fill_uniform(100000, 42, 4, 0.0, 1.0))
    assert np.array_equal(u, orc.fill_uniform(100000, 42, 3, 0.0, 1.0))


def test_init_data_formula_and_parity(orc):
    orc.reset_init_count()
    a = orc.init_data(1000)
    b = orc.init_data(1000)
    i = np.arange(1000, dtype=np.float64)
    assert np.array_equal(a, 0.2 * (i + 1.1) / (i + 1.12345))
    assert np.array_equal(b, 0.1 * (i + 1.1) / (i + 1.12345))
    assert orc.init_count() == 2
    # every initialiser, constant and scalar included, advances the counter
    orc.init_data_const(10, 0.0)
    s = orc.init_scalar()
    assert orc.init_count() == 4
    assert s == 0.1 * 1.1 / 1.12345  # taken at count 3: odd -> factor 0.1


def test_rand_initialisers_match_glibc(orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(4793)
    RAND_MAX = 2147483647
    expect = np.array([libc.rand() / RAND_MAX for _ in range(500)])
    orc.reset_init_count()
    got = orc.init_data_rand_value(500)
    assert np.array_equal(got, expect)
    # signed variant uses the same sequence as the sign source
    orc.reset_init_count()
    orc.init_data_const(1, 0.0)  # odd count -> factor 0.1
    rs = orc.init_data_rand_sign(500)
    i = np.arange(500, dtype=np.float64)
    sign = np.where(expect < 0.5, -1.0, 1.0)
    assert np.array_equal(rs, sign * 0.1 * (i + 1.1) / (i + 1.12345))


def test_int_initialiser_outliers(orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(4793)
    RAND_MAX = 2147483647
    n = 1000
    seq = [libc.rand() / RAND_MAX for _ in range(n + 2)]
    expect = np.where(np.array(seq[:n]) < 0.5, -1, 1)
    expect[int(n * seq[n])] = -58
    expect[int(n * seq[n + 1])] = 19
    assert np.array_equal(orc.init_data_int(n), expect)


def _exact_checksum(x):
    total = Fraction(0)
    for j, v in enumerate(x):
        w = abs(math.sin(j + 1.0)) + 0.5
        total += Fraction(w) * Fraction(float(v))
    return total


def test_checksum_against_exact_sum(orc, rng):
    x = rng.uniform(-1, 1, 3000)
    x[::97] *= 1e12  # large dynamic range: Kahan compensation matters
    exact = _exact_checksum(x)
    got = orc.checksum_exact(x, 1.0)
    scale = sum(abs(Fraction(float(v))) for v in x)
    assert abs(got - exact) <= scale * Fraction(1, 2**60)


def test_checksum_small_hand_values(orc):
    x = np.array([1.0, 2.0, -3.0])
    w = [abs(math.sin(j + 1.0)) + 0.5 for j in range(3)]
    expect = (w[0] * 1 + w[1] * 2 - w[2] * 3) * 0.5
    assert abs(orc.checksum(x, 0.5) - expect) < 1e-14


def test_iweighted_sum_closed_forms_and_order(orc, rng):
    """sum_i (i+1) x[i] in FP64, index order: exact on integer-valued data
    (n(n+1)/2 for ones; a hand-weighted case), and equal bit for bit to the
    sequential left-to-right sum; a reordered (reversed) sum differs on data
    whose rounding depends on order, so the order is pinned."""
    n = 100_000
    assert orc.iweighted_sum(np.ones(n)) == n * (n + 1) / 2
    assert orc.iweighted_sum(np.array([3.0, -1.0, 2.0])) == 3.0 - 2.0 + 6.0
    assert orc.iweighted_sum(np.array([5.0])) == 5.0 and orc.iweighted_sum(np.array([])) == 0.0
    x = rng.uniform(-1, 1, 5000)
    x[::7] *= 1e10
    seq = 0.0
    for i, v in enumerate(x):
        seq += float(i + 1) * float(v)
    assert orc.iweighted_sum(x) == seq
    rev = 0.0
    for i in range(len(x) - 1, -1, -1):
        rev += float(i + 1) * float(x[i])
    assert rev != seq
    exact = sum(Fraction(i + 1) * Fraction(float(v)) for i, v in enumerate(x))
    mag = sum(Fraction(i + 1) * abs(Fraction(float(v))) for i, v in enumerate(x))
    assert abs(Fraction(seq) - exact) <= mag * len(x) * Fraction(1, 2**52)


def test_suite_run_reports_iweighted_sum(orc):
    """orc_run's i-weighted checksum is the sum over the checksummed arrays
    and passes: TRIAD's output array, DOT's scalar per pass."""
    r = orc.run("Stream_TRIAD", "Base_Seq", input_mode=1, size=1000, checkrun=1)
    b = orc.fill_uniform(1000, 42, 1, 0.0, 1.0)
    c = orc.fill_uniform(1000, 42, 2, 0.0, 1.0)
    assert r["iweighted"] == orc.iweighted_sum(orc.triad(b, c, 0.75))
    r1 = orc.run("Stream_DOT", "Base_Seq", input_mode=1, size=1000, checkrun=2, npasses=1)
    r2 = orc.run("Stream_DOT", "Base_Seq", input_mode=1, size=1000, checkrun=2, npasses=3)
    assert r1["iweighted"] == float(r1["checksum"])
    assert r2["iweighted"] == 3 * r1["iweighted"]


def test_fem_basis_is_lagrange_and_symmetric(orc):
    for d1d in range(2, 9):
        q1d = d1d + 1
        B, G = orc.fem_basis(d1d, q1d)
        Bm = B.reshape(d1d, q1d).T  # Bm[q, d]
        Gm = G.reshape(d1d, q1d).T
        # partition of unity and derivative of a constant
        assert np.allclose(Bm.sum(axis=1), 1.0, atol=1e-12)
        assert np.allclose(Gm.sum(axis=1), 0.0, atol=1e-10)
        # exact symmetry required by the half-storage tables
        assert np.array_equal(Bm, Bm[::-1, ::-1])
        assert np.array_equal(Gm, -Gm[::-1, ::-1])
        # interpolation of x^k (k < d1d) is exact at the quadrature points:
        # nodal values at the Gauss-Lobatto nodes come from a Vandermonde solve
        xq = np.polynomial.legendre.leggauss(q1d)[0] * 0.5 + 0.5
        assert np.allclose(np.sort(xq), xq)
        if d1d > 1:
            cheb = np.polynomial.legendre.Legendre.basis(d1d - 1).deriv().roots()
            xd = np.concatenate([[0.0], np.sort(cheb) * 0.5 + 0.5, [1.0]])
            for k in range(d1d):
                assert np.allclose(Bm @ xd ** k, xq ** k, atol=1e-11)
                dexp = k * xq ** (k - 1) if k > 0 else np.zeros(q1d)
                assert np.allclose(Gm @ xd ** k, dexp, atol=1e-9)


def _splitmix64_py(state, count):
    """SplitMix64 (Steele, Lea & Flood 2014), written from the published
    algorithm; the published vectors above pin the oracle's version too."""
    M = (1 << 64) - 1
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def _mix64_py(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


@pytest.mark.parametrize("sym", [6, 9])
def test_fill_sym_tensor_slots_and_ranges(orc, sym):
    # Component names give the diagonal slots: (11,12,13,22,23,33) for the
    # six-component half storage (DIFFUSION3DPA.hpp:384-392), row-major 3x3
    # for nine.  Diagonal entries are uniform in [0.5, 2], the others in
    # [-0.25, 0.25], each from the SplitMix64 value at its flat index.
    comps = ([(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)] if sym == 6
             else [(r, c) for r in (1, 2, 3) for c in (1, 2, 3)])
    diag = np.array([r == c for r, c in comps])
    ne, q3, seed, sid = 3, 8, 42, 7
    t = orc.fill_sym_tensor(ne, q3, sym, seed, sid)
    n = ne * sym * q3
    state0 = _mix64_py(seed ^ ((sid * 0xD1B54A32D192ED03) & ((1 << 64) - 1)))
    u = np.array([(s >> 11) * 2.0 ** -53 for s in _splitmix64_py(state0, n)])
    slot = (np.arange(n) // q3) % sym
    lo = np.where(diag[slot], 0.5, -0.25)
    width = np.where(diag[slot], 1.5, 0.5)
    assert np.array_equal(t, lo + width * u)
    # the same values as the plain uniform generator over the flat index
    uu = orc.fill_uniform(n, seed, sid, 0.0, 1.0)
    assert np.array_equal(u, uu)
    # ranges per slot on a large sample: the extremes approach both ends
    big = orc.fill_sym_tensor(200, 125, sym, seed, sid).reshape(200, sym, 125)
    for s_, is_diag in enumerate(diag):
        v = big[:, s_, :]
        lo_, hi_ = (0.5, 2.0) if is_diag else (-0.25, 0.25)
        assert v.min() >= lo_ and v.max() <= hi_, (s_, v.min(), v.max())
        assert v.min() < lo_ + 0.01 * (hi_ - lo_) and v.max() > hi_ - 0.01 * (hi_ - lo_)
    # a symmetric field of the nine-component layout: diagonal slots 0, 4, 8
    if sym == 9:
        assert [i for i, d in enumerate(diag) if d] == [0, 4, 8]
    else:
        assert [i for i, d in enumerate(diag) if d] == [0, 3, 5]
