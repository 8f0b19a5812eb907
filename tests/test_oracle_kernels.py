"""Pins the oracle's Base_Seq / Base_OpenMP kernels against closed forms,
invariants, exact integer arithmetic, numpy and assembled operators -- each
chosen so that a dropped term, a wrong index or sign, or a transposed operand
fails."""

import numpy as np
import pytest

from oracle import OMP, SEQ

VARIANTS = [SEQ, OMP]


@pytest.mark.parametrize("v", VARIANTS)
def test_stream_closed_forms(orc, rng, v):
    n = 10007
    a = rng.random(n)
    b = rng.random(n)
    alpha = 0.75
    assert np.array_equal(orc.copy(a, v), a)
    assert np.array_equal(orc.mul(a, alpha, v), alpha * a)
    assert np.array_equal(orc.add(a, b, v), a + b)
    assert np.array_equal(orc.triad(a, b, alpha, v), a + alpha * b)  # b + alpha*c roles
    # operand order: triad(b=1, c=2, alpha=0.5) is exactly 2, triad(b=2, c=1) is 2.5
    ones, twos = np.ones(5), 2 * np.ones(5)
    assert np.all(orc.triad(ones, twos, 0.5, v) == 2.0)
    assert np.all(orc.triad(twos, ones, 0.5, v) == 2.5)


@pytest.mark.parametrize("v", VARIANTS)
def test_dot_and_reduce_exact_on_integers(orc, rng, v):
    n = 100003
    a = rng.integers(-1000, 1000, n).astype(np.float64)
    b = rng.integers(-1000, 1000, n).astype(np.float64)
    assert orc.dot(a, b, v) == float(sum(int(x) * int(y) for x, y in zip(a, b)))
    assert orc.reduce_sum(a, v) == float(sum(int(x) for x in a))
    assert orc.dot(np.zeros(0), np.zeros(0), v) == 0.0


@pytest.mark.parametrize("v", VARIANTS)
def test_scan_exclusive_invariants(orc, rng, v):
    n = 50001
    x = rng.integers(-100, 100, n).astype(np.float64)
    y = orc.scan(x, v)
    assert y[0] == 0.0
    assert np.array_equal(np.diff(y), x[:-1])  # exact for integer-valued x
    expect = np.concatenate([[0.0], np.cumsum(x)[:-1]])
    assert np.array_equal(y, expect)


def test_scan_float_close_to_longdouble(orc, rng):
    x = rng.uniform(-1, 1, 20000)
    y = orc.scan(x, SEQ)
    ref = np.concatenate([[0], np.cumsum(x.astype(np.longdouble))[:-1]])
    assert np.max(np.abs(y - ref)) < 1e-12


def test_sort_matches_numpy(orc, rng):
    x = rng.uniform(-1, 1, 30000)
    x[:10] = [0.0, -0.0, 1.0, -1.0, 0.5, 0.5, -0.5, 1e-300, -1e-300, 0.0]
    y = orc.sort(x)
    assert np.array_equal(y, np.sort(x))
    assert np.all(np.diff(y) >= 0)


def _jacobi_numpy(A, B, N, tsteps):
    A = A.reshape(N, N).copy()
    B = B.reshape(N, N).copy()
    for _ in range(tsteps):
        B[1:-1, 1:-1] = 0.2 * ((((A[1:-1, 1:-1] + A[1:-1, :-2]) + A[1:-1, 2:]) + A[2:, 1:-1]) + A[:-2, 1:-1])
        A[1:-1, 1:-1] = 0.2 * ((((B[1:-1, 1:-1] + B[1:-1, :-2]) + B[1:-1, 2:]) + B[2:, 1:-1]) + B[:-2, 1:-1])
    return A.reshape(-1), B.reshape(-1)


@pytest.mark.parametrize("v", VARIANTS)
def test_jacobi_fixed_point_and_bruteforce(orc, rng, v):
    N = 8
    c = np.ones(N * N)
    A, B = orc.jacobi2d(c, c, N, 3, v)
    assert np.array_equal(A, c) and np.array_equal(B, c)  # constant field is a fixed point
    A0 = rng.random(N * N)
    B0 = rng.random(N * N)
    A, B = orc.jacobi2d(A0, B0, N, 4, v)
    An, Bn = _jacobi_numpy(A0, B0, N, 4)
    assert np.array_equal(A, An) and np.array_equal(B, Bn)
    # boundary untouched
    assert np.array_equal(A.reshape(N, N)[0], A0.reshape(N, N)[0])
    assert np.array_equal(B.reshape(N, N)[:, -1], B0.reshape(N, N)[:, -1])


@pytest.mark.parametrize("v", VARIANTS)
def test_gemm_semantics(orc, rng, v):
    ni, nj, nk = 17, 13, 19
    A = rng.uniform(-1, 1, ni * nk)
    B = rng.uniform(-1, 1, nk * nj)
    C0 = rng.uniform(-1, 1, ni * nj)
    Cm = orc.gemm(A, B, 1.5, 1.2, ni, nj, nk, C0=C0, variant=v)
    ref = 1.5 * (A.reshape(ni, nk) @ B.reshape(nk, nj))
    assert np.allclose(Cm.reshape(ni, nj), ref, rtol=0, atol=1e-13)
    # beta never reaches the result (C is overwritten after C *= beta)
    assert np.array_equal(Cm, orc.gemm(A, B, 1.5, 7.0, ni, nj, nk, C0=-C0, variant=v))
    # alpha = 1 with identity B returns A exactly
    Id = np.eye(nk).reshape(-1)
    assert np.array_equal(orc.gemm(A, Id, 1.0, 0.0, ni, nk, nk, variant=v), A)


def _mass_dense(B, Bt, D, X, NE, d1d, q1d):
    """Textbook tensor-product form: Y_e = (Bt'(x)Bt'(x)Bt') diag(D_e) (B(x)B(x)B) X_e,
    with B[q,d] = B[q + q1d*d] and the literal Bt'[d,q] = Bt[q + d1d*d]."""
    Bm = B.reshape(d1d, q1d).T
    Btm = np.array([[Bt[q + d1d * d] for q in range(q1d)] for d in range(d1d)])
    Y = np.zeros(NE * d1d ** 3)
    for e in range(NE):
        Xe = X[e * d1d ** 3:(e + 1) * d1d ** 3].reshape(d1d, d1d, d1d)  # [dz,dy,dx]
        De = D[e * q1d ** 3:(e + 1) * q1d ** 3].reshape(q1d, q1d, q1d)  # [qz,qy,qx]
        Q = np.einsum("ZD,YE,XF,DEF->ZYX", Bm, Bm, Bm, Xe) * De
        Ye = np.einsum("DZ,EY,FX,ZYX->DEF", Btm, Btm, Btm, Q)
        Y[e * d1d ** 3:(e + 1) * d1d ** 3] = Ye.reshape(-1)
    return Y


@pytest.mark.parametrize("v", VARIANTS)
@pytest.mark.parametrize("d1d,q1d", [(4, 5), (2, 3), (3, 3), (5, 7)])
def test_mass3dpa_against_dense(orc, rng, v, d1d, q1d):
    NE = 3
    B = rng.uniform(-1, 1, q1d * d1d)
    Bt = rng.uniform(-1, 1, q1d * d1d)
    D = rng.uniform(0.5, 2, NE * q1d ** 3)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    Y0 = rng.uniform(-1, 1, NE * d1d ** 3)
    Y = orc.mass3dpa(B, Bt, D, X, Y0, NE, d1d, q1d, variant=v)
    assert np.allclose(Y - Y0, _mass_dense(B, Bt, D, X, NE, d1d, q1d), rtol=0, atol=1e-12)


def test_mass3dpa_all_ones_closed_form(orc):
    # Reference inputs (all ones): every entry gains D1D^3 * Q1D^3 per rep.
    r = orc.run("Apps_MASS3DPA", size=125 * 20, checkrun=3)
    expect = np.longdouble(0)
    n = 20 * 64
    w = np.abs(np.sin(np.arange(1, n + 1, dtype=np.float64))) + 0.5
    expect = float(np.sum(w.astype(np.longdouble) * 3 * 8000))
    assert abs(r["checksum"] - expect) <= 1e-12 * expect


def _effective_tables(b, g, d1d, q1d):
    """Simulates the half-storage shared table writes in the Base_Seq order
    (dy outer, qx inner for B/G; d outer, q inner for Bt/Gt) as a memory of
    q1d*d1d cells and returns what each (q,d) read sees, sign applied."""
    mem = {}

    def maps(q, d):
        if q <= d:
            return (q, d), (q1d - 1 - q, d1d - 1 - d), -1.0
        return (q1d - 1 - q, d1d - 1 - d), (q, d), 1.0

    for d in range(d1d):
        for q in range(q1d):
            (i, j), (k, l), s = maps(q, d)
            mem[i * d1d + j] = b[q + q1d * d]
            mem[k * d1d + l] = g[q + q1d * d] * s
    Bf = np.zeros((q1d, d1d))
    Gf = np.zeros((q1d, d1d))
    for d in range(d1d):
        for q in range(q1d):
            (i, j), (k, l), s = maps(q, d)
            Bf[q, d] = mem[i * d1d + j]
            Gf[q, d] = mem[k * d1d + l] * s
    for d in range(d1d):
        for q in range(q1d):
            (i, j), (k, l), s = maps(q, d)
            mem[j * q1d + i] = b[q + q1d * d]
            mem[l * q1d + k] = g[q + q1d * d] * s
    Bb = np.zeros((q1d, d1d))
    Gb = np.zeros((q1d, d1d))
    for d in range(d1d):
        for q in range(q1d):
            (i, j), (k, l), s = maps(q, d)
            Bb[q, d] = mem[j * q1d + i]
            Gb[q, d] = mem[l * q1d + k] * s
    return Bf, Gf, Bb, Gb


def _diffusion_dense(Bf, Gf, Bb, Gb, D, X, NE, d1d, q1d):
    Y = np.zeros(NE * d1d ** 3)
    for e in range(NE):
        Xe = X[e * d1d ** 3:(e + 1) * d1d ** 3].reshape(d1d, d1d, d1d)  # [z,y,x]
        De = D[e * 6 * q1d ** 3:(e + 1) * 6 * q1d ** 3].reshape(6, q1d, q1d, q1d)
        gx = np.einsum("ZD,YE,XF,DEF->ZYX", Bf, Bf, Gf, Xe)
        gy = np.einsum("ZD,YE,XF,DEF->ZYX", Bf, Gf, Bf, Xe)
        gz = np.einsum("ZD,YE,XF,DEF->ZYX", Gf, Bf, Bf, Xe)
        O11, O12, O13, O22, O23, O33 = De
        q0 = O11 * gx + O12 * gy + O13 * gz
        q1 = O12 * gx + O22 * gy + O23 * gz
        q2 = O13 * gx + O23 * gy + O33 * gz
        Ye = (np.einsum("ZD,YE,XF,ZYX->DEF", Bb, Bb, Gb, q0)
              + np.einsum("ZD,YE,XF,ZYX->DEF", Bb, Gb, Bb, q1)
              + np.einsum("ZD,YE,XF,ZYX->DEF", Gb, Bb, Bb, q2))
        Y[e * d1d ** 3:(e + 1) * d1d ** 3] = Ye.reshape(-1)
    return Y


@pytest.mark.parametrize("v", VARIANTS)
@pytest.mark.parametrize("d1d", [2, 3, 4, 6])
def test_diffusion3dpa_symmetric_bases_equal_textbook(orc, rng, v, d1d):
    q1d = d1d + 1
    NE = 3
    b, g = orc.fem_basis(d1d, q1d)
    D = orc.fill_sym_tensor(NE, q1d ** 3, 6, 42, 7)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    Y = orc.diffusion3dpa(b, g, D, X, np.zeros_like(X), NE, d1d, q1d, variant=v)
    Bm = b.reshape(d1d, q1d).T
    Gm = g.reshape(d1d, q1d).T
    assert np.allclose(Y, _diffusion_dense(Bm, Gm, Bm, Gm, D, X, NE, d1d, q1d), rtol=0, atol=1e-11)
    # symmetric positive semi-definite operator: x^T A x >= 0
    assert all(np.dot(X[e * d1d ** 3:(e + 1) * d1d ** 3], Y[e * d1d ** 3:(e + 1) * d1d ** 3]) > -1e-12
               for e in range(NE))


def _diffusion_dense_full(Bm, Gm, D9, X, NE, d1d, q1d):
    """Textbook Y_e = sum_r sum_c G_r^T O_rc G_c X_e with O the full 3x3 tensor
    of each quadrature point given row-major (component 3r+c is O_{r+1,c+1}),
    G_x = B (x) B (x) G' etc.; nothing assumes O symmetric."""
    Y = np.zeros(NE * d1d ** 3)
    for e in range(NE):
        Xe = X[e * d1d ** 3:(e + 1) * d1d ** 3].reshape(d1d, d1d, d1d)  # [z,y,x]
        O = D9[e * 9 * q1d ** 3:(e + 1) * 9 * q1d ** 3].reshape(3, 3, q1d, q1d, q1d)
        grads = [np.einsum("ZD,YE,XF,DEF->ZYX", Bm, Bm, Gm, Xe),   # d/dx
                 np.einsum("ZD,YE,XF,DEF->ZYX", Bm, Gm, Bm, Xe),   # d/dy
                 np.einsum("ZD,YE,XF,DEF->ZYX", Gm, Bm, Bm, Xe)]   # d/dz
        flux = [sum(O[r, c] * grads[c] for c in range(3)) for r in range(3)]
        Ye = (np.einsum("ZD,YE,XF,ZYX->DEF", Bm, Bm, Gm, flux[0])
              + np.einsum("ZD,YE,XF,ZYX->DEF", Bm, Gm, Bm, flux[1])
              + np.einsum("ZD,YE,XF,ZYX->DEF", Gm, Bm, Bm, flux[2]))
        Y[e * d1d ** 3:(e + 1) * d1d ** 3] = Ye.reshape(-1)
    return Y


@pytest.mark.parametrize("v", VARIANTS)
@pytest.mark.parametrize("d1d", [2, 3, 4])
def test_diffusion3dpa_full_tensor_equals_textbook(orc, rng, v, d1d):
    # symmetric=0 reads nine components per point (DIFFUSION3DPA.hpp:384-392).
    # A tensor with independent entries in every slot, O12 != O21 etc., makes
    # a transposed, swapped or symmetrised component read change Y.
    q1d = d1d + 1
    NE = 2
    b, g = orc.fem_basis(d1d, q1d)
    D9 = rng.uniform(-1, 1, NE * 9 * q1d ** 3)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    Y = orc.diffusion3dpa(b, g, D9, X, np.zeros_like(X), NE, d1d, q1d, symmetric=False, variant=v)
    Bm = b.reshape(d1d, q1d).T
    Gm = g.reshape(d1d, q1d).T
    ref = _diffusion_dense_full(Bm, Gm, D9, X, NE, d1d, q1d)
    assert np.allclose(Y, ref, rtol=0, atol=1e-11)
    # the transposed tensor is a different operator here, so the check binds
    D9t = D9.reshape(NE, 3, 3, -1).transpose(0, 2, 1, 3).reshape(-1)
    assert not np.allclose(ref, _diffusion_dense_full(Bm, Gm, D9t, X, NE, d1d, q1d), rtol=0, atol=1e-6)
    # a symmetric 9-component tensor gives the 6-component result
    Dsym = 0.5 * (D9 + D9t)
    D6 = Dsym.reshape(NE, 9, -1)[:, [0, 1, 2, 4, 5, 8], :].reshape(-1)
    Y9 = orc.diffusion3dpa(b, g, Dsym, X, np.zeros_like(X), NE, d1d, q1d, symmetric=False, variant=v)
    Y6 = orc.diffusion3dpa(b, g, D6, X, np.zeros_like(X), NE, d1d, q1d, symmetric=True, variant=v)
    assert np.allclose(Y9, Y6, rtol=0, atol=1e-12)


@pytest.mark.parametrize("d1d", [3, 4])
def test_diffusion3dpa_all_ones_follows_seq_write_order(orc, rng, d1d):
    q1d = d1d + 1
    NE = 2
    b = np.ones(q1d * d1d)
    g = np.ones(q1d * d1d)
    D = np.ones(NE * 6 * q1d ** 3)
    X = rng.uniform(-1, 1, NE * d1d ** 3)
    Y = orc.diffusion3dpa(b, g, D, X, np.zeros_like(X), NE, d1d, q1d)
    tabs = _effective_tables(b, g, d1d, q1d)
    assert np.allclose(Y, _diffusion_dense(*tabs, D, X, NE, d1d, q1d), rtol=0, atol=1e-11)


def test_suite_run_shapes_match_reference_formulas(orc):
    r = orc.run("Polybench_GEMM", size=4096 * 4096, checkrun=1, nk=64)
    assert r["problem_size"] == 4096 * 4096
    r = orc.run("Polybench_JACOBI_2D", size=30 * 30, checkrun=1)
    assert r["problem_size"] == 30 * 30  # N = 32
    r = orc.run("Apps_MASS3DPA", checkrun=1, size=262144 * 125 // 64)
    assert r["problem_size"] % 125 == 0
    with pytest.raises(ValueError):
        orc.run("Algorithm_SORT", "Base_OpenMP", size=100, checkrun=1)


def test_seq_and_openmp_checksums_agree(orc):
    for k in ["Stream_TRIAD", "Stream_DOT", "Algorithm_REDUCE_SUM", "Algorithm_SCAN",
              "Polybench_JACOBI_2D", "Apps_MASS3DPA", "Apps_DIFFUSION3DPA"]:
        s = orc.run(k, "Base_Seq", size=20000, checkrun=2, input_mode=1)
        o = orc.run(k, "Base_OpenMP", size=20000, checkrun=2, input_mode=1)
        assert abs(s["checksum"] - o["checksum"]) <= 1e-7, k  # the suite's own test tolerance
