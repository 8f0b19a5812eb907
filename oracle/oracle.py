"""ctypes wrapper of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker and the timed CPU baseline.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

SEQ, OMP = 0, 1


class Params(C.Structure):
    _fields_ = [("input_mode", C.c_int32), ("npasses", C.c_int32), ("seed", C.c_uint64),
                ("size", C.c_int64), ("sizefact", C.c_double), ("repfact", C.c_double),
                ("checkrun_reps", C.c_int64), ("tsteps", C.c_int64), ("nk", C.c_int64),
                ("d1d", C.c_int32), ("q1d", C.c_int32), ("nthreads", C.c_int32), ("pad_", C.c_int32),
                ("alpha", C.c_double), ("beta", C.c_double), ("has_alpha", C.c_int32), ("has_beta", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("checksum", C.c_longdouble), ("checksum_hi", C.c_double), ("checksum_lo", C.c_double),
                ("time_total_s", C.c_double), ("reps", C.c_int64),
                ("problem_size", C.c_int64), ("bytes_per_rep", C.c_int64), ("flops_per_rep", C.c_int64),
                ("iweighted", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        i64, d, i32, u64 = C.c_int64, C.c_double, C.c_int, C.c_uint64
        sigs = {
            "orc_reset_init_count": (None, []), "orc_get_init_count": (i32, []),
            "orc_init_data": (None, [P, i64]), "orc_init_data_int": (None, [P, i64]),
            "orc_init_data_const": (None, [P, i64, d]), "orc_init_data_rand_sign": (None, [P, i64]),
            "orc_init_data_rand_value": (None, [P, i64]), "orc_init_scalar": (None, [dp]),
            "orc_checksum": (C.c_longdouble, [P, i64, d]), "orc_iweighted_sum": (d, [P, i64]),
            "orc_checksum_parts": (None, [P, i64, d, dp, dp]), "orc_checksum_int": (C.c_longdouble, [P, i64, d]),
            "orc_splitmix64_next": (u64, [C.POINTER(u64)]),
            "orc_fill_uniform": (None, [P, i64, u64, u64, d, d]),
            "orc_fill_sym_tensor": (None, [P, i64, i32, i32, u64, u64]),
            "orc_fem_basis": (None, [i32, i32, P, P]),
            "orc_copy": (None, [i32, P, P, i64]), "orc_mul": (None, [i32, P, P, d, i64]),
            "orc_add": (None, [i32, P, P, P, i64]), "orc_triad": (None, [i32, P, P, P, d, i64]),
            "orc_dot": (d, [i32, P, P, i64]), "orc_reduce_sum": (d, [i32, P, i64]),
            "orc_scan": (None, [i32, P, P, i64]), "orc_sort": (None, [P, i64]),
            "orc_jacobi2d": (None, [i32, P, P, i64, i64]),
            "orc_gemm": (None, [i32, P, P, P, d, d, i64, i64, i64]),
            "orc_mass3dpa": (None, [i32, P, P, P, P, P, i64, i32, i32]),
            "orc_diffusion3dpa": (None, [i32, P, P, P, P, P, i64, i32, i32, i32]),
            "orc_run": (i32, [C.c_char_p, C.c_char_p, C.POINTER(Params), C.POINTER(Result)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data


# ---- data utils ---------------------------------------------------------------

def reset_init_count():
    lib().orc_reset_init_count()


def init_count() -> int:
    return lib().orc_get_init_count()


def init_data(n: int) -> np.ndarray:
    a = np.empty(n)
    lib().orc_init_data(_p(a), n)
    return a


def init_data_int(n: int) -> np.ndarray:
    a = np.empty(n, dtype=np.int32)
    lib().orc_init_data_int(_p(a), n)
    return a


def init_data_const(n: int, v: float) -> np.ndarray:
    a = np.empty(n)
    lib().orc_init_data_const(_p(a), n, v)
    return a


def init_data_rand_sign(n: int) -> np.ndarray:
    a = np.empty(n)
    lib().orc_init_data_rand_sign(_p(a), n)
    return a


def init_data_rand_value(n: int) -> np.ndarray:
    a = np.empty(n)
    lib().orc_init_data_rand_value(_p(a), n)
    return a


def init_scalar() -> float:
    d = C.c_double()
    lib().orc_init_scalar(C.byref(d))
    return d.value


def checksum(a: np.ndarray, scale: float = 1.0) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_checksum(_p(a), a.size, scale)


def checksum_exact(a: np.ndarray, scale: float = 1.0):
    """The long-double checksum as an exact Fraction."""
    from fractions import Fraction
    a = np.ascontiguousarray(a, dtype=np.float64)
    hi, lo = C.c_double(), C.c_double()
    lib().orc_checksum_parts(_p(a), a.size, scale, C.byref(hi), C.byref(lo))
    return Fraction(hi.value) + Fraction(lo.value)


def iweighted_sum(a: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_iweighted_sum(_p(a), a.size)


def checksum_int(a: np.ndarray, scale: float = 1.0) -> float:
    a = np.ascontiguousarray(a, dtype=np.int32)
    return lib().orc_checksum_int(_p(a), a.size, scale)


def splitmix64(state: int, count: int) -> list[int]:
    s = C.c_uint64(state)
    return [lib().orc_splitmix64_next(C.byref(s)) for _ in range(count)]


def fill_uniform(n: int, seed: int, stream_id: int, lo: float, hi: float) -> np.ndarray:
    a = np.empty(n)
    lib().orc_fill_uniform(_p(a), n, seed, stream_id, lo, hi)
    return a


def fill_sym_tensor(ne: int, q3: int, sym: int, seed: int, stream_id: int) -> np.ndarray:
    a = np.empty(ne * q3 * sym)
    lib().orc_fill_sym_tensor(_p(a), ne, q3, sym, seed, stream_id)
    return a


def fem_basis(d1d: int, q1d: int):
    B = np.empty(q1d * d1d)
    G = np.empty(q1d * d1d)
    lib().orc_fem_basis(d1d, q1d, _p(B), _p(G))
    return B, G


def fill_grid(N: int, seed: int, stream_id: int) -> np.ndarray:
    g = fill_uniform(N * N, seed, stream_id, 0.0, 1.0).reshape(N, N)
    g[0, :] = 0.0
    g[-1, :] = 0.0
    g[:, 0] = 0.0
    g[:, -1] = 0.0
    return g.reshape(-1).copy()


# ---- kernels (in place on numpy arrays) ----------------------------------------

def copy(a, variant=SEQ):
    c = np.empty_like(a)
    lib().orc_copy(variant, _p(c), _p(a), a.size)
    return c


def mul(c, alpha, variant=SEQ):
    b = np.empty_like(c)
    lib().orc_mul(variant, _p(b), _p(c), alpha, c.size)
    return b


def add(a, b, variant=SEQ):
    c = np.empty_like(a)
    lib().orc_add(variant, _p(c), _p(a), _p(b), a.size)
    return c


def triad(b, c, alpha, variant=SEQ):
    a = np.empty_like(b)
    lib().orc_triad(variant, _p(a), _p(b), _p(c), alpha, b.size)
    return a


def dot(a, b, variant=SEQ) -> float:
    return lib().orc_dot(variant, _p(a), _p(b), a.size)


def reduce_sum(x, variant=SEQ) -> float:
    return lib().orc_reduce_sum(variant, _p(x), x.size)


def scan(x, variant=SEQ):
    y = np.empty_like(x)
    lib().orc_scan(variant, _p(x), _p(y), x.size)
    return y


def sort(x):
    y = np.array(x, dtype=np.float64, copy=True)
    lib().orc_sort(_p(y), y.size)
    return y


def jacobi2d(A, B, N, tsteps, variant=SEQ):
    A = A.copy()
    B = B.copy()
    lib().orc_jacobi2d(variant, _p(A), _p(B), N, tsteps)
    return A, B


def gemm(A, B, alpha, beta, ni, nj, nk, C0=None, variant=SEQ):
    Cm = np.zeros(ni * nj) if C0 is None else np.array(C0, dtype=np.float64, copy=True)
    lib().orc_gemm(variant, _p(Cm), _p(A), _p(B), alpha, beta, ni, nj, nk)
    return Cm


def mass3dpa(B, Bt, D, X, Y, NE, d1d, q1d, variant=SEQ):
    Y = np.array(Y, dtype=np.float64, copy=True)
    lib().orc_mass3dpa(variant, _p(B), _p(Bt), _p(D), _p(X), _p(Y), NE, d1d, q1d)
    return Y


def diffusion3dpa(Basis, dBasis, D, X, Y, NE, d1d, q1d, symmetric=True, variant=SEQ):
    Y = np.array(Y, dtype=np.float64, copy=True)
    lib().orc_diffusion3dpa(variant, _p(Basis), _p(dBasis), _p(D), _p(X), _p(Y), NE, d1d, q1d,
                            1 if symmetric else 0)
    return Y


# ---- suite-level -----------------------------------------------------------------

def run(kernel: str, variant: str = "Base_Seq", *, input_mode: int = 0, seed: int = 42, size: int = 0,
        sizefact: float = 1.0, repfact: float = 1.0, checkrun: int = 0, npasses: int = 1,
        tsteps: int = 0, nk: int = 0, d1d: int = 0, q1d: int = 0, nthreads: int = 0,
        alpha: float | None = None, beta: float | None = None) -> dict:
    p = Params(input_mode, npasses, seed, size, sizefact, repfact, checkrun, tsteps, nk, d1d, q1d, nthreads, 0,
               0.0 if alpha is None else alpha, 0.0 if beta is None else beta, alpha is not None, beta is not None)
    r = Result()
    rc = lib().orc_run(kernel.encode(), variant.encode(), C.byref(p), C.byref(r))
    if rc != 0:
        raise ValueError(f"oracle has no {variant} for {kernel}")
    return {"checksum": float(r.checksum), "checksum_hi": r.checksum_hi, "checksum_lo": r.checksum_lo, "time_s": r.time_total_s,
            "reps": r.reps, "problem_size": r.problem_size, "bytes_per_rep": r.bytes_per_rep,
            "flops_per_rep": r.flops_per_rep, "iweighted": r.iweighted}
