#!/usr/bin/env python3
"""Benchmark of the MI355X variant on the HBM-bound FP64 workload.

Workload (one GPU, synthetic seeded inputs, seed 42, resident in HBM):
  Stream_TRIAD       N = 2^28 doubles (a, b, c: 2 GiB each)
  Stream_DOT         N = 2^28
  Algorithm_REDUCE_SUM  N = 2^28, x uniform in [-1,1)
  Algorithm_SCAN     N = 2^28 (exclusive), x uniform in [-1,1)
  Polybench_JACOBI_2D  8192 x 8192 grid, interior uniform [0,1), zero
                     boundary, 50 time steps (100 sweeps)
One step is one rep of each kernel, in that order, on one HIP stream.

`value` is the step's aggregate HBM throughput: the bytes each kernel must
move (its inputs read once, its outputs written once; for JACOBI_2D, whose
50 time steps run as 17 temporally blocked launches, each launch reads A and
writes the new A, and the last also writes B) divided by the step time.
`suite_algorithmic_GBps` is the suite's own rate definition (bytes per rep
summed, JACOBI_2D counted per sweep as POLYBENCH_JACOBI_2D.cpp:37-42);
temporal blocking makes it exceed the HBM peak, so it is not a roofline.

The roofline object describes the dominant kernel, the fused JACOBI_2D
launch (three time steps, six sweeps, B not stored): its compulsory HBM
bytes (read A, write the new A) over its average launch time, measured with
HIP events on the launch stream, against the 8 TB/s peak.  Its executed
FP64 work (halo recompute included) is stated beside it and is far below
the FP64 peak's time, so HBM is the bound.  `traffic` is the launch's HBM
bytes from the committed rocprofv3 counter summary.  Every kernel is also
given as a fraction of a COPY of 2^28 doubles timed in the same run (the
measured streaming ceiling).

After timing, the outputs are checked: TRIAD and SCAN on slices at both
ends, DOT and REDUCE_SUM against host sums, and the final JACOBI_2D grid
against the two-sweep path.

The CPU baseline times the oracle's Base_Seq (one thread) and Base_OpenMP
restatements of the same kernels, in this run, on a stated fraction of the
workload, and scales each to a full step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6            # MI355X FP64 vector and matrix peak
SEED = 42
N_STREAM = 1 << 28
N_GRID = 8192
TSTEPS = 50
DEPTH = 3                          # JACOBI_2D time steps per fused launch
ROWS = 99                          # output rows per wavefront chunk (streaming fused kernel)
LAUNCHES = TSTEPS // DEPTH         # 16 launches of DEPTH steps that store no B ...
LAST_DEPTH = TSTEPS - LAUNCHES * DEPTH  # ... then one of 2 steps that stores the final B


def algorithmic_bytes():
    n, N = N_STREAM, N_GRID
    sweep = 8 * (N * N - 4) + 8 * (N - 2) * (N - 2)  # POLYBENCH_JACOBI_2D.cpp:37-42 per sweep
    return {
        "Stream_TRIAD": 24 * n,                   # TRIAD.cpp:30-32
        "Stream_DOT": 8 + 16 * n + 8,             # DOT.cpp:31-35
        "Algorithm_REDUCE_SUM": 8 * (1 + n) + 8,  # REDUCE_SUM.cpp:31-33
        "Algorithm_SCAN": 16 * n,                 # SCAN.cpp:31-32
        "Polybench_JACOBI_2D": 2 * TSTEPS * sweep,
        "_jacobi_sweep": sweep,
    }


def compulsory_bytes():
    """HBM bytes each kernel must move under this launch plan: every input
    read once and every output written once.  A fused JACOBI_2D launch reads
    A and writes the new A (2*8*N^2); the last launch of a step also writes
    B (8*N^2)."""
    n, N = N_STREAM, N_GRID
    launch = 2 * 8 * N * N
    return {
        "Stream_TRIAD": 24 * n,
        "Stream_DOT": 16 * n + 8,
        "Algorithm_REDUCE_SUM": 8 * n + 8,
        "Algorithm_SCAN": 16 * n,
        "Polybench_JACOBI_2D": (LAUNCHES + 1) * launch + 8 * N * N,
        "_jacobi_launch": launch,
    }


def fused_launch_flops(N: int, rows: int, depth: int) -> int:
    """FP64 operations one fused JACOBI_2D launch executes, halo recompute
    included, 5 per point, on 128-column strips of 112 owned columns.
    Tiles (rows < 48): a wavefront runs steps 2 .. rows+4*depth-1 and level
    L computes rows + 4*depth - 2L rows.  Streaming chunks (rows >= 48; the
    first and last strips in chunks of 24 rows): steps 2 .. Q-1 with Q the
    step count rounded up to the 9-step blocks past the first 18, level L
    from step 2L."""
    H = 2 * depth
    strips = -(-(N - 1) // 112)
    if rows < 48:
        chunks = -(-(N - 2) // rows)
        nr = rows + 2 * H
        return strips * chunks * sum(nr - 2 * L for L in range(1, H + 1)) * 128 * 5

    def strip_levels(r):  # level rows of one strip cut into chunks of r rows
        total = 0
        for c in range(-(-(N - 2) // r)):
            nr = min(r, N - 2 - c * r) + 2 * H
            q = 18 + 9 * max(0, -(-(nr - 18) // 9))
            total += sum(q - 2 * L for L in range(1, H + 1))
        return total
    edge = min(strips, 2)
    return (edge * strip_levels(24) + (strips - edge) * strip_levels(rows)) * 128 * 5


def read_traffic(kernel_key: str):
    """HBM bytes per dispatch for a kernel from the committed counter summary
    (profiles/*counters*.txt lines `kernel=<key> hbm_bytes_per_dispatch=<B>`)."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if "counters" not in fn or not fn.endswith(".txt"):
            continue
        for line in open(os.path.join(pdir, fn)):
            m = re.match(r"\s*kernel=(\S+)\s+hbm_bytes_per_dispatch=([0-9.eE+]+)", line)
            if m and m.group(1) == kernel_key:
                return float(m.group(2))
    return None


def gpu_run(steps: int, warmup: int):
    from rajaperf_amd import abi
    arch = abi.init(0)
    L = abi.lib()
    n, N = N_STREAM, N_GRID
    s = None  # null stream: every launch and event below is on it
    B = abi.DeviceBuffer
    a, b, c = B(8 * n), B(8 * n), B(8 * n)
    x, y = B(8 * n), B(8 * n)
    gA, gB, gQ = B(8 * N * N), B(8 * N * N), B(8 * N * N)
    res = B(16)
    L.rpx_fill_uniform_f64(b.ptr, n, SEED, 1, 0.0, 1.0, s)
    L.rpx_fill_uniform_f64(c.ptr, n, SEED, 2, 0.0, 1.0, s)
    L.rpx_fill_const_f64(a.ptr, n, 0.0, s)
    L.rpx_fill_uniform_f64(x.ptr, n, SEED, 0, -1.0, 1.0, s)
    L.rpx_fill_grid_f64(gA.ptr, N, SEED, 0, 0.0, 1.0, 0.0, s)
    L.rpx_fill_grid_f64(gB.ptr, N, SEED, 1, 0.0, 1.0, 0.0, s)
    L.rpx_fill_const_f64(res.ptr, 2, 0.0, s)
    # The new-A grid of each fused launch ping-pongs between gA and gQ; the
    # boundary never changes, so it is copied once.
    L.rpx_jacobi2d_copy_boundary_f64(gQ.ptr, gA.ptr, N, s)
    grids = [gA.ptr, gQ.ptr]
    rsb = L.rpx_reduce_scratch_bytes(n, None)
    rscr = B(rsb)
    L.rpx_reduce_scratch_init(rscr.ptr, rsb, s)
    ssb = L.rpx_scan_scratch_bytes(n, None)
    sscr = B(ssb)
    L.rpx_scan_scratch_init(sscr.ptr, ssb, s)
    abi.synchronize()

    names = ["Stream_TRIAD", "Stream_DOT", "Algorithm_REDUCE_SUM", "Algorithm_SCAN", "Polybench_JACOBI_2D"]
    # One event before each kernel of each timed step, one after the JACOBI_2D
    # launches that store no B, and one at the end.
    evs = [[abi.Event() for _ in range(len(names) + 2)] for _ in range(steps)]
    skip_b = abi.tuning(0, ROWS, 0, abi.RPX_TUNE_JACOBI_SKIP_B)
    last = abi.tuning(0, 0)  # the library's default for a launch that stores B: 16-row tiles

    def step(ev):
        if ev:
            ev[0].record(s)
        L.rpx_triad_f64(a.ptr, b.ptr, c.ptr, 0.75, n, None, s)
        if ev:
            ev[1].record(s)
        L.rpx_dot_f64(a.ptr, b.ptr, n, res.ptr, 1, rscr.ptr, rsb, None, s)
        if ev:
            ev[2].record(s)
        L.rpx_reduce_sum_f64(x.ptr, n, res.ptr + 8, rscr.ptr, rsb, None, s)
        if ev:
            ev[3].record(s)
        L.rpx_scan_exclusive_f64(x.ptr, y.ptr, n, sscr.ptr, ssb, None, s)
        if ev:
            ev[4].record(s)
        # 50 time steps: 16 launches of 3 that leave B unwritten (each of
        # their B's is overwritten before anything reads it), then one of 2
        # that stores the final B.
        for _ in range(LAUNCHES):
            L.rpx_jacobi2d_fused_launch_f64(grids[1], gB.ptr, grids[0], N, DEPTH, skip_b, s)
            grids.reverse()
        if ev:
            ev[5].record(s)
        L.rpx_jacobi2d_fused_launch_f64(grids[1], gB.ptr, grids[0], N, LAST_DEPTH, last, s)
        grids.reverse()
        if ev:
            ev[6].record(s)

    for _ in range(warmup):
        step(None)
    abi.synchronize()
    t_start, t_stop = abi.Event(), abi.Event()
    wall0 = time.perf_counter()
    t_start.record(s)
    for k in range(steps):
        step(evs[k])
    t_stop.record(s)
    abi.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_ms(t_stop)
    per_kernel_ms = {k: 0.0 for k in names}
    skip_b_ms = 0.0
    for ev in evs:
        for i, k in enumerate(names[:-1]):
            per_kernel_ms[k] += ev[i].elapsed_ms(ev[i + 1])
        per_kernel_ms["Polybench_JACOBI_2D"] += ev[4].elapsed_ms(ev[6])
        skip_b_ms += ev[4].elapsed_ms(ev[5])
    if L.rpx_scan_scratch_status(sscr.ptr) != 0:
        raise RuntimeError("scan look-back wait expired")
    reps = warmup + steps
    checks = verify(L, a, b, c, x, y, res, grids[0], gB, reps, rscr, rsb)
    copy_ms = stream_ceiling(L, y, x)
    return {"arch": arch, "total_ms": total_ms, "per_kernel_ms": per_kernel_ms, "skip_b_ms": skip_b_ms,
            "wall": wall, "checks": checks, "copy_ms": copy_ms}


def verify(L, a, b, c, x, y, res, gA_final, gB_final, reps, rscr, rsb) -> dict:
    """Checks the timed run's outputs; raises if one is wrong."""
    from rajaperf_amd import abi
    n, N = N_STREAM, N_GRID
    eps = np.finfo(np.float64).eps
    out = {}
    k = 1 << 16
    # TRIAD, bit-exact on both ends (a = b + alpha*c, no contraction)
    for off in (0, n - k):
        hb, hc, ha = (buf.to_array(count=k, offset_bytes=8 * off) for buf in (b, c, a))
        if not np.array_equal(ha, hb + 0.75 * hc):
            raise RuntimeError(f"TRIAD output differs at elements [{off}, {off + k})")
    out["Stream_TRIAD"] = "bit-exact on the first and last 2^16 elements"
    hx = x.to_array()
    got = res.to_array()
    # REDUCE_SUM keeps the last rep's sum (REDUCE_SUM-Seq.cpp:34-51)
    sum_ref = np.sum(hx.astype(np.longdouble))
    if abs(float(got[1] - sum_ref)) > 64 * eps * float(np.sum(np.abs(hx))):
        raise RuntimeError(f"REDUCE_SUM {got[1]!r} vs {float(sum_ref)!r}")
    out["Algorithm_REDUCE_SUM"] = f"rel err vs 80-bit sum {abs(float(got[1] - sum_ref)) / abs(float(sum_ref)):.1e}"
    # DOT accumulates every rep (m_dot += dot, DOT-Seq.cpp:45); a and b do not change
    ha, hb = a.to_array(), b.to_array()
    dot_ref = float(np.dot(ha, hb))
    dot_abs = dot_ref  # all terms are non-negative
    if abs(got[0] - reps * dot_ref) > (reps + 64) * eps * reps * dot_abs:
        raise RuntimeError(f"DOT {got[0]!r} vs {reps} x {dot_ref!r}")
    out["Stream_DOT"] = f"rel err vs host over {reps} reps {abs(got[0] - reps * dot_ref) / (reps * dot_ref):.1e}"
    del ha, hb
    # SCAN: the first 2^16 prefixes against an 80-bit prefix, and the last one
    hy0 = y.to_array(count=k)
    exact = np.concatenate([[0.0], np.cumsum(hx[:k - 1].astype(np.longdouble))])
    absx = np.concatenate([[0.0], np.cumsum(np.abs(hx[:k - 1]))])
    bound = (n / 4096 + 256) * eps
    if not np.all(np.abs(hy0 - exact).astype(np.float64) <= bound * absx + 1e-300):
        raise RuntimeError("SCAN prefixes differ at the start")
    last = y.to_array(count=1, offset_bytes=8 * (n - 1))[0]
    last_ref = sum_ref - np.longdouble(hx[-1])
    if abs(float(last - last_ref)) > bound * float(np.sum(np.abs(hx[:-1]))):
        raise RuntimeError(f"SCAN last prefix {last!r} vs {float(last_ref)!r}")
    out["Algorithm_SCAN"] = "first 2^16 prefixes and the last one within the look-back chain bound"
    del hx
    # JACOBI_2D: the same reps x 50 time steps on the two-sweep path, bit-exact
    rA, rB = abi.DeviceBuffer(8 * N * N), abi.DeviceBuffer(8 * N * N)
    L.rpx_fill_grid_f64(rA.ptr, N, SEED, 0, 0.0, 1.0, 0.0, None)
    L.rpx_fill_grid_f64(rB.ptr, N, SEED, 1, 0.0, 1.0, 0.0, None)
    L.rpx_jacobi2d_f64(rA.ptr, rB.ptr, N, reps * TSTEPS, None, None)
    if not (np.array_equal(rA.to_array(), _grid(gA_final)) and np.array_equal(rB.to_array(), gB_final.to_array())):
        raise RuntimeError("JACOBI_2D fused grids differ from the two-sweep path")
    out["Polybench_JACOBI_2D"] = f"A and B bit-exact against the two-sweep path after {reps * TSTEPS} time steps"
    return out


def _grid(ptr: int) -> np.ndarray:
    from rajaperf_amd import abi
    N = N_GRID
    h = np.empty(N * N)
    abi.lib().rpx_memcpy_d2h(h.ctypes.data, ptr, h.nbytes)
    return h


def stream_ceiling(L, dst, src, reps: int = 10) -> float:
    """Average ms of a COPY of 2^28 doubles (2 GiB read, 2 GiB written)."""
    from rajaperf_amd import abi
    L.rpx_copy_f64(dst.ptr, src.ptr, N_STREAM, None, None)
    abi.synchronize()
    e0, e1 = abi.Event(), abi.Event()
    e0.record(None)
    for _ in range(reps):
        L.rpx_copy_f64(dst.ptr, src.ptr, N_STREAM, None, None)
    e1.record(None)
    abi.synchronize()
    return e0.elapsed_ms(e1) / reps


def _timed(fn, k, warmup=1):
    from rajaperf_amd import abi
    for _ in range(warmup):
        fn()
    abi.synchronize()
    e0, e1 = abi.Event(), abi.Event()
    e0.record(None)
    for _ in range(k):
        fn()
    e1.record(None)
    abi.synchronize()
    return e0.elapsed_ms(e1) / k


def _vendor_dgemm_ms(dA, dB, dC, n, alpha, reps):
    """rocBLAS dgemm on the same row-major operands (C^T = B^T A^T in its
    column-major terms), the vendor FP64 comparator; None where rocBLAS
    cannot be loaded."""
    import ctypes as C
    try:
        rb = C.CDLL("/opt/rocm/lib/librocblas.so")
    except OSError:
        return None
    h = C.c_void_p()
    if rb.rocblas_create_handle(C.byref(h)) != 0:
        return None
    rb.rocblas_dgemm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(C.c_double), C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                 C.POINTER(C.c_double), C.c_void_p, C.c_int]
    al, be = C.c_double(alpha), C.c_double(0.0)
    none = 111  # rocblas_operation_none

    def run():
        st = rb.rocblas_dgemm(h, none, none, n, n, n, C.byref(al), dB.ptr, n, dA.ptr, n, C.byref(be), dC.ptr, n)
        if st != 0:
            raise RuntimeError(f"rocblas_dgemm status {st}")
    try:
        return _timed(run, reps, warmup=10)
    finally:
        rb.rocblas_destroy_handle(h)


def mixed_vs_fp64(reps: int = 5) -> dict:
    """The metric's second half, GEMM: 4096^3 (alpha 1.5) on int8 MFMA with the
    modular scheme (13 moduli, 44-bit operands) against this build's FP64
    MFMA kernel and the plain FP64
    comparator (the suite's Base_HIP shape).  The errors of both paths are
    taken on 8 sampled rows against 80-bit host arithmetic, relative to
    |alpha|(|A||B|)_ij.  Not part of `value`."""
    from rajaperf_amd import abi
    L = abi.lib()
    B = abi.DeviceBuffer
    n, alpha = 4096, 1.5
    dA, dB, Cf, Cm = B(8 * n * n), B(8 * n * n), B(8 * n * n), B(8 * n * n)
    L.rpx_fill_uniform_f64(dA.ptr, n * n, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(dB.ptr, n * n, SEED, 1, -1.0, 1.0, None)
    sb = L.rpx_gemm_mixed_scratch_bytes(n, n, n, 0)
    bits = L.rpx_gemm_mixed_bits(n, 0)
    scr = B(sb)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)
    # Ten untimed calls of each GEMM first: per-call times fall for the
    # first ~25 mixed and ~7 FP64 calls (1.4 -> 1.1 ms and 2.45 -> 2.25 ms
    # at 4096^3) before they settle; every comparator gets the same warm-up.
    ms_f64 = _timed(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, None, None), reps,
                    warmup=10)
    ms_mix = _timed(lambda: L.rpx_gemm_mixed_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, 0, scr.ptr, sb, None),
                    reps, warmup=10)
    ms_plain = _timed(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, naive, None), 1)
    ms_vendor = _vendor_dgemm_ms(dA, dB, Cf, n, alpha, reps)
    L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, None, None)
    rows = np.sort(np.random.default_rng(0).choice(n, 8, replace=False))
    A = dA.to_array().reshape(n, n)[rows]
    Bh = dB.to_array().reshape(n, n)
    ref = (A.astype(np.longdouble) @ Bh.astype(np.longdouble)) * np.longdouble(alpha)
    scale = alpha * (np.abs(A) @ np.abs(Bh))
    err_mix = float(np.max(np.abs(Cm.to_array().reshape(n, n)[rows] - ref) / scale))
    err_f64 = float(np.max(np.abs(Cf.to_array().reshape(n, n)[rows] - ref) / scale))
    # The textbook form alpha*A*B + beta*C0 (RPX_TUNE_GEMM_BETA), reported
    # beside the suite's form: errors from one call on C0, times over
    # repeated calls after the same warm-up.
    beta = 1.2
    C0 = B(8 * n * n)
    L.rpx_fill_uniform_f64(C0.ptr, n * n, SEED, 4, -1.0, 1.0, None)
    c0 = C0.to_array().reshape(n, n)[rows]
    ref_tb = ref + np.longdouble(beta) * c0.astype(np.longdouble)
    scale_tb = scale + abs(beta) * np.abs(c0)
    tb = abi.tuning(flags=abi.RPX_TUNE_GEMM_BETA)
    tb_err = {}
    for name, C, call in (("fp64_mfma", Cf, lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, tb,
                                                                    None)),
                          ("mixed", Cm, lambda: L.rpx_gemm_mixed_tuned_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, beta, n, n,
                                                                           n, 0, tb, scr.ptr, sb, None))):
        L.rpx_memcpy_d2d(C.ptr, C0.ptr, 8 * n * n, None)
        call()
        tb_err[name] = float(np.max(np.abs(C.to_array().reshape(n, n)[rows] - ref_tb) / scale_tb))
    ms_tb_f64 = _timed(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, tb, None), reps,
                       warmup=10)
    ms_tb_mix = _timed(lambda: L.rpx_gemm_mixed_tuned_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, 0, tb,
                                                          scr.ptr, sb, None), reps, warmup=10)
    textbook = {"form": "alpha*A*B + beta*C, beta 1.2, C0 U[-1,1) (the suite's own form ignores beta)",
                "fp64_mfma_ms": round(ms_tb_f64, 4), "mixed_ms": round(ms_tb_mix, 4),
                "speedup_vs_fp64_mfma": round(ms_tb_f64 / ms_tb_mix, 3),
                "max_err_rel_mixed": tb_err["mixed"], "max_err_rel_fp64": tb_err["fp64_mfma"],
                "error_reference": "80-bit host on 8 sampled rows, relative to |alpha|(|A||B|) + |beta C0|"}
    return {"kernel": "Polybench_GEMM 4096^3, alpha 1.5, U[-1,1)", "moduli": 13, "operand_bits": bits,
            "fp64_mfma_ms": round(ms_f64, 4),
            "plain_fp64_ms": round(ms_plain, 3), "mixed_ms": round(ms_mix, 4),
            "vendor_dgemm_ms": None if ms_vendor is None else round(ms_vendor, 4),
            "speedup_vs_fp64_mfma": round(ms_f64 / ms_mix, 3), "speedup_vs_plain_fp64": round(ms_plain / ms_mix, 2),
            "speedup_vs_vendor_dgemm": None if ms_vendor is None else round(ms_vendor / ms_mix, 3),
            "max_err_rel_abs_ab_mixed": err_mix, "max_err_rel_abs_ab_fp64": err_f64,
            "error_reference": "80-bit host sums on 8 sampled rows", "textbook_beta": textbook}


def mixed_mass_vs_fp64(reps: int = 10) -> dict:
    """MASS3DPA at NE = 262,144, (D1D, Q1D) = (4, 5), with its basis
    contractions on int8 MFMA (7 slices), against this build's FP64 kernel and
    the plain FP64 comparator (the suite's Base_HIP shape).  Error per output
    against the FP64 kernel, relative to the operator applied to |B|, |Bt|, D,
    |X|.  Not part of `value`."""
    from rajaperf_amd import abi
    L = abi.lib()
    B = abi.DeviceBuffer
    NE, D, Q = 262144, 4, 5
    Bm, Bt, Ba, Bta = B(8 * Q * D), B(8 * Q * D), B(8 * Q * D), B(8 * Q * D)
    L.rpx_fill_uniform_f64(Bm.ptr, Q * D, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(Bt.ptr, Q * D, SEED, 1, -1.0, 1.0, None)
    Ba.upload(np.abs(Bm.to_array()))
    Bta.upload(np.abs(Bt.to_array()))
    Dm = B(8 * NE * Q ** 3)
    X, Xa = B(8 * NE * D ** 3), B(8 * NE * D ** 3)
    L.rpx_fill_uniform_f64(Dm.ptr, NE * Q ** 3, SEED, 2, 0.5, 2.0, None)
    L.rpx_fill_uniform_f64(X.ptr, NE * D ** 3, SEED, 3, -1.0, 1.0, None)
    Xa.upload(np.abs(X.to_array()))
    Yf, Ym, Yz = B(8 * NE * D ** 3), B(8 * NE * D ** 3), B(8 * NE * D ** 3)
    sb = L.rpx_mass3dpa_mixed_scratch_bytes(D, Q, 7)
    scr = B(sb)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)
    ms_f64 = _timed(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, None, None), reps)
    ms_plain = _timed(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, naive, None), reps)
    ms_mix = _timed(lambda: L.rpx_mass3dpa_mixed_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, 7, scr.ptr,
                                                     sb, None), reps)
    for y in (Yf, Ym, Yz):
        L.rpx_fill_const_f64(y.ptr, NE * D ** 3, 0.0, None)
    L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yf.ptr, NE, D, Q, None, None)
    L.rpx_mass3dpa_mixed_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Ym.ptr, NE, D, Q, 7, scr.ptr, sb, None)
    L.rpx_mass3dpa_f64(Ba.ptr, Bta.ptr, Dm.ptr, Xa.ptr, Yz.ptr, NE, D, Q, None, None)
    err = float(np.max(np.abs(Ym.to_array() - Yf.to_array()) / Yz.to_array()))
    return {"kernel": "Apps_MASS3DPA NE=262144, D1D=4, Q1D=5", "slices": 7, "fp64_ms": round(ms_f64, 4),
            "plain_fp64_ms": round(ms_plain, 4), "mixed_ms": round(ms_mix, 4),
            "speedup_vs_fp64": round(ms_f64 / ms_mix, 3), "speedup_vs_plain_fp64": round(ms_plain / ms_mix, 3),
            "max_err_rel_magnitude_operator": err}


def cpu_baseline(gpu_ms_per_step=None):
    """The oracle's Base_Seq (one thread) and Base_OpenMP, timed in this run on
    this host, on a fixed fraction of the workload: TRIAD/DOT/REDUCE_SUM/SCAN
    at N = 2^27 (half of 2^28) and JACOBI_2D on the full 8192 x 8192 grid for
    5 of the 50 time steps, seeded inputs, two reps each.  Each kernel's time
    per rep is scaled to the full workload (linear in N and in time steps) to
    give an equivalent step time; `value` is Base_OpenMP's rate in the suite's
    bytes-per-rep definition."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    n, ts = 1 << 27, 5
    runs = [("Stream_TRIAD", dict(size=n), N_STREAM / n), ("Stream_DOT", dict(size=n), N_STREAM / n),
            ("Algorithm_REDUCE_SUM", dict(size=n), N_STREAM / n), ("Algorithm_SCAN", dict(size=n), N_STREAM / n),
            ("Polybench_JACOBI_2D", dict(size=(N_GRID - 2) ** 2, tsteps=ts), TSTEPS / ts)]
    legs = {}
    for variant, nt in (("Base_Seq", 1), ("Base_OpenMP", threads)):
        total_bytes, total_s, step_ms = 0, 0.0, 0.0
        for k, kw, scale in runs:
            r = orc.run(k, variant, input_mode=1, seed=SEED, checkrun=2, nthreads=nt, **kw)
            total_bytes += r["bytes_per_rep"] * r["reps"]
            total_s += r["time_s"]
            step_ms += r["time_s"] / r["reps"] * scale * 1e3
        leg = {"GB/s": round(total_bytes / total_s / 1e9, 2), "threads": nt, "timed_s": round(total_s, 2),
               "ms_per_step_full_workload": round(step_ms, 1)}
        if gpu_ms_per_step:
            leg["gpu_speedup"] = round(step_ms / gpu_ms_per_step, 1)
        legs[variant] = leg
    return {"value": legs["Base_OpenMP"]["GB/s"], "unit": "GB/s (suite bytes per rep)", "cores": threads,
            "kind": "port",
            "sample": f"oracle Base_Seq (1 thread) and Base_OpenMP ({threads} threads) in this run: 2 reps each of "
                      f"TRIAD/DOT/REDUCE_SUM/SCAN at N=2^27 and JACOBI_2D 8192x8192 x {ts} of 50 tsteps, seeded "
                      "inputs; value = Base_OpenMP",
            "Base_Seq": legs["Base_Seq"], "Base_OpenMP": legs["Base_OpenMP"]}


def host_info() -> dict:
    """nproc, CPU model, sockets and NUMA nodes of this host."""
    info = {"nproc": os.cpu_count() or 0}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        txt = open("/proc/cpuinfo").read()
        models = [l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("model name")]
        phys = {l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("physical id")}
        info["model"] = models[0] if models else "unknown"
        info["sockets"] = len(phys) or 1
    except OSError:
        pass
    try:
        info["numa_nodes"] = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")])
    except OSError:
        pass
    return info


def cpu_table() -> str:
    """The oracle's Base_Seq (one thread) and Base_OpenMP on every kernel of
    the hot path, at bounded sample sizes (about a second of CPU work each),
    seeded inputs, two reps.  Rates use the kernels' algorithmic bytes (and
    2MNK FLOPs for GEMM).  SORT has no Base_OpenMP in the suite."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    hi = host_info()
    threads = max(1, min(16, hi.get("affinity", hi["nproc"])))
    n = 1 << 24
    runs = [("Stream_COPY", dict(size=n)), ("Stream_MUL", dict(size=n)), ("Stream_ADD", dict(size=n)),
            ("Stream_TRIAD", dict(size=n)), ("Stream_DOT", dict(size=n)), ("Algorithm_REDUCE_SUM", dict(size=n)),
            ("Algorithm_SCAN", dict(size=n)), ("Algorithm_SORT", dict(size=1 << 21)),
            ("Polybench_JACOBI_2D", dict(size=2046 ** 2, tsteps=5)), ("Polybench_GEMM", dict(size=512 * 512, nk=512)),
            ("Apps_MASS3DPA", dict(size=125 * 16384, d1d=4, q1d=5)),
            ("Apps_DIFFUSION3DPA", dict(size=125 * 8192, d1d=4, q1d=5))]
    lines = [f"# host: {hi.get('model', '?')}; nproc {hi['nproc']}, affinity {hi.get('affinity', '?')}, "
             f"sockets {hi.get('sockets', '?')}, NUMA nodes {hi.get('numa_nodes', '?')}; "
             f"Base_OpenMP on {threads} threads",
             "# command: python bench.py --cpu-table   (oracle Base_Seq / Base_OpenMP, seeded inputs, 2 reps)",
             f"{'kernel':<22} {'problem size':>13} {'variant':<12} {'ms/rep':>10} {'GB/s':>9} {'GFLOP/s':>9}"]
    for k, kw in runs:
        for variant, nt in (("Base_Seq", 1), ("Base_OpenMP", threads)):
            try:
                r = orc.run(k, variant, input_mode=1, seed=SEED, checkrun=2, nthreads=nt, **kw)
            except ValueError:
                continue
            t = r["time_s"] / r["reps"]
            flops = 2 * 512 ** 3 if k == "Polybench_GEMM" else r["flops_per_rep"]
            lines.append(f"{k:<22} {r['problem_size']:>13} {variant:<12} {t * 1e3:>10.3f} "
                         f"{r['bytes_per_rep'] / t / 1e9:>9.2f} {flops / t / 1e9:>9.2f}")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-table", action="store_true",
                    help="print the oracle's Base_Seq/Base_OpenMP timings of every kernel (CPU only) and exit")
    args = ap.parse_args()
    if args.cpu_table:
        print(cpu_table())
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the kernels are single-device; multi-GPU sharding is a design note).", file=sys.stderr)
        sys.exit(2)

    g = gpu_run(args.steps, args.warmup)
    ab, cb = algorithmic_bytes(), compulsory_bytes()
    total_ms = g["total_ms"]
    ms_per_step = total_ms / args.steps
    step_s = total_ms / 1e3 / args.steps
    value = sum(v for k, v in cb.items() if not k.startswith("_")) / step_s / 1e9
    suite_gbs = sum(v for k, v in ab.items() if not k.startswith("_")) / step_s / 1e9
    copy_gbs = 2 * 8 * N_STREAM / (g["copy_ms"] / 1e3) / 1e9

    # The dominant kernel is the fused JACOBI_2D launch that stores no B (16 of
    # the 17 launches of a step): it reads A and writes the new A.
    launch_ms = g["skip_b_ms"] / (args.steps * LAUNCHES)
    launch_bytes = cb["_jacobi_launch"]
    achieved = launch_bytes / (launch_ms / 1e3) / 1e9
    flops = fused_launch_flops(N_GRID, ROWS, DEPTH)
    hbm_floor_ms = launch_bytes / (HBM_PEAK_GBS * 1e9) * 1e3
    fp64_floor_ms = flops / (FP64_PEAK_TFLOPS * 1e12) * 1e3
    traffic = read_traffic(f"jacobi2d_fused_depth{DEPTH}_rows{ROWS}_skip_b_8192")
    kernels = {}
    for k, ms in g["per_kernel_ms"].items():
        t = ms / args.steps
        gbs = cb[k] / (t / 1e3) / 1e9
        kernels[k] = {"ms_per_rep": round(t, 4), "hbm_GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                      "frac_copy_ceiling": round(gbs / copy_gbs, 4),
                      "suite_GB/s": round(ab[k] / (t / 1e3) / 1e9, 1), "check": g["checks"][k]}

    out = {
        "metric": "per-kernel GB/s or GFLOP/s vs roofline; mixed vs FP64 speedup at stated error",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 42, generated on device)",
        "config": {"workload": "HBM-bound plain FP64: TRIAD/DOT/REDUCE_SUM/SCAN at N=2^28, "
                               "JACOBI_2D 8192x8192 x 50 tsteps",
                   "n_stream": N_STREAM, "jacobi_n": N_GRID, "jacobi_tsteps": TSTEPS,
                   "variant": "MI355X_Native", "arch": g["arch"]},
        "value_definition": "compulsory HBM bytes of the step (each kernel's inputs read once, outputs written "
                            "once; JACOBI_2D per fused launch) / step time",
        "suite_algorithmic_GBps": round(suite_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"jacobi_stream_kernel<{DEPTH},false> ({DEPTH} time steps = "
                               f"{2 * DEPTH} sweeps per launch, chunks of {ROWS} rows, B not stored, 8192x8192)",
                     "compulsory_bytes_per_launch": launch_bytes,
                     "avg_launch_ms": round(launch_ms, 5),
                     "hbm_floor_ms": round(hbm_floor_ms, 5),
                     "executed_fp64_flops_per_launch": flops,
                     "fp64_floor_ms": round(fp64_floor_ms, 5),
                     "frac_copy_ceiling": round(achieved / copy_gbs, 4)},
        "stream_ceiling": {"kernel": "COPY of 2^28 doubles", "ms": round(g["copy_ms"], 4),
                           "GB/s": round(copy_gbs, 1)},
        "kernels": kernels,
        "wall_s": round(g["wall"], 3),
        "mixed": mixed_vs_fp64(),
        "mixed_mass3dpa": mixed_mass_vs_fp64(),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ms_per_step)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
