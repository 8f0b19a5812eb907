"""Per-kernel roofline report of the MI355X variant.

    python -m rajaperf_amd.roofline [--configs 2,3,4,5] [--reps 5] [--csv FILE]

Times each C-ABI entry point on device-resident seeded inputs with HIP
events (after warm-up) and reports, per kernel, the time per rep and the
throughput against its roofline: algorithmic bytes (the suite's bytes per
rep) over time against 8 TB/s HBM for the memory-bound kernels, 2*M*N*K
flops over time against the MFMA peak for GEMM.  Config 4 also reports the
mixed-precision GEMM's speedup over the FP64 paths and its error, per
element, relative to |alpha|(|A||B|)_ij on sampled rows checked in 80-bit
arithmetic on the host.

Configurations (BASELINE workloads):
  2  TRIAD/DOT/REDUCE_SUM/SCAN (+COPY/MUL/ADD) at N=2^28; JACOBI_2D 8192^2
  3  MASS3DPA and DIFFUSION3DPA, NE=262,144, D1D=4, Q1D=5
  4  GEMM 4096^3 (FP64 MFMA, the plain FP64 comparator, mixed precision);
     MASS3DPA of config 3 with its contractions on int8 MFMA
  5  SORT and SCAN over 256 segments, lengths log-uniform in [2^10, 2^20];
     MASS3DPA over 64 ragged batches of orders 2..8 from one plan (one launch
     per order, and one persistent launch for all orders);
     one pass of all kernels at random sizes, eager and as one HIP graph
"""

from __future__ import annotations

import argparse
import csv
import sys

import numpy as np

from . import abi

HBM_GBS = 8000.0
FP64_TFLOPS = 78.6
INT8_TOPS = 5000.0
SEED = 42


def _time(fn, reps, warmup=2):
    for _ in range(warmup):
        fn()
    abi.synchronize()
    e0, e1 = abi.Event(), abi.Event()
    e0.record(None)
    for _ in range(reps):
        fn()
    e1.record(None)
    abi.synchronize()
    return e0.elapsed_ms(e1) / reps


def _row(rows, kernel, config, ms, nbytes=None, flops=None, bound="hbm", peak=None, **extra):
    r = {"kernel": kernel, "config": config, "ms_per_rep": round(ms, 5)}
    if nbytes is not None:
        gbs = nbytes / (ms / 1e3) / 1e9
        r.update({"GB/s": round(gbs, 1)})
        if bound == "hbm":
            r.update({"bound": "hbm", "frac": round(gbs / HBM_GBS, 4)})
    if flops is not None:
        tf = flops / (ms / 1e3) / 1e12
        r.update({"TFLOP/s": round(tf, 2)})
        if bound != "hbm":
            r.update({"bound": bound, "frac": round(tf / peak, 4)})
    r.update(extra)
    rows.append(r)
    print("  ".join(f"{k}={v}" for k, v in r.items()), flush=True)


def config2(rows, reps):
    L = abi.lib()
    n = 1 << 28
    B = abi.DeviceBuffer
    a, b, c = B(8 * n), B(8 * n), B(8 * n)
    L.rpx_fill_uniform_f64(b.ptr, n, SEED, 1, 0.0, 1.0, None)
    L.rpx_fill_uniform_f64(c.ptr, n, SEED, 2, 0.0, 1.0, None)
    _row(rows, "Stream_COPY", 2, _time(lambda: L.rpx_copy_f64(a.ptr, b.ptr, n, None, None), reps), 16 * n)
    _row(rows, "Stream_MUL", 2, _time(lambda: L.rpx_mul_f64(a.ptr, b.ptr, 0.75, n, None, None), reps), 16 * n)
    _row(rows, "Stream_ADD", 2, _time(lambda: L.rpx_add_f64(a.ptr, b.ptr, c.ptr, n, None, None), reps), 24 * n)
    _row(rows, "Stream_TRIAD", 2, _time(lambda: L.rpx_triad_f64(a.ptr, b.ptr, c.ptr, 0.75, n, None, None), reps),
         24 * n)
    cached = abi.tuning(0, 0, 0, abi.RPX_TUNE_CACHED)
    _row(rows, "Stream_TRIAD(cached accesses)", 2,
         _time(lambda: L.rpx_triad_f64(a.ptr, b.ptr, c.ptr, 0.75, n, cached, None), reps), 24 * n)
    sb = L.rpx_reduce_scratch_bytes(n, None)
    scr, res = B(sb), B(16)
    L.rpx_reduce_scratch_init(scr.ptr, sb, None)
    _row(rows, "Stream_DOT", 2,
         _time(lambda: L.rpx_dot_f64(b.ptr, c.ptr, n, res.ptr, 0, scr.ptr, sb, None, None), reps), 16 * n + 16)
    _row(rows, "Algorithm_REDUCE_SUM", 2,
         _time(lambda: L.rpx_reduce_sum_f64(b.ptr, n, res.ptr, scr.ptr, sb, None, None), reps), 8 * n + 16)
    _row(rows, "Stream_DOT(cached accesses)", 2,
         _time(lambda: L.rpx_dot_f64(b.ptr, c.ptr, n, res.ptr, 0, scr.ptr, sb, cached, None), reps), 16 * n + 16)
    ssb = L.rpx_scan_scratch_bytes(n, None)
    sscr = B(ssb)
    L.rpx_scan_scratch_init(sscr.ptr, ssb, None)
    for ipt in (16, 32, 64):
        t = abi.tuning(0, ipt)
        _row(rows, f"Algorithm_SCAN(ipt {ipt})", 2,
             _time(lambda: L.rpx_scan_exclusive_f64(b.ptr, a.ptr, n, sscr.ptr, ssb, t, None), reps), 16 * n)
    t512 = abi.tuning(512, 64)
    _row(rows, "Algorithm_SCAN(ipt 64, 512 work-items per tile)", 2,
         _time(lambda: L.rpx_scan_exclusive_f64(b.ptr, a.ptr, n, sscr.ptr, ssb, t512, None), reps), 16 * n)
    _row(rows, "Algorithm_SCAN(ipt 64, cached accesses)", 2,
         _time(lambda: L.rpx_scan_exclusive_f64(b.ptr, a.ptr, n, sscr.ptr, ssb, cached, None), reps), 16 * n)
    _row(rows, "Algorithm_SCAN(inclusive form, ipt 64)", 2,
         _time(lambda: L.rpx_scan_inclusive_f64(b.ptr, a.ptr, n, sscr.ptr, ssb, None, None), reps), 16 * n)
    if L.rpx_scan_scratch_status(sscr.ptr) != 0:
        raise RuntimeError("scan look-back wait expired")
    del a, b, c
    N = 8192
    gA, gB = B(8 * N * N), B(8 * N * N)
    L.rpx_fill_grid_f64(gA.ptr, N, SEED, 0, 0.0, 1.0, 0.0, None)
    L.rpx_fill_grid_f64(gB.ptr, N, SEED, 1, 0.0, 1.0, 0.0, None)
    sweep = 8 * (N * N - 4) + 8 * (N - 2) * (N - 2)
    ms = _time(lambda: L.rpx_jacobi2d_f64(gA.ptr, gB.ptr, N, 50, None, None), max(1, reps // 2))
    _row(rows, "Polybench_JACOBI_2D(two-sweep, per sweep)", 2, ms / 100, sweep)
    # Temporally blocked launches: `depth` time steps (2*depth sweeps) each,
    # the new A ping-ponging between two grids.
    gQ = B(L.rpx_jacobi2d_fused_scratch_bytes(N))
    L.rpx_jacobi2d_copy_boundary_f64(gQ.ptr, gA.ptr, N, None)
    # Per launch: `depth` time steps on tiles of `rows` output rows per
    # wavefront.  "B skipped": the launches of a run other than the last
    # store no B (each run's B is overwritten before it is read).
    skip = abi.RPX_TUNE_JACOBI_SKIP_B
    # Rows of 48 and more run the streaming kernel: a wavefront walks a chunk
    # of that many rows (the first and last strips in chunks of 24).
    for depth, rows_t, flags in ((1, 8, 0), (2, 8, 0), (3, 8, 0), (2, 16, 0), (3, 16, 0), (3, 99, 0), (4, 144, 0),
                                 (1, 8, skip), (2, 4, skip), (2, 8, skip), (2, 12, skip), (2, 16, skip),
                                 (2, 99, skip), (3, 4, skip), (3, 8, skip), (3, 12, skip), (3, 16, skip),
                                 (3, 99, skip), (4, 144, skip)):
        grids = [gA.ptr, gQ.ptr]
        t = abi.tuning(0, rows_t, 0, flags)

        def launch(depth=depth, grids=grids, t=t):
            L.rpx_jacobi2d_fused_launch_f64(grids[1], gB.ptr, grids[0], N, depth, t, None)
            grids.reverse()
        ms = _time(launch, 4 * reps)
        kind = (f"rows {rows_t}" if rows_t < 48 else f"streaming chunks of {rows_t} rows") + \
            (", B skipped" if flags else "")
        _row(rows, f"Polybench_JACOBI_2D(fused depth {depth}, {kind}, per launch)", 2, ms, 2 * depth * sweep)
    sb = L.rpx_jacobi2d_fused_scratch_bytes(N)
    for depth, rows_t in ((2, 16), (3, 8), (3, 16), (3, 99), (4, 144)):
        t = abi.tuning(0, rows_t, 0, 0, depth)
        ms = _time(lambda: L.rpx_jacobi2d_fused_f64(gA.ptr, gB.ptr, N, 50, gQ.ptr, sb, t, None), max(1, reps // 2))
        _row(rows, f"Polybench_JACOBI_2D(50 time steps at depth {depth}, rows {rows_t}, per time step)", 2,
             ms / 50, 2 * sweep)


def config3(rows, reps):
    L = abi.lib()
    NE, D, Q = 262144, 4, 5
    B = abi.DeviceBuffer
    Bm, Bt = B(8 * Q * D), B(8 * Q * D)
    L.rpx_fill_uniform_f64(Bm.ptr, Q * D, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(Bt.ptr, Q * D, SEED, 1, -1.0, 1.0, None)
    Dm = B(8 * NE * Q ** 3)
    X, Y = B(8 * NE * D ** 3), B(8 * NE * D ** 3)
    L.rpx_fill_uniform_f64(Dm.ptr, NE * Q ** 3, SEED, 2, 0.5, 2.0, None)
    L.rpx_fill_uniform_f64(X.ptr, NE * D ** 3, SEED, 3, -1.0, 1.0, None)
    L.rpx_fill_const_f64(Y.ptr, NE * D ** 3, 0.0, None)
    nbytes = 2 * 8 * Q * D + 2 * 8 * D ** 3 * NE + 8 * Q ** 3 * NE + 8 * D ** 3 * NE
    flops = NE * (2 * D ** 3 * Q + 2 * D * D * Q * Q + 2 * D * Q ** 3 + Q ** 3 + 2 * Q ** 3 * D + 2 * Q * Q * D * D
                  + 2 * Q * D ** 3 + D ** 3)
    ms = _time(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Y.ptr, NE, D, Q, None, None), reps)
    _row(rows, "Apps_MASS3DPA", 3, ms, nbytes, flops)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)
    ms_naive = _time(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Y.ptr, NE, D, Q, naive, None), reps)
    _row(rows, "Apps_MASS3DPA(plain FP64, one element per Q x Q workgroup)", 3, ms_naive, nbytes, flops)
    del Dm
    Dd = B(8 * NE * 6 * Q ** 3)
    L.rpx_fill_sym_tensor_f64(Dd.ptr, NE, Q ** 3, 6, SEED, 2, None)
    nbytes = 2 * 8 * Q * D + 2 * 8 * D ** 3 * NE + 6 * 8 * Q ** 3 * NE + 8 * D ** 3 * NE
    flops = NE * (Q * D + 5 * D * D * Q * D + 7 * D * D * Q * Q + 7 * Q * D * Q * Q + 15 * Q ** 3 + Q * D
                  + 7 * Q * Q * D * Q + 7 * Q * Q * D * D + 7 * D * Q * D * D + 3 * D ** 3)
    ms = _time(lambda: L.rpx_diffusion3dpa_f64(Bm.ptr, Bt.ptr, Dd.ptr, X.ptr, Y.ptr, NE, D, Q, 1, None, None), reps)
    _row(rows, "Apps_DIFFUSION3DPA", 3, ms, nbytes, flops)


def config4(rows, reps, sample_rows=8):
    L = abi.lib()
    n = 4096
    B = abi.DeviceBuffer
    dA, dB = B(8 * n * n), B(8 * n * n)
    L.rpx_fill_uniform_f64(dA.ptr, n * n, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(dB.ptr, n * n, SEED, 1, -1.0, 1.0, None)
    Cf, Cm = B(8 * n * n), B(8 * n * n)
    alpha = 1.5
    flops = 2 * n ** 3

    def suite_rate(ms):  # secondary rate with the suite's own count (1 + 3nk) ni nj, POLYBENCH_GEMM.cpp:47-48
        return {"suite_TFLOP/s": round((1 + 3 * n) * n * n / (ms / 1e3) / 1e12, 2)}
    # Ten untimed calls before each timed GEMM: per-call times settle only
    # after several calls (clocks rising under the sustained load).
    ms_f64 = _time(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, None, None), reps,
                   warmup=10)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)
    ms_naive = _time(lambda: L.rpx_gemm_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, naive, None), 1, warmup=1)
    # error on sampled rows against 80-bit host arithmetic
    A = dA.to_array().reshape(n, n)
    Bh = dB.to_array().reshape(n, n)
    rng = np.random.default_rng(0)
    rows_idx = np.sort(rng.choice(n, sample_rows, replace=False))
    Al = A[rows_idx].astype(np.longdouble)
    ref = (Al @ Bh.astype(np.longdouble)) * np.longdouble(alpha)
    scale = alpha * (np.abs(A[rows_idx]) @ np.abs(Bh))
    C64 = Cf.to_array().reshape(n, n)[rows_idx]
    err_f64 = float(np.max(np.abs(C64 - ref) / scale))
    _row(rows, "Polybench_GEMM(fp64-mfma)", 4, ms_f64, None, flops, "mfma", FP64_TFLOPS, err_vs_exact=err_f64,
         **suite_rate(ms_f64))
    _row(rows, "Polybench_GEMM(fp64-naive)", 4, ms_naive, None, flops, "mfma", FP64_TFLOPS, **suite_rate(ms_naive))
    # T moduli: T int8 GEMMs; the error bound is 2^(2-b) relative to
    # |alpha| rowmax colmax nk, b from the library (13 moduli: b = 44,
    # 12 moduli: b = 40 at nk = 4096).
    for t_ in (0, 12):
        sb = L.rpx_gemm_mixed_scratch_bytes(n, n, n, t_)
        bits = L.rpx_gemm_mixed_bits(n, t_)
        nmod = 13 if t_ == 0 else t_
        scr = B(sb)
        ms_mix = _time(lambda: L.rpx_gemm_mixed_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, t_, scr.ptr, sb,
                                                    None), reps, warmup=10)
        Cmix = Cm.to_array().reshape(n, n)[rows_idx]
        err_mix = float(np.max(np.abs(Cmix - ref) / scale))
        _row(rows, f"Polybench_GEMM(mixed-int8-{nmod}mod)", 4, ms_mix, None, flops, "mfma", FP64_TFLOPS,
             err_vs_exact=err_mix, err_bound=2.0 ** (2 - bits), int8_gemms=nmod,
             int8_tops=round(nmod * flops / (ms_mix / 1e3) / 1e12, 1),
             speedup_vs_fp64_mfma=round(ms_f64 / ms_mix, 3), speedup_vs_naive=round(ms_naive / ms_mix, 2),
             **suite_rate(ms_mix))
        del scr
    # The textbook form C = alpha*A*B + beta*C0 (RPX_TUNE_GEMM_BETA), a
    # separately labelled output: errors from one call on C0, against
    # 80-bit host arithmetic relative to |alpha|(|A||B|) + |beta C0|; times
    # over repeated calls (C then carries the previous result).
    beta = 1.2
    C0 = B(8 * n * n)
    L.rpx_fill_uniform_f64(C0.ptr, n * n, SEED, 4, -1.0, 1.0, None)
    c0 = C0.to_array().reshape(n, n)[rows_idx]
    ref_tb = ref + np.longdouble(beta) * c0.astype(np.longdouble)
    scale_tb = scale + abs(beta) * np.abs(c0)
    tb = abi.tuning(flags=abi.RPX_TUNE_GEMM_BETA)
    L.rpx_memcpy_d2d(Cf.ptr, C0.ptr, 8 * n * n, None)
    L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, tb, None)
    err_tb = float(np.max(np.abs(Cf.to_array().reshape(n, n)[rows_idx] - ref_tb) / scale_tb))
    ms_tb = _time(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, tb, None), reps, warmup=10)
    _row(rows, "Polybench_GEMM(fp64-mfma, textbook alpha*AB+beta*C)", 4, ms_tb, None, flops, "mfma", FP64_TFLOPS,
         err_vs_exact=err_tb, **suite_rate(ms_tb))
    sb = L.rpx_gemm_mixed_scratch_bytes(n, n, n, 0)
    scr = B(sb)
    L.rpx_memcpy_d2d(Cm.ptr, C0.ptr, 8 * n * n, None)
    L.rpx_gemm_mixed_tuned_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, 0, tb, scr.ptr, sb, None)
    err_tbm = float(np.max(np.abs(Cm.to_array().reshape(n, n)[rows_idx] - ref_tb) / scale_tb))
    ms_tbm = _time(lambda: L.rpx_gemm_mixed_tuned_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, beta, n, n, n, 0, tb, scr.ptr,
                                                      sb, None), reps, warmup=10)
    _row(rows, "Polybench_GEMM(mixed-int8-13mod, textbook alpha*AB+beta*C)", 4, ms_tbm, None, flops, "mfma",
         FP64_TFLOPS, err_vs_exact=err_tbm, speedup_vs_fp64_mfma=round(ms_tb / ms_tbm, 3), **suite_rate(ms_tbm))
    del scr, C0
    del dA, dB, Cf, Cm
    mass_mixed(rows, reps)
    diffusion_mixed(rows, reps)


def diffusion_mixed(rows, reps):
    """DIFFUSION3DPA of config 3 (symmetric-tensor D) with its contractions on
    int8 MFMA, against the FP64 kernel.  Error: the largest deviation from
    the FP64 kernel's result relative to the largest |Y| (the per-output
    magnitude-bound check is in the GPU tests)."""
    L = abi.lib()
    NE, D, Q = 262144, 4, 5
    B = abi.DeviceBuffer
    Bm, G = B(8 * Q * D), B(8 * Q * D)
    L.rpx_fill_uniform_f64(Bm.ptr, Q * D, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(G.ptr, Q * D, SEED, 1, -1.0, 1.0, None)
    Dd = B(8 * NE * 6 * Q ** 3)
    L.rpx_fill_sym_tensor_f64(Dd.ptr, NE, Q ** 3, 6, SEED, 2, None)
    X = B(8 * NE * D ** 3)
    L.rpx_fill_uniform_f64(X.ptr, NE * D ** 3, SEED, 3, -1.0, 1.0, None)
    Yf, Ym = B(8 * NE * D ** 3), B(8 * NE * D ** 3)
    nbytes = 2 * 8 * Q * D + 2 * 8 * D ** 3 * NE + 6 * 8 * Q ** 3 * NE + 8 * D ** 3 * NE
    flops = NE * (Q * D + 5 * D * D * Q * D + 7 * D * D * Q * Q + 7 * Q * D * Q * Q + 15 * Q ** 3 + Q * D
                  + 7 * Q * Q * D * Q + 7 * Q * Q * D * D + 7 * D * Q * D * D + 3 * D ** 3)
    sb = L.rpx_diffusion3dpa_mixed_scratch_bytes(D, Q, 7)
    scr = B(sb)
    ms_f64 = _time(lambda: L.rpx_diffusion3dpa_f64(Bm.ptr, G.ptr, Dd.ptr, X.ptr, Yf.ptr, NE, D, Q, 1, None, None),
                   reps)
    ms_mix = _time(lambda: L.rpx_diffusion3dpa_mixed_f64(Bm.ptr, G.ptr, Dd.ptr, X.ptr, Ym.ptr, NE, D, Q, 1, 7,
                                                         scr.ptr, sb, None), reps)
    for y in (Yf, Ym):
        L.rpx_fill_const_f64(y.ptr, NE * D ** 3, 0.0, None)
    L.rpx_diffusion3dpa_f64(Bm.ptr, G.ptr, Dd.ptr, X.ptr, Yf.ptr, NE, D, Q, 1, None, None)
    L.rpx_diffusion3dpa_mixed_f64(Bm.ptr, G.ptr, Dd.ptr, X.ptr, Ym.ptr, NE, D, Q, 1, 7, scr.ptr, sb, None)
    yf, ym = Yf.to_array(), Ym.to_array()
    err = float(np.max(np.abs(ym - yf)) / np.max(np.abs(yf)))
    _row(rows, "Apps_DIFFUSION3DPA(mixed-int8x7)", 4, ms_mix, nbytes, flops, err_vs_fp64_normwise=err,
         speedup_vs_fp64=round(ms_f64 / ms_mix, 3), fp64_ms=round(ms_f64, 5))


def mass_mixed(rows, reps):
    """MASS3DPA of config 3 with the basis contractions on int8 MFMA, against
    the FP64 kernel and the plain FP64 comparator.  Error per output relative
    to the operator applied to |B|, |Bt|, D, |X| (its magnitude bound), with
    the FP64 kernel's result as reference."""
    L = abi.lib()
    NE, D, Q = 262144, 4, 5
    B = abi.DeviceBuffer
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
    # The suite's bytes and FLOPs per rep (MASS3DPA.cpp:38-49), as in config 3.
    nbytes = 2 * 8 * Q * D + 2 * 8 * D ** 3 * NE + 8 * Q ** 3 * NE + 8 * D ** 3 * NE
    flops = NE * (2 * D ** 3 * Q + 2 * D * D * Q * Q + 2 * D * Q ** 3 + Q ** 3 + 2 * Q ** 3 * D + 2 * Q * Q * D * D
                  + 2 * Q * D ** 3 + D ** 3)
    sb = L.rpx_mass3dpa_mixed_scratch_bytes(D, Q, 7)
    scr = B(sb)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)
    ms_f64 = _time(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, None, None), reps)
    ms_naive = _time(lambda: L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, naive, None), reps)
    ms_mix = _time(lambda: L.rpx_mass3dpa_mixed_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yz.ptr, NE, D, Q, 7, scr.ptr, sb,
                                                    None), reps)
    # One application each from Y = 0: FP64 reference, mixed, and the bound.
    for y in (Yf, Ym, Yz):
        L.rpx_fill_const_f64(y.ptr, NE * D ** 3, 0.0, None)
    L.rpx_mass3dpa_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Yf.ptr, NE, D, Q, None, None)
    L.rpx_mass3dpa_mixed_f64(Bm.ptr, Bt.ptr, Dm.ptr, X.ptr, Ym.ptr, NE, D, Q, 7, scr.ptr, sb, None)
    L.rpx_mass3dpa_f64(Ba.ptr, Bta.ptr, Dm.ptr, Xa.ptr, Yz.ptr, NE, D, Q, None, None)
    yf, ym, mag = Yf.to_array(), Ym.to_array(), Yz.to_array()
    err = float(np.max(np.abs(ym - yf) / mag))
    _row(rows, "Apps_MASS3DPA(mixed-int8x7)", 4, ms_mix, nbytes, flops, err_vs_fp64=err,
         speedup_vs_fp64=round(ms_f64 / ms_mix, 3), speedup_vs_plain_fp64=round(ms_naive / ms_mix, 3),
         fp64_ms=round(ms_f64, 5), plain_fp64_ms=round(ms_naive, 5))


def config5(rows, reps):
    L = abi.lib()
    rng = np.random.default_rng(SEED)
    lens = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), 256)).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)])
    n = int(offs[-1])
    B = abi.DeviceBuffer
    src, x, y = B(8 * n), B(8 * n), B(8 * n)
    L.rpx_fill_uniform_f64(src.ptr, n, SEED, 0, -1.0, 1.0, None)
    o = abi.offsets_array(offs.tolist())
    sb = L.rpx_sort_segmented_scratch_bytes(o, 256)
    scr = B(sb)

    def sort_once():
        L.rpx_memcpy_d2d(x.ptr, src.ptr, 8 * n, None)
        L.rpx_sort_segmented_f64(x.ptr, o, 256, scr.ptr, sb, None)
    ms_copy = _time(lambda: L.rpx_memcpy_d2d(x.ptr, src.ptr, 8 * n, None), reps)
    ms = _time(sort_once, reps) - ms_copy
    _row(rows, "Algorithm_SORT(256 ragged)", 5, ms, 8 * 16 * n, None, keys=n,
         Gkeys_per_s=round(n / (ms / 1e3) / 1e9, 3))
    ssb = L.rpx_scan_segmented_scratch_bytes(o, 256)
    sscr = B(ssb)
    L.rpx_scan_scratch_init(sscr.ptr, ssb, None)
    ms = _time(lambda: L.rpx_scan_exclusive_segmented_f64(src.ptr, y.ptr, o, 256, sscr.ptr, ssb, None), reps)
    _row(rows, "Algorithm_SCAN(256 ragged)", 5, ms, 16 * n)
    del src, x, y, scr, sscr
    mass_batches(rows, reps)
    graph_pass(rows, reps)


def graph_pass(rows, reps):
    """One pass of every kernel run eagerly and as one captured HIP graph:
    at randomly drawn sizes, and at the smallest sizes, where the pass is
    bound by launches rather than by the kernels."""
    from .suite_pass import draw_sizes
    _graph_pass_at(rows, reps, draw_sizes(np.random.default_rng(SEED)), "random sizes")
    small = {"n_stream": 4096, "n_scan": 4096, "n_sort": 1024, "gemm": (64, 64, 64), "jacobi": (130, 2),
             "mass": (64, 3), "diffusion": (64, 3)}
    _graph_pass_at(rows, reps, small, "smallest sizes, launch-bound")


def _graph_pass_at(rows, reps, sizes, label):
    from .suite_pass import SuitePass
    p = SuitePass(sizes)
    st = abi.Stream()
    with abi.Graph(st) as graph:
        p.run(st)
    st.synchronize()

    def timed(fn):
        for _ in range(2):
            fn()
        st.synchronize()
        e0, e1 = abi.Event(), abi.Event()
        e0.record(st.ptr)
        for _ in range(reps):
            fn()
        e1.record(st.ptr)
        st.synchronize()
        return e0.elapsed_ms(e1) / reps
    ms_eager = timed(lambda: p.run(st))
    ms_graph = timed(graph.launch)
    _row(rows, f"Suite pass of all kernels ({label}), one HIP graph", 5, ms_graph, p.bytes_per_pass(),
         eager_ms=round(ms_eager, 5), graph_speedup=round(ms_eager / ms_graph, 2), sizes=str(sizes).replace(" ", ""))


def mass_batches(rows, reps, nbatch=64):
    """64 MASS3DPA batches, NE log-uniform in [1024, 65536], D1D uniform in
    {2..8}, Q1D = D1D + 1, data as in config 3: one persistent launch against
    one launch per batch."""
    L = abi.lib()
    rng = np.random.default_rng(SEED)
    nes = np.exp(rng.uniform(np.log(1024), np.log(65536), nbatch)).astype(np.int64)
    ds = rng.integers(2, 9, nbatch)
    B = abi.DeviceBuffer
    keep, arr = [], (abi.MassBatch * nbatch)()
    nbytes = 0
    for i, (ne, d) in enumerate(zip(nes.tolist(), ds.tolist())):
        q = d + 1
        bufs = [B(8 * q * d), B(8 * q * d), B(8 * ne * q ** 3), B(8 * ne * d ** 3), B(8 * ne * d ** 3)]
        for k, (buf, n, lo, hi) in enumerate(zip(bufs, [q * d, q * d, ne * q ** 3, ne * d ** 3],
                                                 [-1.0, -1.0, 0.5, -1.0], [1.0, 1.0, 2.0, 1.0])):
            L.rpx_fill_uniform_f64(buf.ptr, n, SEED + i, k, lo, hi, None)
        L.rpx_fill_const_f64(bufs[4].ptr, ne * d ** 3, 0.0, None)
        keep.append(bufs)
        arr[i] = abi.MassBatch(*[b.ptr for b in bufs], ne, d, q)
        nbytes += 2 * 8 * q * d + 8 * (3 * d ** 3 + q ** 3) * ne  # MASS3DPA.cpp:38-42 per batch
    sb = L.rpx_mass3dpa_batched_scratch_bytes(arr, nbatch)
    scr = B(sb)
    L.rpx_mass3dpa_batched_plan(arr, nbatch, scr.ptr, sb, None)
    ms = _time(lambda: L.rpx_mass3dpa_batched_run(scr.ptr, None, None), reps)

    def per_batch():
        for i in range(nbatch):
            b = arr[i]
            L.rpx_mass3dpa_f64(b.B, b.Bt, b.D, b.X, b.Y, b.ne, b.d1d, b.q1d, None, None)
    ms_each = _time(per_batch, reps)
    one = abi.tuning(flags=abi.RPX_TUNE_MASS_SINGLE_LAUNCH)
    ms_one = _time(lambda: L.rpx_mass3dpa_batched_run(scr.ptr, one, None), reps)
    _row(rows, "Apps_MASS3DPA(64 ragged batches, D1D 2..8, one launch per order, one workgroup per chunk)", 5, ms, nbytes,
         elements=int(nes.sum()), per_batch_launches_ms=round(ms_each, 5), speedup=round(ms_each / ms, 3))
    _row(rows, "Apps_MASS3DPA(64 ragged batches, D1D 2..8, one persistent launch for all orders)", 5, ms_one,
         nbytes, elements=int(nes.sum()), per_batch_launches_ms=round(ms_each, 5), speedup=round(ms_each / ms_one, 3))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="2,3,4,5")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--csv", default=None)
    args = ap.parse_args(argv)
    arch = abi.init(0)
    print(f"device {arch}", flush=True)
    rows = []
    for c in args.configs.split(","):
        {"2": config2, "3": config3, "4": config4, "5": config5}[c.strip()](rows, args.reps)
    if args.csv:
        keys = sorted({k for r in rows for k in r})
        with open(args.csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=keys)
            w.writeheader()
            w.writerows(rows)
    return rows


if __name__ == "__main__":
    main(sys.argv[1:])
