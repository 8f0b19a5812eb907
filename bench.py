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

`value` is the aggregate algorithmic throughput of a step (the suite's
bytes-per-rep of every kernel summed, divided by the step time).  The
roofline object describes the dominant kernel, the temporally blocked
JACOBI_2D launch (three time steps, six sweeps, B not stored -- only the
last launch of the 50 steps stores B): the suite's algorithmic bytes of its
six sweeps over its average launch time measured with HIP events on the
launch stream; `traffic` is its HBM bytes per launch from the committed
rocprofv3 counter summary (temporal blocking keeps it far below the
algorithmic bytes, so `frac` can exceed 1).  The CPU baseline times the oracle's
Base_OpenMP restatement of the same kernels on a bounded sample.

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
SEED = 42
N_STREAM = 1 << 28
N_GRID = 8192
TSTEPS = 50
DEPTH = 3                          # JACOBI_2D time steps per fused launch
ROWS = 16                          # output rows per wavefront tile
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
    last = abi.tuning(0, ROWS)

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
    return arch, total_ms, per_kernel_ms, skip_b_ms, wall


def mixed_vs_fp64(reps: int = 5) -> dict:
    """The metric's second half: GEMM 4096^3 (alpha 1.5) on int8 MFMA with 7
    Ozaki slices against this build's FP64 MFMA kernel and the plain FP64
    comparator (the suite's Base_HIP shape), with the mixed result's largest
    deviation from the FP64 MFMA result relative to |alpha|(|A||B|)_ij on
    8 sampled rows.  Not part of `value`."""
    from rajaperf_amd import abi
    L = abi.lib()
    B = abi.DeviceBuffer
    n, alpha = 4096, 1.5
    dA, dB, Cf, Cm = B(8 * n * n), B(8 * n * n), B(8 * n * n), B(8 * n * n)
    L.rpx_fill_uniform_f64(dA.ptr, n * n, SEED, 0, -1.0, 1.0, None)
    L.rpx_fill_uniform_f64(dB.ptr, n * n, SEED, 1, -1.0, 1.0, None)
    sb = L.rpx_gemm_mixed_scratch_bytes(n, n, n, 7)
    scr = B(sb)
    naive = abi.tuning(flags=abi.RPX_TUNE_NAIVE)

    def timed(fn, k):
        fn()
        abi.synchronize()
        e0, e1 = abi.Event(), abi.Event()
        e0.record(None)
        for _ in range(k):
            fn()
        e1.record(None)
        abi.synchronize()
        return e0.elapsed_ms(e1) / k
    ms_f64 = timed(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, None, None), reps)
    ms_mix = timed(lambda: L.rpx_gemm_mixed_f64(Cm.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, 7, scr.ptr, sb, None),
                   reps)
    ms_plain = timed(lambda: L.rpx_gemm_f64(Cf.ptr, dA.ptr, dB.ptr, alpha, 1.2, n, n, n, naive, None), 1)
    rows = np.sort(np.random.default_rng(0).choice(n, 8, replace=False))
    A = dA.to_array().reshape(n, n)[rows]
    Bh = dB.to_array().reshape(n, n)
    scale = alpha * (np.abs(A) @ np.abs(Bh))
    err = float(np.max(np.abs(Cm.to_array().reshape(n, n)[rows] - Cf.to_array().reshape(n, n)[rows]) / scale))
    return {"kernel": "Polybench_GEMM 4096^3, alpha 1.5, U[-1,1)", "slices": 7, "fp64_mfma_ms": round(ms_f64, 4),
            "plain_fp64_ms": round(ms_plain, 3), "mixed_ms": round(ms_mix, 4),
            "speedup_vs_fp64_mfma": round(ms_f64 / ms_mix, 3), "speedup_vs_plain_fp64": round(ms_plain / ms_mix, 2),
            "max_err_vs_fp64_rel_abs_ab": err, "stated_error_bound": (7 + 3) * 2.0 ** -49}


def cpu_baseline():
    """Oracle Base_OpenMP on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    n = 1 << 25
    Ng = 4096
    ts = 5
    runs = [("Stream_TRIAD", dict(size=n)), ("Stream_DOT", dict(size=n)),
            ("Algorithm_REDUCE_SUM", dict(size=n)), ("Algorithm_SCAN", dict(size=n)),
            ("Polybench_JACOBI_2D", dict(size=(Ng - 2) ** 2, tsteps=ts))]
    total_bytes, total_s = 0, 0.0
    for k, kw in runs:
        r = orc.run(k, "Base_OpenMP", input_mode=1, seed=SEED, checkrun=2, nthreads=threads, **kw)
        total_bytes += r["bytes_per_rep"] * r["reps"]
        total_s += r["time_s"]
    return {"value": total_bytes / total_s / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"oracle Base_OpenMP ({threads} threads), 2 reps each of TRIAD/DOT/REDUCE_SUM/SCAN "
                      f"at N=2^25 and JACOBI_2D {Ng}x{Ng} x {ts} tsteps, seeded inputs"}


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

    arch, total_ms, per_ms, skip_b_ms, wall = gpu_run(args.steps, args.warmup)
    ab = algorithmic_bytes()
    step_bytes = sum(v for k, v in ab.items() if not k.startswith("_"))
    ms_per_step = total_ms / args.steps
    value = step_bytes * args.steps / (total_ms / 1e3) / 1e9

    # The dominant kernel is the fused JACOBI_2D launch that stores no B (16 of
    # the 17 launches of a step).  One launch = DEPTH time steps = 2*DEPTH
    # sweeps of the suite's algorithmic bytes.
    launch_ms = skip_b_ms / (args.steps * LAUNCHES)
    launch_bytes = 2 * DEPTH * ab["_jacobi_sweep"]
    achieved = launch_bytes / (launch_ms / 1e3) / 1e9
    traffic = read_traffic(f"jacobi2d_fused_depth{DEPTH}_rows{ROWS}_skip_b_8192")
    kernels = {}
    for k, ms in per_ms.items():
        t = ms / args.steps
        gbs = ab[k] / (t / 1e3) / 1e9
        kernels[k] = {"ms_per_rep": round(t, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}

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
                   "variant": "MI355X_Native", "arch": arch},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"jacobi_fused_kernel<{ROWS},{DEPTH},false> ({DEPTH} time steps = "
                               f"{2 * DEPTH} sweeps per launch, B not stored, 8192x8192)",
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "avg_launch_ms": round(launch_ms, 5)},
        "kernels": kernels,
        "wall_s": round(wall, 3),
        "mixed": mixed_vs_fp64(),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
