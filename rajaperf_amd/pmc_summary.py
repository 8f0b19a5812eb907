"""HBM traffic per dispatch from two rocprofv3 counter runs, and a
kernel-time table from a rocprofv3 --stats run.

    python -m rajaperf_amd.pmc_summary FETCH.csv WRITE.csv > profiles/<name>.txt
    python -m rajaperf_amd.pmc_summary --stats KERNEL_STATS.csv > profiles/<name>.txt

FETCH.csv and WRITE.csv are the `*_counter_collection.csv` files of

    rocprofv3 --pmc FETCH_SIZE --output-format csv -- python3 -m rajaperf_amd.roofline --configs 2,3 --reps 1
    rocprofv3 --pmc WRITE_SIZE --output-format csv -- python3 -m rajaperf_amd.roofline --configs 2,3 --reps 1

Both counters are in KiB per dispatch.  Each is scaled by the factor that
makes the stream COPY kernel (2 GiB read, 2 GiB written at N = 2^28) read
exactly its bytes.  Every line gives the measured bytes per dispatch (the
mean over a kernel's dispatches), the least bytes the kernel's design must
move (`ideal_bytes`: each array read once and written once) and their
ratio; a ratio well above 1 means re-reads that reach HBM.
"""

from __future__ import annotations

import collections
import csv
import sys

N_STREAM = 1 << 28
N_GRID = 8192
NE, D, Q = 262144, 4, 5


def _mean_by_kernel(path):
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def _kernels():
    n, N = N_STREAM, N_GRID
    a_read = 8 * (N * N - 4)    # one read of A (POLYBENCH_JACOBI_2D.cpp:37-42)
    interior = 8 * (N - 2) ** 2  # one write of an interior
    return [
        # (substring of the demangled name, key, ideal bytes)
        ("stream_vec_kernel<0, 1, 256, true>", "stream_copy_2^28", 16 * n),
        ("stream_vec_kernel<1, 1, 256, true>", "stream_mul_2^28", 16 * n),
        ("stream_vec_kernel<2, 1, 256, true>", "stream_add_2^28", 24 * n),
        ("stream_vec_kernel<3, 1, 256, true>", "stream_triad_2^28", 24 * n),
        ("reduce_kernel<true, 256, 4, true>", "dot_2^28", 16 * n),
        ("reduce_kernel<false, 256, 4, true>", "reduce_sum_2^28", 8 * n),
        ("scan_kernel<32, true>", "scan_2^28", 16 * n),
        ("jacobi_vec_kernel<16>", "jacobi2d_sweep_8192", a_read + interior),
        ("jacobi_fused_kernel<8, 1, true>", "jacobi2d_fused_depth1_rows8_8192", a_read + 2 * interior),
        ("jacobi_fused_kernel<8, 2, true>", "jacobi2d_fused_depth2_rows8_8192", a_read + 2 * interior),
        ("jacobi_fused_kernel<8, 3, true>", "jacobi2d_fused_depth3_rows8_8192", a_read + 2 * interior),
        ("jacobi_fused_kernel<8, 1, false>", "jacobi2d_fused_depth1_rows8_skip_b_8192", a_read + interior),
        ("jacobi_fused_kernel<8, 2, false>", "jacobi2d_fused_depth2_rows8_skip_b_8192", a_read + interior),
        ("jacobi_fused_kernel<8, 3, false>", "jacobi2d_fused_depth3_rows8_skip_b_8192", a_read + interior),
        ("jacobi_fused_kernel<16, 2, false>", "jacobi2d_fused_depth2_rows16_skip_b_8192", a_read + interior),
        ("jacobi_fused_kernel<16, 3, false>", "jacobi2d_fused_depth3_rows16_skip_b_8192", a_read + interior),
        ("jacobi_fused_kernel<16, 2, true>", "jacobi2d_fused_depth2_rows16_8192", a_read + 2 * interior),
        # The streaming kernel, timed at chunks of 99 rows (depths 2, 3) and 144 (depth 4).
        ("jacobi_stream_kernel<4, false>", "jacobi2d_fused_depth4_rows144_skip_b_8192", a_read + interior),
        ("jacobi_stream_kernel<4, true>", "jacobi2d_fused_depth4_rows144_8192", a_read + 2 * interior),
        ("jacobi_stream_kernel<3, false>", "jacobi2d_fused_depth3_rows99_skip_b_8192", a_read + interior),
        ("jacobi_stream_kernel<2, false>", "jacobi2d_fused_depth2_rows99_skip_b_8192", a_read + interior),
        ("jacobi_stream_kernel<3, true>", "jacobi2d_fused_depth3_rows99_8192", a_read + 2 * interior),
        ("mass3dpa_kernel<4, 5>", "mass3dpa_262144_d4_q5", 8 * (3 * D ** 3 + Q ** 3) * NE),
        ("mass3dpa_naive_kernel<4, 5>", "mass3dpa_plain_262144_d4_q5", 8 * (3 * D ** 3 + Q ** 3) * NE),
        ("diffusion3dpa_kernel<4, 5, true>", "diffusion3dpa_262144_d4_q5", 8 * (3 * D ** 3 + 6 * Q ** 3) * NE),
    ]


def summarize(fetch_csv: str, write_csv: str) -> str:
    f, w = _mean_by_kernel(fetch_csv), _mean_by_kernel(write_csv)
    copy = [k for k in f if "stream_vec_kernel<0, 1, 256, true>" in k]
    if not copy:
        raise SystemExit("the COPY kernel is missing from the counter run; it calibrates the counters")
    ff = 8 * N_STREAM / (f[copy[0]] * 1024)
    wf = 8 * N_STREAM / (w[copy[0]] * 1024)
    out = [
        "# HBM traffic per dispatch from rocprofv3 counters, MI355X (gfx950), config-2 and config-3 sizes",
        "# commands (two runs, one counter each; repository root on PYTHONPATH):",
        "#   cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch> -o run -- "
        "python3 -m rajaperf_amd.roofline --configs 2,3 --reps 1",
        "#   cd /tmp && rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write> -o run -- "
        "python3 -m rajaperf_amd.roofline --configs 2,3 --reps 1",
        "#   python3 -m rajaperf_amd.pmc_summary <fetch>/run_counter_collection.csv <write>/run_counter_collection.csv",
        f"# calibration on the COPY kernel (reads 2 GiB, writes 2 GiB): FETCH_SIZE x {ff:.4f}, "
        f"WRITE_SIZE x {wf:.4f}",
        "# hbm_bytes_per_dispatch = FETCH_SIZE[KiB]*1024*fetch_factor + WRITE_SIZE[KiB]*1024*write_factor, "
        "averaged over dispatches",
        "# ideal_bytes = the least a kernel of this design must move: each array read once and written once",
        "#   (fused JACOBI_2D: read A once, write the new A, and the final B unless skip_b, whatever the depth)",
        "",
    ]
    for sub, key, ideal in _kernels():
        ks = [k for k in f if sub in k]
        if not ks:
            continue
        fb = f[ks[0]] * 1024 * ff
        wb = w.get(ks[0], 0.0) * 1024 * wf
        out.append(f"kernel={key} hbm_bytes_per_dispatch={fb + wb:.6e} read_bytes={fb:.6e} write_bytes={wb:.6e} "
                   f"ideal_bytes={ideal:.6e} ratio={(fb + wb) / ideal:.4f}")
    return "\n".join(out)


def kernel_stats(stats_csv: str) -> str:
    """Per-kernel calls and durations (microseconds) of a `--stats` run."""
    rows = list(csv.DictReader(open(stats_csv)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"{'kernel':<96} {'calls':>6} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'pct':>6}"]
    for r in rows:
        name = r["Name"][:96]
        out.append(f"{name:<96} {int(r['Calls']):>6} {float(r['AverageNs']) / 1e3:>10.2f} "
                   f"{float(r['MinNs']) / 1e3:>10.2f} {float(r['MaxNs']) / 1e3:>10.2f} "
                   f"{float(r['Percentage']):>6.2f}")
    return "\n".join(out)


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--stats":
        print(kernel_stats(sys.argv[2]))
    elif len(sys.argv) == 3:
        print(summarize(sys.argv[1], sys.argv[2]))
    else:
        raise SystemExit(__doc__)
