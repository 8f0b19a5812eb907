"""One pass of every kernel of the MI355X variant, for graph capture.

A `SuitePass` owns device-resident seeded inputs for TRIAD, DOT,
REDUCE_SUM, SCAN, SORT, GEMM, JACOBI_2D, MASS3DPA and DIFFUSION3DPA at the
sizes it is given, and `run(stream)` issues one rep of each on a stream --
the work of one pass of the suite's rep loops (Executor.cpp:645-670), with
each kernel's state carried from pass to pass as the suite carries it
(DOT accumulates, MASS3DPA/DIFFUSION3DPA add into Y, JACOBI_2D advances A
and B).  Every call in `run` is capturable: no host synchronisation, no
allocation.
"""

from __future__ import annotations

import numpy as np

from . import abi

KERNELS = ("Stream_TRIAD", "Stream_DOT", "Algorithm_REDUCE_SUM", "Algorithm_SCAN", "Algorithm_SORT",
           "Polybench_GEMM", "Polybench_JACOBI_2D", "Apps_MASS3DPA", "Apps_DIFFUSION3DPA")


def draw_sizes(rng: np.random.Generator) -> dict:
    """Random sizes for one pass (log-uniform within each kernel's range)."""
    lu = lambda lo, hi: int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    d_mass, d_diff = int(rng.integers(2, 9)), int(rng.integers(2, 9))
    return {"n_stream": lu(1 << 16, 1 << 26), "n_scan": lu(1 << 16, 1 << 26), "n_sort": lu(1 << 12, 1 << 22),
            "gemm": (lu(64, 2048), lu(64, 2048), lu(64, 2048)), "jacobi": (2 * lu(64, 2048), int(rng.integers(1, 8))),
            "mass": (lu(1024, 65536), d_mass), "diffusion": (lu(1024, 32768), d_diff)}


class SuitePass:
    def __init__(self, sizes: dict, seed: int = 42):
        L = abi.lib()
        B = abi.DeviceBuffer
        self.sizes = sizes
        self.keep = []

        def buf(n, sid=None, lo=-1.0, hi=1.0):
            b = B(8 * max(int(n), 1))
            if sid is None:
                L.rpx_fill_const_f64(b.ptr, int(n), 0.0, None)
            else:
                L.rpx_fill_uniform_f64(b.ptr, int(n), seed, sid, lo, hi, None)
            self.keep.append(b)
            return b

        n = sizes["n_stream"]
        self.a, self.b, self.c = buf(n), buf(n, 1, 0.0, 1.0), buf(n, 2, 0.0, 1.0)
        self.res = buf(2)
        self.rsb = L.rpx_reduce_scratch_bytes(n, None)
        self.rscr = B(self.rsb)
        L.rpx_reduce_scratch_init(self.rscr.ptr, self.rsb, None)
        ns = sizes["n_scan"]
        self.sx, self.sy = buf(ns, 3), buf(ns)
        self.ssb = L.rpx_scan_scratch_bytes(ns, None)
        self.sscr = B(self.ssb)
        L.rpx_scan_scratch_init(self.sscr.ptr, self.ssb, None)
        nt = sizes["n_sort"]
        self.tsrc, self.tx = buf(nt, 4), buf(nt)
        self.toffs = abi.offsets_array([0, nt])
        self.tsb = L.rpx_sort_segmented_scratch_bytes(self.toffs, 1)
        self.tscr = B(self.tsb)
        ni, nj, nk = sizes["gemm"]
        self.gA, self.gB, self.gC = buf(ni * nk, 5), buf(nk * nj, 6), buf(ni * nj)
        N, _ = sizes["jacobi"]
        self.jA, self.jB = B(8 * N * N), B(8 * N * N)
        self.keep += [self.jA, self.jB]
        L.rpx_fill_grid_f64(self.jA.ptr, N, seed, 7, 0.0, 1.0, 0.0, None)
        L.rpx_fill_grid_f64(self.jB.ptr, N, seed, 8, 0.0, 1.0, 0.0, None)
        self.jsb = L.rpx_jacobi2d_fused_scratch_bytes(N)
        self.jscr = B(self.jsb)
        ne, d = sizes["mass"]
        q = d + 1
        self.mass = [buf(q * d, 9), buf(q * d, 10), buf(ne * q ** 3, 11, 0.5, 2.0), buf(ne * d ** 3, 12),
                     buf(ne * d ** 3)]
        ne, d = sizes["diffusion"]
        q = d + 1
        self.diff = [buf(q * d, 13), buf(q * d, 14), buf(ne * 6 * q ** 3, 15, 0.5, 2.0), buf(ne * d ** 3, 16),
                     buf(ne * d ** 3)]
        abi.synchronize()

    def bytes_per_pass(self) -> int:
        """The suite's algorithmic bytes of one pass (each kernel's .cpp)."""
        s = self.sizes
        n, ns, nt = s["n_stream"], s["n_scan"], s["n_sort"]
        ni, nj, nk = s["gemm"]
        N, ts = s["jacobi"]
        ne_m, dm = s["mass"]
        ne_d, dd = s["diffusion"]
        qm, qd = dm + 1, dd + 1
        return (24 * n + 16 * n + 16 + 8 * n + 16 + 16 * ns + 16 * nt
                + 8 * (ni * nk + nk * nj + ni * nj)
                + 2 * ts * (8 * (N * N - 4) + 8 * (N - 2) ** 2)
                + 8 * (3 * dm ** 3 + qm ** 3) * ne_m + 8 * (3 * dd ** 3 + 6 * qd ** 3) * ne_d)

    def run(self, stream) -> None:
        L = abi.lib()
        s = self.sizes
        st = stream.ptr if stream is not None else None
        n, ns, nt = s["n_stream"], s["n_scan"], s["n_sort"]
        L.rpx_triad_f64(self.a.ptr, self.b.ptr, self.c.ptr, 0.75, n, None, st)
        L.rpx_dot_f64(self.a.ptr, self.b.ptr, n, self.res.ptr, 1, self.rscr.ptr, self.rsb, None, st)
        L.rpx_reduce_sum_f64(self.c.ptr, n, self.res.ptr + 8, self.rscr.ptr, self.rsb, None, st)
        L.rpx_scan_exclusive_f64(self.sx.ptr, self.sy.ptr, ns, self.sscr.ptr, self.ssb, None, st)
        L.rpx_memcpy_d2d(self.tx.ptr, self.tsrc.ptr, 8 * nt, st)
        L.rpx_sort_segmented_f64(self.tx.ptr, self.toffs, 1, self.tscr.ptr, self.tsb, st)
        ni, nj, nk = s["gemm"]
        L.rpx_gemm_f64(self.gC.ptr, self.gA.ptr, self.gB.ptr, 1.5, 1.2, ni, nj, nk, None, st)
        N, ts = s["jacobi"]
        L.rpx_jacobi2d_fused_f64(self.jA.ptr, self.jB.ptr, N, ts, self.jscr.ptr, self.jsb, None, st)
        ne, d = s["mass"]
        L.rpx_mass3dpa_f64(*[b.ptr for b in self.mass], ne, d, d + 1, None, st)
        ne, d = s["diffusion"]
        L.rpx_diffusion3dpa_f64(*[b.ptr for b in self.diff], ne, d, d + 1, 1, None, st)

    def outputs(self) -> dict:
        """Host copies of every output and carried state."""
        return {"triad": self.a.to_array(), "dot_reduce": self.res.to_array(), "scan": self.sy.to_array(),
                "sort": self.tx.to_array(), "gemm": self.gC.to_array(), "jacobi_A": self.jA.to_array(),
                "jacobi_B": self.jB.to_array(), "mass": self.mass[4].to_array(), "diffusion": self.diff[4].to_array()}
