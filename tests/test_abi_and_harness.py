"""CPU tests of the boundary: the C ABI library loads and exports every symbol
the header declares (no compute calls), and the suite harness parses its
command line, selects kernels/variants and reports bad input like the suite
(src/common/RunParams.cpp, Executor.cpp:360-373)."""

import ctypes
import os
import subprocess

import pytest

from rajaperf_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "rajaperf_amd", "bin", "rajaperf-mi355x")


def test_library_exports_every_declared_symbol():
    declared = abi.declared_symbols()
    assert len(declared) >= 40
    dll = ctypes.CDLL(abi.KERNEL_LIB)
    missing = [s for s in declared if not hasattr(dll, s)]
    assert missing == []
    # and the binding knows every declared entry point
    assert sorted(set(declared) - set(abi._SIGS)) == []


def test_abi_version_and_status_strings():
    L = abi.lib()
    assert L.rpx_abi_version() == 1
    assert L.rpx_status_string(abi.RPX_ERR_SCRATCH_TOO_SMALL) == b"scratch buffer too small"
    assert L.rpx_status_string(0) == b"success"


def test_argument_errors_need_no_device():
    L = abi.lib().raw
    # invalid arguments are rejected before any device call
    assert L("rpx_triad_f64")(None, None, None, 1.0, 10, None, None) == abi.RPX_ERR_INVALID_ARGUMENT
    assert L("rpx_copy_f64")(None, None, -1, None, None) == abi.RPX_ERR_INVALID_ARGUMENT
    assert L("rpx_diffusion3dpa_f64")(8, 8, 8, 8, 8, 1, 3, 5, 1, None, None) == abi.RPX_ERR_UNSUPPORTED
    assert L("rpx_mass3dpa_f64")(8, 8, 8, 8, 8, 1, 9, 9, None, None) == abi.RPX_ERR_UNSUPPORTED
    o = abi.offsets_array([5, 2])
    assert L("rpx_sort_segmented_f64")(8, o, 1, 8, 1 << 20, None) == abi.RPX_ERR_INVALID_ARGUMENT
    assert L("rpx_jacobi2d_f64")(8, 8, 10, 1, None, None) == abi.RPX_ERR_INVALID_ARGUMENT  # A == B


def test_scratch_queries():
    L = abi.lib()
    assert L.rpx_reduce_scratch_bytes(1 << 28, None) >= 256 + 8 * 2048
    assert L.rpx_scan_scratch_bytes(1 << 28, None) >= 3 * 8 * (1 << 16)
    offs = abi.offsets_array([0, 100, 100, 5000])
    assert L.rpx_sort_segmented_scratch_bytes(offs, 3) >= 8 * 5000


def test_mixed_gemm_bit_count_follows_the_moduli_product():
    """b is the largest integer <= 46 with 2 nk 4^b < M, M the product of
    the first T of the 16 largest pairwise coprime integers <= 256; the
    default T is the smallest with b >= 42.  The moduli are recomputed here
    from their definition."""
    import math
    mods = []
    for m in range(256, 1, -1):
        if all(math.gcd(m, x) == 1 for x in mods):
            mods.append(m)
        if len(mods) == 16:
            break
    L = abi.lib()
    for nk in (1, 7, 64, 1000, 4096, 19020, 131071):
        auto = None
        for t in range(1, 17):
            M = math.prod(mods[:t])
            want = max((b for b in range(47) if 2 * nk * 4 ** b < M), default=-1)
            got = L.rpx_gemm_mixed_bits(nk, t)
            assert got == (want if want >= 1 else abi.RPX_ERR_UNSUPPORTED), (nk, t)
            if auto is None and want >= 42:
                auto = want
        assert L.rpx_gemm_mixed_bits(nk, 0) == auto
    assert L.rpx_gemm_mixed_bits(4096, 0) == 44 and L.rpx_gemm_mixed_bits(4096, 13) == 44
    assert L.rpx_gemm_mixed_bits(131072, 0) == abi.RPX_ERR_INVALID_ARGUMENT
    assert L.rpx_gemm_mixed_scratch_bytes(64, 64, 131072, 0) == 0
    assert L.rpx_gemm_mixed_scratch_bytes(64, 64, 64, 17) == 0


def run_harness(*args):
    return subprocess.run([HARNESS, *args], capture_output=True, text=True, timeout=120)


def test_harness_lists_kernels_and_variant():
    out = run_harness("--print-kernels").stdout
    for k in ["Stream_TRIAD", "Stream_DOT", "Polybench_GEMM", "Polybench_JACOBI_2D", "Apps_MASS3DPA",
              "Apps_DIFFUSION3DPA", "Algorithm_SCAN", "Algorithm_REDUCE_SUM", "Algorithm_SORT"]:
        assert k in out
    assert "MI355X_Native" in run_harness("--print-variants").stdout


def test_harness_bad_input_runs_nothing():
    r = run_harness("-k", "NotAKernel")
    assert r.returncode == 1 and "Bad input" in r.stdout and "NotAKernel" in r.stdout
    r = run_harness("-v", "Base_Seq")  # provided by the suite, not by this harness
    assert r.returncode == 1 and "not built into this harness" in r.stdout
    r = run_harness("--size", "10", "--sizefact", "2")  # Direct and Factor sizes conflict
    assert r.returncode == 1


def test_harness_combiners_and_graph_flags_parse():
    r = run_harness("--npasses-combiners", "Average", "Median")  # RunParams.cpp:1858-1878 accepts three
    assert r.returncode == 1 and "Median" in r.stdout
    r = run_harness("--dryrun", "-k", "Stream_TRIAD", "Algorithm_SORT", "--graph",
                    "--npasses-combiners", "Average", "Minimum", "Maximum")
    assert r.returncode == 0
    assert "--graph" in run_harness("--help").stdout


def test_harness_dryrun_shapes_match_reference_formulas():
    r = run_harness("--dryrun", "-k", "Polybench", "Apps", "--size", str(8190 * 8190))
    assert r.returncode == 0
    lines = {l.split()[0]: l.split() for l in r.stdout.splitlines() if l.strip().startswith(("Polybench", "Apps"))}
    assert lines["Polybench_JACOBI_2D"][2] == str(8190 * 8190)       # N = 8192
    assert lines["Polybench_GEMM"][2] == str(8190 * 8190)            # ni = nj = 8190
    r = run_harness("--dryrun", "-k", "GEMM", "--size", str(4096 * 4096), "--nk", "4096")
    gemm = [l for l in r.stdout.splitlines() if "Polybench_GEMM" in l][0].split()
    assert gemm[6] == str(8 * (2 * 4096 * 4096) + 8 * 4096 * 4096)  # bytes/rep with nk = 4096
    assert gemm[8] == str((1 + 3 * 4096) * 4096 * 4096)               # POLYBENCH_GEMM.cpp:47-48


def test_harness_gemm_alpha_beta_flags(orc):
    """--alpha / --beta set the GEMM scalars (the suite fixes them at
    POLYBENCH_GEMM.cpp:35-36); non-numbers and non-finite values are bad
    input.  In the oracle, doubling alpha doubles the checksum exactly (2x
    is exact through alpha*A, the sums and the weighted checksum), and beta
    never reaches the result."""
    assert run_harness("--dryrun", "-k", "GEMM", "--alpha", "2.5", "--beta", "-1").returncode == 0
    for bad in (["--alpha", "nan"], ["--beta", "x"], ["--alpha"]):
        r = run_harness("--dryrun", "-k", "GEMM", *bad)
        assert r.returncode != 0 and "Bad input" in r.stdout
    base = orc.run("Polybench_GEMM", size=40 * 40)
    assert orc.run("Polybench_GEMM", size=40 * 40, alpha=2 * 0.62)["checksum"] == 2 * base["checksum"]
    assert orc.run("Polybench_GEMM", size=40 * 40, beta=5.0)["checksum"] == base["checksum"]


def test_harness_without_device_fails_cleanly():
    if abi.device_count() > 0:
        pytest.skip("a device is present")
    r = run_harness("-k", "TRIAD", "--checkrun")
    assert r.returncode == 3 and "no gfx950 device" in r.stderr
