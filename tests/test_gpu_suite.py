"""GPU: the suite harness end to end, in the pattern of the suite's own test
(test/test-raja-perf-suite.cpp:50-178): run every kernel through
KernelBase::execute for the MI355X_Native variant with --checkrun and
several passes, then compare each checksum with the oracle's Base_Seq
checksum for the same inputs.  The suite's tolerance is 1e-7 absolute; the
element-wise, stencil, FEM and sort kernels are held to equality of the
long-double checksum."""

import csv
import os

import pytest

from rajaperf_amd.suite import run_suite

pytestmark = pytest.mark.gpu

BIT_EXACT = {"Stream_COPY", "Stream_MUL", "Stream_ADD", "Stream_TRIAD", "Polybench_JACOBI_2D", "Apps_MASS3DPA",
             "Apps_DIFFUSION3DPA", "Algorithm_SORT"}

CASES = [
    # (kernel, size, extra args)
    ("Stream_COPY", 100_003, []), ("Stream_MUL", 100_003, []), ("Stream_ADD", 100_003, []),
    ("Stream_TRIAD", 100_003, []), ("Stream_DOT", 100_003, []),
    ("Algorithm_REDUCE_SUM", 100_003, []), ("Algorithm_SCAN", 100_003, []), ("Algorithm_SORT", 20_011, []),
    ("Polybench_JACOBI_2D", 130 * 130, ["--tsteps", "5"]), ("Polybench_GEMM", 100 * 100, []),
    ("Polybench_GEMM", 100 * 100, ["--alpha", "-0.5", "--beta", "3"]),
    ("Apps_MASS3DPA", 125 * 300, []), ("Apps_DIFFUSION3DPA", 64 * 300, []),
    ("Apps_MASS3DPA", 216 * 50, ["--d1d", "5", "--q1d", "6"]),
    ("Apps_DIFFUSION3DPA", 125 * 50, ["--d1d", "4"]),
]


@pytest.mark.parametrize("mode", ["reference", "seeded"])
@pytest.mark.parametrize("kernel,size,extra", CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_harness_checksum_matches_base_seq(rpx, orc, kernel, size, extra, mode):
    args = ["-k", kernel, "--size", str(size), "--checkrun", "3", "--npasses", "2", "--input-mode", mode,
            "-t", "default", *extra]
    res = run_suite(args)
    assert res["rc"] == 0, res.get("failures")
    (run,) = [r for r in res["runs"] if r["kernel"] == kernel]
    kw = dict(size=size, checkrun=3, npasses=2, input_mode=1 if mode == "seeded" else 0)
    for a, b in zip(extra[::2], extra[1::2]):
        kw[a.lstrip("-")] = float(b) if a in ("--alpha", "--beta") else int(b)
    ref = orc.run(kernel, "Base_Seq", **kw)
    assert run["problem_size"] == ref["problem_size"] and run["reps"] == ref["reps"]
    assert run["bytes_per_rep"] == ref["bytes_per_rep"]
    from fractions import Fraction
    ref_cs = Fraction(ref["checksum_hi"]) + Fraction(ref["checksum_lo"])
    diff = abs(float(run["checksum"] - ref_cs))
    assert diff < 1e-7, (kernel, float(run["checksum"]), float(ref_cs))  # test-raja-perf-suite.cpp:167
    if kernel in BIT_EXACT:
        assert run["checksum"] == ref_cs
    else:
        assert diff <= 1e-12 * max(1.0, abs(float(ref_cs)))


def test_harness_all_tunings_run_and_agree(rpx):
    res = run_suite(["-k", "Stream", "Polybench_JACOBI_2D", "Algorithm_REDUCE_SUM", "Polybench_GEMM",
                     "--size", "40000", "--checkrun", "2", "--input-mode", "seeded"])
    assert res["rc"] == 0, res.get("failures")
    by_kernel = {}
    for r in res["runs"]:
        by_kernel.setdefault(r["kernel"], []).append(r)
    for k, runs in by_kernel.items():
        assert len(runs) >= 2
        cs = [float(r["checksum"]) for r in runs]
        assert max(cs) - min(cs) < 1e-7, k
        assert all(r["time_total_s"] > 0 for r in runs)


def test_harness_autotuner_decides_caches_and_keeps_checksums(rpx, orc, tmp_path):
    """The "auto" tuning times its candidates before the timed reps, keeps
    the checksum of Base_Seq, records one decision per (kernel, shape, arch)
    in the cache file, and a second run reads the decisions back."""
    cache = str(tmp_path / "tuning.txt")
    kernels = ["Stream_TRIAD", "Stream_DOT", "Algorithm_SCAN", "Polybench_JACOBI_2D"]
    args = ["-k", *kernels, "--size", "16900", "--checkrun", "2", "--input-mode", "seeded", "--tsteps", "4",
            "-t", "auto", "--tuning-cache", cache]
    first = run_suite(args)
    assert first["rc"] == 0, first.get("failures")
    assert sorted(d["kernel"] for d in first["autotune"]) == sorted(kernels)
    assert not any(d["from_cache"] for d in first["autotune"])
    lines = open(cache).read().splitlines()
    assert len(lines) == len(kernels)
    second = run_suite(args)
    assert second["rc"] == 0
    assert all(d["from_cache"] for d in second["autotune"])
    assert {d["kernel"]: d["choice"] for d in second["autotune"]} == {d["kernel"]: d["choice"] for d in first["autotune"]}
    for r in first["runs"]:
        assert r["tuning"] == "auto"
        kw = dict(size=16900, checkrun=2, npasses=1, input_mode=1)
        if r["kernel"] == "Polybench_JACOBI_2D":
            kw["tsteps"] = 4
        ref = orc.run(r["kernel"], "Base_Seq", **kw)
        from fractions import Fraction
        ref_cs = Fraction(ref["checksum_hi"]) + Fraction(ref["checksum_lo"])
        if r["kernel"] in BIT_EXACT:
            assert r["checksum"] == ref_cs, r["kernel"]
        else:
            assert abs(float(r["checksum"] - ref_cs)) <= 1e-12 * max(1.0, abs(float(ref_cs))), r["kernel"]


def test_harness_progress_jsonl_and_resume(rpx, tmp_path):
    """--progress appends one JSON line per finished run; --resume restores
    the recorded runs of the same shape (checksums and times as recorded)
    and runs only the rest.  A record of another problem size is not
    restored."""
    import json
    prog = str(tmp_path / "progress.jsonl")
    common = ["--size", "100000", "--checkrun", "2", "--input-mode", "seeded", "-t", "default", "--progress", prog]
    first = run_suite(["-k", "Stream_TRIAD", "Stream_COPY", *common])
    assert first["rc"] == 0, first.get("failures")
    recs = [json.loads(line) for line in open(prog).read().splitlines()]
    assert sorted(r["kernel"] for r in recs) == ["Stream_COPY", "Stream_TRIAD"]
    by_kernel = {r["kernel"]: r for r in recs}

    second = run_suite(["-k", "Stream_TRIAD", "Stream_COPY", "Stream_MUL", *common, "--resume"])
    assert second["rc"] == 0, second.get("failures")
    recs2 = [json.loads(line) for line in open(prog).read().splitlines()]
    assert len(recs2) == 3 and recs2[-1]["kernel"] == "Stream_MUL"  # only the new kernel ran
    runs = {r["kernel"]: r for r in second["runs"]}
    assert set(runs) == {"Stream_TRIAD", "Stream_COPY", "Stream_MUL"}
    for k in ("Stream_TRIAD", "Stream_COPY"):
        for f in ("checksum_hi", "checksum_lo", "time_total_s", "passes"):
            assert runs[k][f] == by_kernel[k][f], (k, f)

    third_args = ["-k", "Stream_TRIAD", "--size", "50000", "--checkrun", "2", "--input-mode", "seeded", "-t",
                  "default", "--progress", prog, "--resume"]
    third = run_suite(third_args)
    assert third["rc"] == 0
    recs3 = [json.loads(line) for line in open(prog).read().splitlines()]
    assert len(recs3) == 4 and recs3[-1]["problem_size"] == 50000  # reran at the new size


def test_harness_roofline_csv_joins_counters(rpx, tmp_path):
    """<prefix>-roofline.csv: per run, the kernel's algorithmic bytes and
    FLOPs, their rate, the bound they imply and the fraction of it, the
    checksum difference from the first run; --counters joins each counter's
    mean value per dispatch over the device kernels behind the suite kernel."""
    import csv
    import os
    cnt = tmp_path / "counter_collection.csv"
    triad = '"void (anonymous namespace)::stream_vec_kernel<3, 1, 256, true>(double*, double const*, double const*)"'
    dot = '"void (anonymous namespace)::reduce_kernel<true, 256, 4, true>(double const*, double const*)"'
    cnt.write_text("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
                   f"1,{triad},FETCH_SIZE,10\n2,{triad},FETCH_SIZE,30\n3,{dot},FETCH_SIZE,7\n"
                   f"1,{triad},WRITE_SIZE,4\n")
    n = 100_000
    res = run_suite(["-k", "Stream_TRIAD", "Stream_COPY", "--size", str(n), "--checkrun", "2", "--input-mode",
                     "seeded", "--counters", str(cnt)])
    assert res["rc"] == 0, res.get("failures")
    rows = list(csv.DictReader(open(os.path.join(res["outdir"], "RAJAPerf-roofline.csv"))))
    triad_rows = [r for r in rows if r["Kernel"] == "Stream_TRIAD"]
    copy_rows = [r for r in rows if r["Kernel"] == "Stream_COPY"]
    assert triad_rows and copy_rows
    assert float(triad_rows[0]["Checksum diff"]) == 0.0  # the first run is the reference
    for r in triad_rows:
        assert float(r["FETCH_SIZE per dispatch"]) == 20.0
        assert float(r["WRITE_SIZE per dispatch"]) == 4.0
        assert int(r["Bytes/rep"]) == 24 * n and int(r["FLOPs/rep"]) == 2 * n
        gbs = int(r["Bytes/rep"]) / float(r["Time/rep (s)"]) / 1e9
        assert abs(float(r["GB/s"]) - gbs) <= 1e-6 * gbs
        assert r["Bound"] == "hbm" and abs(float(r["Frac of bound"]) - gbs / 8000.0) <= 1e-6 * gbs
        assert float(r["Checksum diff"]) == 0.0  # every TRIAD tuning is bit-exact
    for r in copy_rows:
        assert r["FETCH_SIZE per dispatch"] == "" and r["WRITE_SIZE per dispatch"] == ""


@pytest.mark.parametrize("kernel,size,extra", [("Polybench_GEMM", 100 * 100, []), ("Polybench_GEMM", 97 * 97, ["--nk", "131"]),
                                               ("Apps_MASS3DPA", 125 * 300, []),
                                               ("Apps_MASS3DPA", 64 * 200, ["--d1d", "3", "--q1d", "4"])])
@pytest.mark.parametrize("mode", ["reference", "seeded"])
def test_harness_mixed_tunings_match_base_seq(rpx, orc, kernel, size, extra, mode):
    """The "mixed" tunings (int8 MFMA on 7 Ozaki slices) run through
    KernelBase like the others and keep Base_Seq's checksum to 1e-12."""
    from fractions import Fraction
    args = ["-k", kernel, "--size", str(size), "--checkrun", "2", "--input-mode", mode, "-t", "mixed", *extra]
    res = run_suite(args)
    assert res["rc"] == 0, res.get("failures")
    (run,) = [r for r in res["runs"] if r["kernel"] == kernel]
    assert run["tuning"] == "mixed"
    kw = dict(size=size, checkrun=2, input_mode=1 if mode == "seeded" else 0)
    for a, b in zip(extra[::2], extra[1::2]):
        kw[a.lstrip("-")] = float(b) if a in ("--alpha", "--beta") else int(b)
    ref = orc.run(kernel, "Base_Seq", **kw)
    ref_cs = Fraction(ref["checksum_hi"]) + Fraction(ref["checksum_lo"])
    assert abs(float(run["checksum"] - ref_cs)) <= 1e-12 * max(1.0, abs(float(ref_cs)))


MIXED_CASES = [("Apps_MASS3DPA", 125 * 300, []), ("Apps_DIFFUSION3DPA", 64 * 300, []),
               ("Apps_DIFFUSION3DPA", 125 * 200, ["--d1d", "4"])]


@pytest.mark.parametrize("mode", ["reference", "seeded"])
@pytest.mark.parametrize("kernel,size,extra", MIXED_CASES, ids=[f"{c[0]}-{i}" for i, c in enumerate(MIXED_CASES)])
def test_harness_mixed_tuning_matches_base_seq(rpx, orc, kernel, size, extra, mode):
    """The "mixed" tuning (int8 MFMA slices of the basis contractions) through
    KernelBase::execute: the suite's 1e-7 against Base_Seq, and 1e-11
    relative to the checksum of |Y|-scale."""
    args = ["-k", kernel, "--size", str(size), "--checkrun", "3", "--input-mode", mode, "-t", "mixed", *extra]
    res = run_suite(args)
    assert res["rc"] == 0, res.get("failures")
    (run,) = [r for r in res["runs"] if r["kernel"] == kernel]
    assert run["tuning"] == "mixed"
    kw = dict(size=size, checkrun=3, npasses=1, input_mode=1 if mode == "seeded" else 0)
    for a, b in zip(extra[::2], extra[1::2]):
        kw[a.lstrip("-")] = float(b) if a in ("--alpha", "--beta") else int(b)
    ref = orc.run(kernel, "Base_Seq", **kw)
    from fractions import Fraction
    ref_cs = Fraction(ref["checksum_hi"]) + Fraction(ref["checksum_lo"])
    diff = abs(float(run["checksum"] - ref_cs))
    assert diff < 1e-7, (kernel, float(run["checksum"]), float(ref_cs))
    assert diff <= 1e-11 * max(1.0, abs(float(ref_cs)))


GRAPH_KERNELS = ["Stream_COPY", "Stream_MUL", "Stream_ADD", "Stream_TRIAD", "Stream_DOT", "Algorithm_REDUCE_SUM",
                 "Algorithm_SCAN", "Polybench_JACOBI_2D", "Polybench_GEMM", "Apps_MASS3DPA", "Apps_DIFFUSION3DPA"]


@pytest.mark.parametrize("mode", ["reference", "seeded"])
def test_harness_graph_tunings_match_base_seq(rpx, orc, mode):
    """--graph: every tuning also runs as "<tuning>-graph", its rep loop
    captured in a HIP graph before the timer and replayed inside it, with
    setUp, checksum and tearDown outside the graph.  Each graph tuning's
    checksum is held to the oracle's Base_Seq like the eager ones (three
    passes, so the same rep loop is captured and replayed three times with
    eager work between); SORT, whose calls upload host tables, has none."""
    size = 40_000
    args = ["-k", *GRAPH_KERNELS, "Algorithm_SORT", "--size", str(size), "--checkrun", "3", "--npasses", "3",
            "--input-mode", mode, "--tsteps", "4", "--graph", "-t", "default", "default-graph", "mixed",
            "mixed-graph"]
    res = run_suite(args)
    assert res["rc"] == 0, res.get("failures")
    by = {(r["kernel"], r["tuning"]): r for r in res["runs"]}
    assert ("Algorithm_SORT", "default") in by and ("Algorithm_SORT", "default-graph") not in by
    for k in GRAPH_KERNELS:
        g = by[(k, "default-graph")]
        assert g["passes"] == 3 and g["time_total_s"] > 0
        kw = dict(size=size, checkrun=3, npasses=3, input_mode=1 if mode == "seeded" else 0)
        if k == "Polybench_JACOBI_2D":
            kw["tsteps"] = 4
        ref = orc.run(k, "Base_Seq", **kw)
        from fractions import Fraction
        ref_cs = Fraction(ref["checksum_hi"]) + Fraction(ref["checksum_lo"])
        for t in ("default-graph", "mixed-graph"):
            if (k, t) not in by:
                continue
            cs = by[(k, t)]["checksum"]
            if k in BIT_EXACT and t == "default-graph":
                assert cs == ref_cs, (k, t)
            else:
                assert abs(float(cs - ref_cs)) <= 1e-11 * max(1.0, abs(float(ref_cs))), (k, t)
    for k in ("Polybench_GEMM", "Apps_MASS3DPA", "Apps_DIFFUSION3DPA"):
        assert (k, "mixed-graph") in by
    timing = open(os.path.join(res["outdir"], "RAJAPerf-timing-Average.csv")).read()
    assert "default-graph" in timing


def test_harness_timing_combiners_and_iweighted_checksum(rpx, orc):
    """--npasses-combiners writes one timing report per combiner (seconds per
    pass: mean, minimum, maximum over the passes), and every run carries the
    i-weighted FP64 sum, equal to the oracle's Base_Seq for bit-exact kernels
    (the same sequential sum over the same values)."""
    kernels = ["Stream_TRIAD", "Stream_COPY", "Polybench_JACOBI_2D", "Apps_MASS3DPA", "Stream_DOT",
               "Algorithm_REDUCE_SUM"]
    res = run_suite(["-k", *kernels, "--size", "30000", "--checkrun", "2", "--npasses", "3", "--input-mode",
                     "seeded", "--tsteps", "3", "-t", "default", "--npasses-combiners", "Average", "Minimum",
                     "Maximum"])
    assert res["rc"] == 0, res.get("failures")
    tables = {}
    for c in ("Average", "Minimum", "Maximum"):
        rows = list(csv.reader(open(os.path.join(res["outdir"], f"RAJAPerf-timing-{c}.csv"))))
        assert rows[0][0].startswith({"Average": "Mean", "Minimum": "Min", "Maximum": "Max"}[c])
        tables[c] = {(r[0], r[2]): float(r[3]) for r in rows[2:] if r[3] not in ("", "Not run")}
    for key, avg in tables["Average"].items():
        assert tables["Minimum"][key] <= avg <= tables["Maximum"][key]
    for r in res["runs"]:
        kw = dict(size=30000, checkrun=2, npasses=3, input_mode=1)
        if r["kernel"] == "Polybench_JACOBI_2D":
            kw["tsteps"] = 3
        ref = orc.run(r["kernel"], "Base_Seq", **kw)
        if r["kernel"] in BIT_EXACT:
            assert r["iweighted"] == ref["iweighted"], r["kernel"]
        else:
            assert abs(r["iweighted"] - ref["iweighted"]) <= 1e-12 * abs(ref["iweighted"]) * 30000, r["kernel"]
    report = open(os.path.join(res["outdir"], "RAJAPerf-checksum.txt")).read()
    assert "i-weighted FP64 sum" in report
