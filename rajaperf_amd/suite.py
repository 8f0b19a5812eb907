"""In-process driver of the C++ suite harness (lib/librajaperf_suite.so).

`run_suite(args)` runs the harness with a command line, exactly like the
rajaperf-mi355x executable, and returns the parsed `--json` results: one
record per executed (kernel, variant, tuning) with its checksum (as an exact
hi+lo double pair of the long-double checksum), timings and shape.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import tempfile
from fractions import Fraction
from typing import Sequence

from .abi import SUITE_LIB, ExtensionMissing

_suite = None


def lib():
    global _suite
    if _suite is None:
        if not os.path.exists(SUITE_LIB):
            raise ExtensionMissing(f"{SUITE_LIB} is not built")
        L = C.CDLL(SUITE_LIB)
        L.rajaperf_run.restype = C.c_int
        L.rajaperf_run.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        _suite = L
    return _suite


def run_suite(args: Sequence[str]) -> dict:
    """Runs the harness in this process; returns {"rc", "status", "runs", ...}."""
    fd, path = tempfile.mkstemp(suffix=".json")
    os.close(fd)
    outdir = tempfile.mkdtemp()
    argv = ["rajaperf-mi355x", *args, "--json", path, "-od", outdir]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    rc = lib().rajaperf_run(len(argv), arr)
    result = {"rc": rc, "runs": [], "outdir": outdir}
    try:
        if os.path.getsize(path):
            result.update(json.load(open(path)))
    finally:
        os.unlink(path)
    for r in result["runs"]:
        r["checksum"] = Fraction(r["checksum_hi"]) + Fraction(r["checksum_lo"])
    return result
