"""MI355X-native execution variant for a subset of the RAJA Performance Suite.

Layout:
  csrc/    hand-written gfx950 HIP kernels behind the C ABI declared in
           include/rajaperf_mi355x.h  ->  lib/librajaperf_mi355x.so
  suite/   C++ harness mirroring the suite's KernelBase plugin API with the
           new MI355X_Native variant  ->  lib/librajaperf_suite.so,
           bin/rajaperf-mi355x
  abi.py   ctypes binding of the C ABI (tests, bench)
  suite.py in-process driver of the C++ harness (tests)
"""

from .abi import ExtensionMissing, RpxError, lib  # noqa: F401

__all__ = ["ExtensionMissing", "RpxError", "lib"]
