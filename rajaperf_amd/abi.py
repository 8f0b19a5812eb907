"""ctypes binding of the MI355X variant's C ABI (include/rajaperf_mi355x.h).

The binding loads ``rajaperf_amd/lib/librajaperf_mi355x.so`` and raises
``ExtensionMissing`` when it has not been built: there is no CPU or
library fallback behind any entry point.  Device memory is handled through
the library's own runtime entry points (``DeviceBuffer``), so no other GPU
framework is loaded into the process.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
KERNEL_LIB = os.path.join(LIB_DIR, "librajaperf_mi355x.so")
SUITE_LIB = os.path.join(LIB_DIR, "librajaperf_suite.so")
SUITE_BIN = os.path.join(_HERE, "bin", "rajaperf-mi355x")

RPX_SUCCESS = 0
RPX_ERR_INVALID_ARGUMENT = -1001
RPX_ERR_SCRATCH_TOO_SMALL = -1002
RPX_ERR_UNSUPPORTED = -1003
RPX_ERR_NO_DEVICE = -1004
RPX_ERR_WAIT_EXPIRED = -1005

RPX_TUNE_CACHED = 0x1
RPX_TUNE_NAIVE = 0x2
RPX_TUNE_JACOBI_SKIP_B = 0x4
RPX_TUNE_MASS_SINGLE_LAUNCH = 0x8
RPX_TUNE_GEMM_BETA = 0x10


class ExtensionMissing(RuntimeError):
    """The HIP library has not been built (run __graft_entry__.build())."""


class RpxError(RuntimeError):
    def __init__(self, func: str, status: int, msg: str):
        super().__init__(f"{func} failed with status {status}: {msg}")
        self.status = status


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("items_per_thread", C.c_int32),
                ("grid_size", C.c_int32), ("flags", C.c_int32), ("depth", C.c_int32)]


class MassBatch(C.Structure):
    _fields_ = [("B", C.c_void_p), ("Bt", C.c_void_p), ("D", C.c_void_p), ("X", C.c_void_p), ("Y", C.c_void_p),
                ("ne", C.c_int64), ("d1d", C.c_int32), ("q1d", C.c_int32)]


_vp, _i64, _u64, _d, _i32, _sz = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int, C.c_size_t
_tp = C.POINTER(Tuning)
_i64p = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
_SIGS = {
    "rpx_abi_version": (_i32, []),
    "rpx_status_string": (C.c_char_p, [_i32]),
    "rpx_init": (_i32, [_i32, C.c_char_p, _sz]),
    "rpx_device_count": (_i32, [C.POINTER(_i32)]),
    "rpx_device_malloc": (_i32, [C.POINTER(_vp), _sz]),
    "rpx_device_free": (_i32, [_vp]),
    "rpx_host_malloc_pinned": (_i32, [C.POINTER(_vp), _sz]),
    "rpx_host_free_pinned": (_i32, [_vp]),
    "rpx_memcpy_h2d": (_i32, [_vp, _vp, _sz]),
    "rpx_memcpy_d2h": (_i32, [_vp, _vp, _sz]),
    "rpx_memcpy_d2d": (_i32, [_vp, _vp, _sz, _vp]),
    "rpx_stream_create": (_i32, [C.POINTER(_vp)]),
    "rpx_stream_destroy": (_i32, [_vp]),
    "rpx_stream_synchronize": (_i32, [_vp]),
    "rpx_device_synchronize": (_i32, []),
    "rpx_graph_begin_capture": (_i32, [_vp]),
    "rpx_graph_end_capture": (_i32, [_vp, C.POINTER(_vp)]),
    "rpx_graph_launch": (_i32, [_vp, _vp]),
    "rpx_graph_destroy": (_i32, [_vp]),
    "rpx_event_create": (_i32, [C.POINTER(_vp)]),
    "rpx_event_destroy": (_i32, [_vp]),
    "rpx_event_record": (_i32, [_vp, _vp]),
    "rpx_event_elapsed_ms": (_i32, [C.POINTER(C.c_float), _vp, _vp]),
    "rpx_fill_uniform_f64": (_i32, [_vp, _i64, _u64, _u64, _d, _d, _vp]),
    "rpx_fill_const_f64": (_i32, [_vp, _i64, _d, _vp]),
    "rpx_fill_suite_formula_f64": (_i32, [_vp, _i64, _d, _vp]),
    "rpx_fill_grid_f64": (_i32, [_vp, _i64, _u64, _u64, _d, _d, _d, _vp]),
    "rpx_fill_sym_tensor_f64": (_i32, [_vp, _i64, _i32, _i32, _u64, _u64, _vp]),
    "rpx_copy_f64": (_i32, [_vp, _vp, _i64, _tp, _vp]),
    "rpx_mul_f64": (_i32, [_vp, _vp, _d, _i64, _tp, _vp]),
    "rpx_add_f64": (_i32, [_vp, _vp, _vp, _i64, _tp, _vp]),
    "rpx_triad_f64": (_i32, [_vp, _vp, _vp, _d, _i64, _tp, _vp]),
    "rpx_reduce_scratch_bytes": (_sz, [_i64, _tp]),
    "rpx_reduce_scratch_init": (_i32, [_vp, _sz, _vp]),
    "rpx_dot_f64": (_i32, [_vp, _vp, _i64, _vp, _i32, _vp, _sz, _tp, _vp]),
    "rpx_reduce_sum_f64": (_i32, [_vp, _i64, _vp, _vp, _sz, _tp, _vp]),
    "rpx_dot_noalloc_max_scratch_bytes": (_sz, []),
    "rpx_dot_f64_noalloc": (_i32, [_vp, _vp, _i64, _vp, _i32, _tp, _vp]),
    "rpx_scan_scratch_bytes": (_sz, [_i64, _tp]),
    "rpx_scan_scratch_init": (_i32, [_vp, _sz, _vp]),
    "rpx_scan_scratch_status": (_i32, [_vp]),
    "rpx_scan_exclusive_f64": (_i32, [_vp, _vp, _i64, _vp, _sz, _tp, _vp]),
    "rpx_scan_inclusive_f64": (_i32, [_vp, _vp, _i64, _vp, _sz, _tp, _vp]),
    "rpx_scan_segmented_scratch_bytes": (_sz, [_i64p, _i64]),
    "rpx_scan_exclusive_segmented_f64": (_i32, [_vp, _vp, _i64p, _i64, _vp, _sz, _vp]),
    "rpx_sort_segmented_scratch_bytes": (_sz, [_i64p, _i64]),
    "rpx_sort_segmented_f64": (_i32, [_vp, _i64p, _i64, _vp, _sz, _vp]),
    "rpx_jacobi2d_sweep_f64": (_i32, [_vp, _vp, _i64, _tp, _vp]),
    "rpx_jacobi2d_f64": (_i32, [_vp, _vp, _i64, _i64, _tp, _vp]),
    "rpx_jacobi2d_fused_scratch_bytes": (_sz, [_i64]),
    "rpx_jacobi2d_fused_f64": (_i32, [_vp, _vp, _i64, _i64, _vp, _sz, _tp, _vp]),
    "rpx_jacobi2d_fused_launch_f64": (_i32, [_vp, _vp, _vp, _i64, _i32, _tp, _vp]),
    "rpx_jacobi2d_copy_boundary_f64": (_i32, [_vp, _vp, _i64, _vp]),
    "rpx_gemm_f64": (_i32, [_vp, _vp, _vp, _d, _d, _i64, _i64, _i64, _tp, _vp]),
    "rpx_gemm_mixed_scratch_bytes": (_sz, [_i64, _i64, _i64, _i32]),
    "rpx_gemm_mixed_bits": (_i32, [_i64, _i32]),
    "rpx_gemm_mixed_f64": (_i32, [_vp, _vp, _vp, _d, _d, _i64, _i64, _i64, _i32, _vp, _sz, _vp]),
    "rpx_gemm_mixed_tuned_f64": (_i32, [_vp, _vp, _vp, _d, _d, _i64, _i64, _i64, _i32, _tp, _vp, _sz, _vp]),
    "rpx_mass3dpa_f64": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _tp, _vp]),
    "rpx_mass3dpa_batched_scratch_bytes": (_sz, [C.POINTER(MassBatch), _i32]),
    "rpx_mass3dpa_batched_plan": (_i32, [C.POINTER(MassBatch), _i32, _vp, _sz, _vp]),
    "rpx_mass3dpa_batched_run": (_i32, [_vp, _tp, _vp]),
    "rpx_mass3dpa_batched_f64": (_i32, [C.POINTER(MassBatch), _i32, _vp, _sz, _tp, _vp]),
    "rpx_mass3dpa_mixed_scratch_bytes": (_sz, [_i32, _i32, _i32]),
    "rpx_mass3dpa_mixed_f64": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp, _sz, _vp]),
    "rpx_diffusion3dpa_f64": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _tp, _vp]),
    "rpx_diffusion3dpa_mixed_scratch_bytes": (_sz, [_i32, _i32, _i32]),
    "rpx_diffusion3dpa_mixed_f64": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _vp, _sz, _vp]),
}

# Entry points whose return value is a size, not a status.
_SIZE_FUNCS = {name for name, (res, _) in _SIGS.items() if res is _sz}


def declared_symbols(header: Optional[str] = None) -> list[str]:
    """Names of the functions declared in include/rajaperf_mi355x.h."""
    import re
    header = header or os.path.join(os.path.dirname(_HERE), "include", "rajaperf_mi355x.h")
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rpx_[a-z0-9_]+)\s*\(", text)))


class _Lib:
    def __init__(self, path: str):
        if not os.path.exists(path):
            raise ExtensionMissing(
                f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        self._dll = C.CDLL(path)
        self.path = path
        self.missing = []
        for name, (res, args) in _SIGS.items():
            try:
                fn = getattr(self._dll, name)
            except AttributeError:
                self.missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args

    def raw(self, name: str):
        return getattr(self._dll, name)

    def __getattr__(self, name: str):
        if not name.startswith("rpx_"):
            raise AttributeError(name)
        fn = getattr(self._dll, name)
        if name in _SIZE_FUNCS or name in ("rpx_abi_version", "rpx_status_string", "rpx_gemm_mixed_bits",
                                           "rpx_scan_scratch_status"):
            return fn

        def checked(*args):
            st = fn(*args)
            if st != RPX_SUCCESS:
                msg = self._dll.rpx_status_string(st).decode()
                raise RpxError(name, st, msg)
            return st
        return checked


_LIB: Optional[_Lib] = None


def lib() -> _Lib:
    global _LIB
    if _LIB is None:
        _LIB = _Lib(KERNEL_LIB)
    return _LIB


def tuning(block_size=0, items_per_thread=0, grid_size=0, flags=0, depth=0):
    return C.pointer(Tuning(block_size, items_per_thread, grid_size, flags, depth))


# ---- device helpers ---------------------------------------------------------

def init(device: int = 0) -> str:
    buf = C.create_string_buffer(64)
    lib().rpx_init(device, buf, 64)
    return buf.value.decode()


def device_count() -> int:
    c = C.c_int(0)
    st = lib().raw("rpx_device_count")(C.byref(c))
    return c.value if st == RPX_SUCCESS else 0


class DeviceBuffer:
    """A device allocation owned by Python (freed on close / GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        lib().rpx_device_malloc(C.byref(p), self.nbytes)
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            lib().rpx_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes)
        return b

    def to_array(self, dtype=np.float64, count: Optional[int] = None, offset_bytes: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset_bytes) // itemsize
        out = np.empty(count, dtype=dtype)
        if count:
            lib().rpx_memcpy_d2h(out.ctypes.data, self.ptr + offset_bytes, count * itemsize)
        return out

    def upload(self, a: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset_bytes + a.nbytes <= self.nbytes
        if a.nbytes:
            lib().rpx_memcpy_h2d(self.ptr + offset_bytes, a.ctypes.data, a.nbytes)

    def close(self) -> None:
        if self.ptr:
            lib().rpx_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __int__(self):
        return self.ptr


def synchronize() -> None:
    lib().rpx_device_synchronize()


def offsets_array(offsets: Sequence[int]):
    arr = (C.c_int64 * len(offsets))(*[int(o) for o in offsets])
    return arr


class Stream:
    """A created (non-blocking) HIP stream; graphs can only capture these."""

    def __init__(self):
        p = C.c_void_p()
        lib().rpx_stream_create(C.byref(p))
        self.ptr = p.value

    def synchronize(self):
        lib().rpx_stream_synchronize(self.ptr)

    def __del__(self):
        try:
            if self.ptr:
                lib().rpx_stream_destroy(self.ptr)
        except Exception:
            pass


class Graph:
    """Calls made on `stream` inside `with Graph(stream) as g:` are captured;
    g.launch() replays them on the stream."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.ptr = None

    def __enter__(self):
        lib().rpx_graph_begin_capture(self.stream.ptr)
        return self

    def __exit__(self, exc_type, exc, tb):
        p = C.c_void_p()
        lib().rpx_graph_end_capture(self.stream.ptr, C.byref(p))
        self.ptr = p.value
        return False

    def launch(self):
        lib().rpx_graph_launch(self.ptr, self.stream.ptr)

    def __del__(self):
        try:
            if self.ptr:
                lib().rpx_graph_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        lib().rpx_event_create(C.byref(p))
        self.ptr = p.value

    def record(self, stream=None):
        lib().rpx_event_record(self.ptr, stream)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        lib().rpx_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr)
        return float(ms.value)

    def __del__(self):
        try:
            if self.ptr:
                lib().rpx_event_destroy(self.ptr)
        except Exception:
            pass
