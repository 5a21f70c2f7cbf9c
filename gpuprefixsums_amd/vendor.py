"""The vendor comparator: rocPRIM's device scan (ctypes binding of
libgps_vendor.so, the C entry points in include/gps_vendor.h).

Benchmark and test comparator only, standing in for the reference's CUB
comparator (GPUPrefixSumsCUDA/CubDispatcher.cuh:21-76). It lives in its own
library so the product library libgps.so carries no vendor scan; the
comparator links libgps.so for its fill and validation kernels and reports
errors through gps_last_error().
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import PKG, GpsError

VENDOR_LIB_PATH = PKG / "libgps_vendor.so"

_vp, _u64, _u32, _int = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
_fp = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes) for every function in include/gps_vendor.h.
SIGNATURES = {
    "gps_vendor_scan_temp_bytes": (_int, [_u64, _int, ctypes.POINTER(_u64)]),
    "gps_vendor_scan_u32": (_int, [_vp, _vp, _u64, _int, _vp, _u64, _vp]),
    "gps_vendor_time_scan_u32": (_int, [_vp, _vp, _u64, _int, _vp, _u64, _vp, _u32, _fp]),
    "gps_vendor_batch_timing": (_int, [_u32, _u32, _u32, ctypes.POINTER(ctypes.c_double)]),
}

_vlib = None


def load() -> ctypes.CDLL:
    """Loads libgps_vendor.so (after libgps.so, which it links); raises if it
    has not been built."""
    global _vlib
    if _vlib is not None:
        return _vlib
    _lib.load()
    if not VENDOR_LIB_PATH.exists():
        raise ImportError(f"{VENDOR_LIB_PATH} is not built; run `python -m gpuprefixsums_amd.build`")
    L = ctypes.CDLL(str(VENDOR_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _vlib = L
    return L


def _check(rc: int) -> None:
    if rc != _lib.GPS_OK:
        raise GpsError(rc, _lib.load().gps_last_error().decode())


def _ptr(x) -> int:
    from . import _ptr as p
    return p(x)


def _stream(s, *bufs):
    from . import _stream as st
    return st(s, *bufs)


def vendor_scan_temp_bytes(n: int, inclusive: bool = True) -> int:
    b = ctypes.c_uint64()
    _check(load().gps_vendor_scan_temp_bytes(n, int(inclusive), ctypes.byref(b)))
    return b.value


def vendor_scan_u32(d_in, d_out, n: int, d_temp, temp_bytes: int, inclusive: bool = True, stream=None) -> None:
    """rocPRIM's device scan of n uint32 (d_in == d_out allowed); d_temp holds
    at least vendor_scan_temp_bytes(n, inclusive) bytes."""
    _check(load().gps_vendor_scan_u32(_ptr(d_in), _ptr(d_out), n, int(inclusive), _ptr(d_temp), temp_bytes,
                                      _stream(stream, d_in, d_out)))


def vendor_time_scan_u32(d_in, d_out, n: int, d_temp, temp_bytes: int, inclusive: bool = True, iters: int = 10,
                         stream=None) -> list:
    """Per-launch milliseconds of `iters` back-to-back rocPRIM scans."""
    ms = (ctypes.c_float * iters)()
    _check(load().gps_vendor_time_scan_u32(_ptr(d_in), _ptr(d_out), n, int(inclusive), _ptr(d_temp), temp_bytes,
                                           _stream(stream, d_in, d_out), iters, ms))
    return list(ms)


class CubDispatcher:
    """The reference's vendor comparator (CubDispatcher.cuh:21-76) on rocPRIM:
    BatchTimingCubChainedScan times the vendor inclusive scan in place over
    all-ones input, first iteration discarded, and returns keys/s."""

    def __init__(self, max_size: int):
        self.max_size = max_size

    def batch_timing_cub_chained_scan(self, size: int, batch_count: int) -> float:
        keys = ctypes.c_double()
        _check(load().gps_vendor_batch_timing(self.max_size, size, batch_count, ctypes.byref(keys)))
        return keys.value


__all__ = ["vendor_scan_temp_bytes", "vendor_scan_u32", "vendor_time_scan_u32", "CubDispatcher", "load"]
