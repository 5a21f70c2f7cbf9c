"""gpuprefixsums_amd: MI355X (gfx950) device-wide prefix sums.

Python host mirror of the C ABI in include/gps.h (libgps.so, hand-written HIP
kernels). Device buffers are passed as raw device addresses or as any object
with .data_ptr() (e.g. torch tensors on a ROCm device); PyTorch is only used
by callers for memory and streams. There is no CPU fallback: if libgps.so is
not built, importing the binding raises.

Reference interface mirrored (b0nes164/GPUPrefixSums):
  ChainedScanDecoupledLookbackDispatcher  GPUPrefixSumsCUDA/ChainedScanDecoupledLookbackDispatcher.cuh:16-163
  ReduceThenScanDispatcher                GPUPrefixSumsCUDA/ReduceThenScanDispatcher.cuh:16-152
  EmulatedDeadlockingDispatcher           GPUPrefixSumsCUDA/EmulatedDeadlockingDispatcher.cuh:19-163
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

from . import _lib
from ._lib import (ALGO_CSDL, ALGO_CSDLDF, ALGO_RTS, SCHEDULE_CONTROL, SCHEDULE_DEFAULT, SCHEDULE_PERSISTENT,
                   SCHEDULE_SPLIT, SCHEDULE_TILE, GpsError, GpsTuning, check, load)

__all__ = [
    "ALGO_CSDL", "ALGO_RTS", "ALGO_CSDLDF", "GpsError", "ScanHandle",
    "ChainedScanDecoupledLookbackDispatcher", "ReduceThenScanDispatcher",
    "EmulatedDeadlockingDispatcher", "fill_u32", "fill_random_u32", "validate_scan_u32",
    "validate_ones_u32", "load", "SCHEDULE_DEFAULT", "SCHEDULE_TILE", "SCHEDULE_PERSISTENT", "SCHEDULE_CONTROL",
    "SCHEDULE_SPLIT",
    "FP_PLAIN", "FP_PAIR", "WAVE_INCLUSIVE_SCAN", "WAVE_SUM", "WAVE_EXCLUSIVE_SCAN",
    "wave_primitive_u32", "SURVEY_NETWORKS", "survey_segment", "survey_scan_u32",
]


def _ptr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    raise TypeError(f"expected a device pointer or tensor, got {type(x)!r}")


def _stream(s, *bufs) -> Optional[int]:
    """The stream to launch on: `s` if given, else torch's current stream of
    the first device tensor among `bufs` (so tensor callers get stream-ordered
    launches), else the null stream."""
    if s is None:
        for b in bufs:
            if getattr(b, "is_cuda", False):
                import torch
                return int(torch.cuda.current_stream(b.device).cuda_stream)
        return None
    if isinstance(s, int):
        return s
    if hasattr(s, "cuda_stream"):
        return int(s.cuda_stream)
    raise TypeError(f"expected a stream handle, got {type(s)!r}")


def _numel(x) -> int:
    if hasattr(x, "numel"):
        return int(x.numel())
    raise TypeError("n must be given for raw pointers")


def _need(what: str, nbytes: int, *bufs) -> None:
    """Raises ValueError when a sized buffer (anything with numel() and
    element_size(), e.g. a torch tensor) holds fewer than nbytes bytes, so an
    explicit n can never make a kernel run past a buffer. Raw addresses are
    the caller's responsibility."""
    for b in bufs:
        if b is not None and hasattr(b, "numel") and hasattr(b, "element_size"):
            have = int(b.numel()) * int(b.element_size())
            if have < nbytes:
                raise ValueError(f"{what}: a buffer holds {have} bytes but the call needs {nbytes}")


class ScanHandle:
    """One C-ABI handle: owns the status words / ticket / tile-sum scratch for
    scans of up to max_n elements. Not re-entrant: use one per stream."""

    def __init__(self, algo: int = ALGO_CSDL, max_n: int = 1 << 28, *, block_threads: int = 0,
                 vec_per_thread: int = 0, spin_limit: int = 0, stall_period: int = 0, schedule: int = 0):
        L = load()
        t = GpsTuning(block_threads, vec_per_thread, spin_limit, stall_period, schedule)
        h = ctypes.c_void_p()
        check(L.gps_create(algo, max_n, ctypes.byref(t), ctypes.byref(h)))
        self._h = h
        self.algo = algo
        self.max_n = max_n
        out = GpsTuning()
        tile = ctypes.c_uint32()
        check(L.gps_get_tuning(h, ctypes.byref(out), ctypes.byref(tile)))
        self.block_threads = out.block_threads
        self.vec_per_thread = out.vec_per_thread
        self.spin_limit = out.spin_limit
        self.stall_period = out.stall_period
        self.schedule = out.schedule
        self.tile = tile.value

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("handle closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            check(load().gps_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def inclusive(self, d_in, d_out, n: Optional[int] = None, stream=None) -> None:
        n = _numel(d_in) if n is None else n
        _need("inclusive", 4 * n, d_in, d_out)
        check(load().gps_inclusive_scan_u32(self.handle, _ptr(d_in), _ptr(d_out), n, _stream(stream, d_in, d_out)))

    def exclusive(self, d_in, d_out, n: Optional[int] = None, stream=None) -> None:
        n = _numel(d_in) if n is None else n
        _need("exclusive", 4 * n, d_in, d_out)
        check(load().gps_exclusive_scan_u32(self.handle, _ptr(d_in), _ptr(d_out), n, _stream(stream, d_in, d_out)))

    def scan(self, d_in, d_out, n: Optional[int] = None, inclusive: bool = True, stream=None) -> None:
        n = _numel(d_in) if n is None else n
        _need("scan", 4 * n, d_in, d_out)
        check(load().gps_scan_u32(self.handle, _ptr(d_in), _ptr(d_out), n, int(inclusive),
                                  _stream(stream, d_in, d_out)))

    def scan_carry(self, d_in, d_out, n: int, inclusive: bool = True, d_carry_in=None, d_carry_out=None,
                   stream=None) -> None:
        """Scan continued from a device carry word (d_carry_in, may be None);
        d_carry_out (may be None, may equal d_carry_in) receives the total plus
        the carry."""
        _need("scan_carry", 4 * n, d_in, d_out)
        _need("scan_carry", 4, d_carry_in, d_carry_out)
        check(load().gps_scan_u32_carry(self.handle, _ptr(d_in), _ptr(d_out), n, int(inclusive),
                                        _ptr(d_carry_in), _ptr(d_carry_out), _stream(stream, d_in, d_out)))

    def time(self, d_in, d_out, n: int, inclusive: bool = True, iters: int = 10, stream=None) -> list:
        ms = (ctypes.c_float * iters)()
        _need("time", 4 * n, d_in, d_out)
        check(load().gps_time_scan_u32(self.handle, _ptr(d_in), _ptr(d_out), n, int(inclusive),
                                       _stream(stream, d_in, d_out), iters, ms))
        return list(ms)

    def scan_u32x4(self, d_in, d_out, n: int, inclusive: bool = True, stream=None) -> None:
        """Member-wise scan of n 4-member uint32 structs (uint4 elements)."""
        fn = load().gps_inclusive_scan_u32x4 if inclusive else load().gps_exclusive_scan_u32x4
        _need("scan_u32x4", 16 * n, d_in, d_out)
        check(fn(self.handle, _ptr(d_in), _ptr(d_out), n, _stream(stream, d_in, d_out)))

    def estimate_occupancy(self, d_in, d_out, n: int, struct_scan: bool = False, stream=None) -> tuple:
        """Runs the persistent inclusive scan with one workgroup per tile and
        returns (workgroups that took a tile, resident workgroups of the normal
        persistent launch)."""
        measured, resident = ctypes.c_uint32(), ctypes.c_uint32()
        _need("estimate_occupancy", (16 if struct_scan else 4) * n, d_in, d_out)
        check(load().gps_estimate_occupancy(self.handle, _ptr(d_in), _ptr(d_out), n, int(struct_scan),
                                            _stream(stream, d_in, d_out), ctypes.byref(measured),
                                            ctypes.byref(resident)))
        return measured.value, resident.value

    def inclusive_f32(self, d_in, d_out, n: Optional[int] = None, compensated=True, stream=None) -> None:
        """FP32 -> FP32 inclusive scan. compensated: False/0 plain FP32 carries,
        True/1 two-float TwoSum pairs."""
        n = _numel(d_in) if n is None else n
        _need("inclusive_f32", 4 * n, d_in, d_out)
        check(load().gps_inclusive_scan_f32(self.handle, _ptr(d_in), _ptr(d_out), n, int(compensated),
                                            _stream(stream, d_in, d_out)))

    def inclusive_bf16_f32(self, d_in, d_out, n: Optional[int] = None, compensated=True,
                           stream=None) -> None:
        """BF16 (uint16 bits) -> FP32 inclusive scan (compensated as inclusive_f32)."""
        n = _numel(d_in) if n is None else n
        _need("inclusive_bf16_f32", 2 * n, d_in)
        _need("inclusive_bf16_f32", 4 * n, d_out)
        check(load().gps_inclusive_scan_bf16_f32(self.handle, _ptr(d_in), _ptr(d_out), n, int(compensated),
                                                 _stream(stream, d_in, d_out)))

    def segmented(self, d_in, d_out, d_offsets, num_segments: int, n: int, inclusive: bool = False,
                  stream=None) -> None:
        """Batched segmented scan in one launch (d_offsets: device uint64[num_segments + 1])."""
        _need("segmented", 4 * n, d_in, d_out)
        _need("segmented", 8 * (num_segments + 1), d_offsets)
        check(load().gps_segmented_scan_u32(self.handle, _ptr(d_in), _ptr(d_out), _ptr(d_offsets), num_segments,
                                            n, int(inclusive), _stream(stream, d_in, d_out)))

    def segmented_per_segment(self, d_in, d_out, h_offsets, inclusive: bool = False, stream=None) -> None:
        """Baseline: one plain scan launch per segment (h_offsets: host uint64 sequence)."""
        import numpy as np
        off = np.ascontiguousarray(h_offsets, dtype=np.uint64)
        ptr = off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        _need("segmented_per_segment", 4 * int(off[-1]) if off.size else 0, d_in, d_out)
        check(load().gps_segmented_scan_u32_per_segment(self.handle, _ptr(d_in), _ptr(d_out), ptr, off.size - 1,
                                                        int(inclusive), _stream(stream, d_in, d_out)))

    def segmented_fp(self, d_in, d_out, d_offsets, num_segments: int, n: int, bf16: bool = False,
                     compensated: bool = True, stream=None) -> None:
        """Batched segmented FP inclusive scan in one launch (FP32 or BF16-bits
        input, FP32 output; d_offsets: device uint64[num_segments + 1])."""
        fn = load().gps_segmented_scan_bf16_f32 if bf16 else load().gps_segmented_scan_f32
        _need("segmented_fp", (2 if bf16 else 4) * n, d_in)
        _need("segmented_fp", 4 * n, d_out)
        _need("segmented_fp", 8 * (num_segments + 1), d_offsets)
        check(fn(self.handle, _ptr(d_in), _ptr(d_out), _ptr(d_offsets), num_segments, n, int(compensated),
                 _stream(stream, d_in, d_out)))

    def fp_per_segment(self, d_in, d_out, h_offsets, bf16: bool = False, compensated: bool = True,
                       stream=None) -> None:
        """One FP inclusive scan launch per segment (h_offsets: host uint64 sequence)."""
        import numpy as np
        off = np.ascontiguousarray(h_offsets, dtype=np.uint64)
        ptr = off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        total = int(off[-1]) if off.size else 0
        _need("fp_per_segment", (2 if bf16 else 4) * total, d_in)
        _need("fp_per_segment", 4 * total, d_out)
        check(load().gps_inclusive_scan_fp_per_segment(self.handle, _ptr(d_in), _ptr(d_out), ptr, off.size - 1,
                                                       int(bf16), int(compensated), _stream(stream, d_in, d_out)))

    def segmented_graph(self, d_in, d_out, d_offsets, num_segments: int, n: int,
                        inclusive: bool = False) -> "Graph":
        g = ctypes.c_void_p()
        _need("segmented_graph", 4 * n, d_in, d_out)
        _need("segmented_graph", 8 * (num_segments + 1), d_offsets)
        check(load().gps_segmented_graph_create(self.handle, _ptr(d_in), _ptr(d_out), _ptr(d_offsets),
                                                num_segments, n, int(inclusive), ctypes.byref(g)))
        return Graph(g)

    def counters(self) -> dict:
        c = (ctypes.c_uint64 * 4)()
        check(load().gps_get_counters(self.handle, c))
        return {"polls": c[0], "fallbacks": c[1], "insertions": c[2], "stall_timeouts": c[3]}

    def reset_counters(self) -> None:
        check(load().gps_reset_counters(self.handle))

    def device_error(self) -> int:
        bits = ctypes.c_uint32()
        check(load().gps_device_error(self.handle, ctypes.byref(bits)))
        return bits.value

    def reset(self, stream=None) -> None:
        check(load().gps_reset(self.handle, _stream(stream)))


class Graph:
    """An instantiated hipGraph (from ScanHandle.segmented_graph)."""

    def __init__(self, g: ctypes.c_void_p):
        self._g = g

    def launch(self, stream=None) -> None:
        check(load().gps_graph_launch(self._g, _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_g", None) is not None:
            check(load().gps_graph_destroy(self._g))
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def status_memory() -> tuple:
    """(held, free): bytes of uncached status memory the library holds on all
    devices, and how many of them sit on its free lists for reuse."""
    held, free = ctypes.c_uint64(0), ctypes.c_uint64(0)
    check(load().gps_status_memory(ctypes.byref(held), ctypes.byref(free)))
    return held.value, free.value


# ---- utility kernels -----------------------------------------------------

def fill_u32(d_ptr, n: int, value: int, stream=None) -> None:
    check(load().gps_fill_u32(_ptr(d_ptr), n, value & 0xFFFFFFFF, _stream(stream)))


def copy_u32(d_dst, d_src, n: int, stream=None) -> None:
    check(load().gps_copy_u32(_ptr(d_dst), _ptr(d_src), n, _stream(stream)))


def fill_random_u32(d_ptr, n: int, seed: int, bits: int = 32, first: int = 0, stream=None) -> None:
    check(load().gps_fill_random_u32(_ptr(d_ptr), n, seed, first, bits, _stream(stream)))


def validate_scan_u32(d_in, d_scan, n: int, inclusive: bool = True, stream=None) -> int:
    e = ctypes.c_uint64()
    check(load().gps_validate_scan_u32(_ptr(d_in), _ptr(d_scan), n, int(inclusive), _stream(stream),
                                       ctypes.byref(e)))
    return e.value


def validate_ones_u32(d_scan, n: int, inclusive: bool = True, stream=None) -> int:
    e = ctypes.c_uint64()
    check(load().gps_validate_ones_u32(_ptr(d_scan), n, int(inclusive), _stream(stream), ctypes.byref(e)))
    return e.value


WAVE_INCLUSIVE_SCAN, WAVE_SUM, WAVE_EXCLUSIVE_SCAN = 0, 1, 2

# FP scan carries (gps.h GPS_FP_*).
FP_PLAIN, FP_PAIR = 0, 1


def wave_primitive_u32(d_in, d_out, n: int, op: int, stream=None) -> None:
    """Runs one wave64 primitive (Utils.cuh:22-71 counterparts) on every run
    of 64 consecutive elements: WAVE_INCLUSIVE_SCAN, WAVE_SUM, WAVE_EXCLUSIVE_SCAN."""
    check(load().gps_wave_primitive_u32(_ptr(d_in), _ptr(d_out), n, op, _stream(stream)))


# ---- survey of wave and block scan networks, Survey.hlsl:49-855 -------------

SURVEY_NETWORKS = {
    "serial": 0, "wave_kogge_stone_lds": 1, "wave_kogge_stone_shuffle": 2, "wave_dpp": 3, "wave_raking": 4,
    "block_kogge_stone": 5, "block_sklansky": 6, "block_brent_kung": 7, "block_wave_raking": 8,
    "block_wave_sklansky": 9, "tile_registers": 10,
}


def survey_segment(alg: int) -> int:
    """Segment length (elements) of a survey network."""
    s = ctypes.c_uint32()
    check(load().gps_survey_segment_u32(alg, ctypes.byref(s)))
    return s.value


def survey_scan_u32(alg: int, d_in, d_out, n: int, inclusive: bool = True, stream=None) -> None:
    """Scans each segment of n elements independently with one survey network."""
    check(load().gps_survey_scan_u32(alg, _ptr(d_in), _ptr(d_out), n, int(inclusive), _stream(stream)))


# ---- reference dispatcher classes (self-tests run in the C++ classes) -----

class _Dispatcher:
    _algo = ALGO_CSDL

    def __init__(self, max_size: int):
        self.max_size = max_size

    def _test_all(self, inclusive: bool) -> tuple:
        passed, total = ctypes.c_uint32(), ctypes.c_uint32()
        check(load().gps_dispatcher_test_all(self._algo, self.max_size, int(inclusive), ctypes.byref(passed),
                                             ctypes.byref(total)))
        return passed.value, total.value

    def test_all_inclusive(self) -> tuple:
        return self._test_all(True)

    def test_all_exclusive(self) -> tuple:
        return self._test_all(False)


class ChainedScanDecoupledLookbackDispatcher(_Dispatcher):
    _algo = ALGO_CSDL


class ReduceThenScanDispatcher(_Dispatcher):
    _algo = ALGO_RTS


class EmulatedDeadlockingDispatcher(_Dispatcher):
    _algo = ALGO_CSDLDF
