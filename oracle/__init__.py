"""CPU oracle for the prefix sums: numpy-facing wrappers over oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the reported CPU
baseline, never as the thing measured or shipped. The product library
(gpuprefixsums_amd/libgps.so) does not link or call it and has no CPU fallback.

What each function restates (reference files relative to the reference root):
  scan_u32 / validate_u32   the serial uint32 validator loop,
                            GPUPrefixSumsD3D12/Shaders/Utility.hlsl:99-125
                            (copy: GPUPrefixSumsUnity/Shaders/Utility.compute:131-187)
  count_ones_errors         the all-ones closed forms,
                            GPUPrefixSumsCUDA/UtilityKernels.cuh:22-38
  segmented_scan_u32        per-segment serial loop (no reference counterpart)
  scan_f32_f64 / bf16       FP64 Neumaier-compensated running sum (no counterpart)
  segmented_scan_fp_f64     the FP64 scans restarted at every segment offset
  generators                counter-based splitmix64 streams (no counterpart); the
                            reference's D3D12 and Unity test generators,
                            GPUPrefixSumsD3D12/Shaders/Utility.hlsl:37-66 and
                            GPUPrefixSumsUnity/Shaders/Utility.compute:50-72

How each is pinned (tests/test_oracle.py): closed forms, brute force with
numpy uint64 cumsum, invariants (excl[0] = 0, excl[i+1] = incl[i], last =
sum mod 2^32, linearity mod 2^32), splitmix64 published test vectors,
bf16 rounding against exhaustive float32 cases, math.fsum for FP64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64 = ctypes.c_uint64


def build() -> Path:
    """Compiles liboracle.so with the committed Makefile (gcc)."""
    res = subprocess.run(["make", "-s", "-C", str(_HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        build()
    L = ctypes.CDLL(str(_LIB_PATH))
    L.orc_scan_u32.argtypes = [_u32p, _u32p, _u64, ctypes.c_int]
    L.orc_scan_u32.restype = None
    L.orc_validate_u32.argtypes = [_u32p, _u32p, _u64, ctypes.c_int]
    L.orc_validate_u32.restype = _u64
    L.orc_count_ones_errors.argtypes = [_u32p, _u64, ctypes.c_int]
    L.orc_count_ones_errors.restype = _u64
    L.orc_scan_u32_mt.argtypes = [_u32p, _u32p, _u64, ctypes.c_int, ctypes.c_int]
    L.orc_scan_u32_mt.restype = None
    L.orc_segmented_scan_u32.argtypes = [_u32p, _u32p, _u64p, _u64, ctypes.c_int]
    L.orc_segmented_scan_u32.restype = ctypes.c_int
    L.orc_scan_f32_f64.argtypes = [_f32p, _f64p, _u64]
    L.orc_scan_f32_f64.restype = None
    L.orc_scan_bf16_f64.argtypes = [_u16p, _f64p, _u64]
    L.orc_scan_bf16_f64.restype = None
    L.orc_abs_scan_f32_f64.argtypes = [_f32p, _f64p, _u64]
    L.orc_abs_scan_f32_f64.restype = None
    L.orc_abs_scan_bf16_f64.argtypes = [_u16p, _f64p, _u64]
    L.orc_abs_scan_bf16_f64.restype = None
    L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
    L.orc_bf16_to_f32.restype = ctypes.c_float
    L.orc_f32_to_bf16_rne.argtypes = [ctypes.c_float]
    L.orc_f32_to_bf16_rne.restype = ctypes.c_uint16
    L.orc_splitmix64_at.argtypes = [_u64, _u64]
    L.orc_splitmix64_at.restype = _u64
    L.orc_gen_u32_bits.argtypes = [_u64, _u64, _u32p, _u64, ctypes.c_int]
    L.orc_gen_u32_bits.restype = None
    L.orc_normal_at.argtypes = [_u64, _u64]
    L.orc_normal_at.restype = ctypes.c_double
    L.orc_gen_f32_normal.argtypes = [_u64, _f32p, _u64]
    L.orc_gen_f32_normal.restype = None
    L.orc_gen_bf16_logmag.argtypes = [_u64, _u16p, _u64]
    L.orc_gen_bf16_logmag.restype = None
    L.orc_gen_lognormal_offsets.argtypes = [_u64, _u64, ctypes.c_double, ctypes.c_double, _u64, _u64, _u64p]
    L.orc_gen_lognormal_offsets.restype = _u64
    L.orc_gen_d3d12_taus.argtypes = [ctypes.c_uint32, _u32p, _u64]
    L.orc_gen_d3d12_taus.restype = None
    L.orc_gen_unity_taus.argtypes = [ctypes.c_uint32, _u32p, _u64]
    L.orc_gen_unity_taus.restype = None
    _lib = L
    return L


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- uint32 scans --------------------------------------------------------

def scan_u32(x, inclusive: bool = True) -> np.ndarray:
    x = _c(x, np.uint32)
    out = np.empty_like(x)
    lib().orc_scan_u32(x, out, x.size, int(inclusive))
    return out


def validate_u32(x, scan, inclusive: bool = True) -> int:
    x, scan = _c(x, np.uint32), _c(scan, np.uint32)
    assert x.size == scan.size
    return int(lib().orc_validate_u32(x, scan, x.size, int(inclusive)))


def count_ones_errors(scan, inclusive: bool = True) -> int:
    scan = _c(scan, np.uint32)
    return int(lib().orc_count_ones_errors(scan, scan.size, int(inclusive)))


def scan_u32_mt(x, inclusive: bool = True, threads: int = 0) -> np.ndarray:
    x = _c(x, np.uint32)
    out = np.empty_like(x)
    lib().orc_scan_u32_mt(x, out, x.size, int(inclusive), int(threads))
    return out


def scan_u32x4(x, inclusive: bool = True) -> np.ndarray:
    """Member-wise scan of 4-member uint32 structs: x is (n, 4) or 4n
    interleaved uint32; each member column is scanned on its own, mod 2^32
    (GPUPrefixSumsWebGPUapis/SharedShaders/csdldf_struct.wgsl:153-291 keeps
    one reduction per member per tile; validate.wgsl:51-60 checks all-ones
    input against (i >> 2) + 1 per uint32)."""
    x = _c(x, np.uint32).reshape(-1, 4)
    out = np.empty_like(x)
    for m in range(4):
        out[:, m] = scan_u32(x[:, m], inclusive)
    return out


def segmented_scan_u32(x, offsets, inclusive: bool = False) -> np.ndarray:
    x, offsets = _c(x, np.uint32), _c(offsets, np.uint64)
    assert offsets.size >= 1 and int(offsets[-1]) == x.size
    out = np.empty_like(x)
    rc = lib().orc_segmented_scan_u32(x, out, offsets, offsets.size - 1, int(inclusive))
    if rc != 0:
        raise ValueError("bad segment offsets")
    return out


# ---- floating point ------------------------------------------------------

def scan_f32_f64(x) -> np.ndarray:
    x = _c(x, np.float32)
    out = np.empty(x.size, np.float64)
    lib().orc_scan_f32_f64(x, out, x.size)
    return out


def scan_bf16_f64(bits) -> np.ndarray:
    bits = _c(bits, np.uint16)
    out = np.empty(bits.size, np.float64)
    lib().orc_scan_bf16_f64(bits, out, bits.size)
    return out


def abs_scan_f32_f64(x) -> np.ndarray:
    x = _c(x, np.float32)
    out = np.empty(x.size, np.float64)
    lib().orc_abs_scan_f32_f64(x, out, x.size)
    return out


def abs_scan_bf16_f64(bits) -> np.ndarray:
    bits = _c(bits, np.uint16)
    out = np.empty(bits.size, np.float64)
    lib().orc_abs_scan_bf16_f64(bits, out, bits.size)
    return out


def segmented_scan_fp_f64(x, offsets, bf16: bool = False, absolute: bool = False) -> np.ndarray:
    """Per-segment FP64 inclusive scan (or scan of |x|) of FP32 values or BF16
    bits: the scans above restarted at every offset (no reference counterpart)."""
    offsets = _c(offsets, np.uint64)
    assert offsets.size >= 1 and int(offsets[0]) == 0 and int(offsets[-1]) == np.asarray(x).size
    f = (abs_scan_bf16_f64 if absolute else scan_bf16_f64) if bf16 else \
        (abs_scan_f32_f64 if absolute else scan_f32_f64)
    out = np.empty(np.asarray(x).size, np.float64)
    for s in range(offsets.size - 1):
        a, b = int(offsets[s]), int(offsets[s + 1])
        if b < a:
            raise ValueError("bad segment offsets")
        if b > a:
            out[a:b] = f(x[a:b])
    return out


def bf16_to_f32(b: int) -> float:
    return float(lib().orc_bf16_to_f32(int(b)))


def f32_to_bf16(f: float) -> int:
    return int(lib().orc_f32_to_bf16_rne(float(f)))


def bf16_bits_to_f32(bits) -> np.ndarray:
    """Vector form of bf16_to_f32 (bit shift; same definition)."""
    return (_c(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


# ---- generators ----------------------------------------------------------

def splitmix64_at(seed: int, i: int) -> int:
    return int(lib().orc_splitmix64_at(seed, i))


def gen_u32(seed: int, n: int, bits: int = 32, first: int = 0) -> np.ndarray:
    out = np.empty(n, np.uint32)
    if n:
        lib().orc_gen_u32_bits(seed, first, out, n, bits)
    return out


def normal_at(seed: int, i: int) -> float:
    return float(lib().orc_normal_at(seed, i))


def gen_f32_normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    if n:
        lib().orc_gen_f32_normal(seed, out, n)
    return out


def gen_bf16_logmag(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint16)
    if n:
        lib().orc_gen_bf16_logmag(seed, out, n)
    return out


def gen_lognormal_offsets(seed: int, num_segments: int, median: float = 4096.0, sigma: float = 1.5,
                          lo: int = 1, hi: int = 1 << 22) -> np.ndarray:
    off = np.empty(num_segments + 1, np.uint64)
    lib().orc_gen_lognormal_offsets(seed, num_segments, median, sigma, lo, hi, off)
    return off


def gen_unity_taus(seed: int, n: int) -> np.ndarray:
    """Inputs of the reference's Unity random tests (bytes of a per-thread
    hybrid Tausworthe stream, values in [0, 255]); the reference's own tests
    run it with seed 0 (GPUPrefixSumsUnity/Shaders/Utility.compute:50-72)."""
    out = np.empty(n, dtype=np.uint32)
    if n:
        lib().orc_gen_unity_taus(seed & 0xFFFFFFFF, out, n)
    return out


def gen_d3d12_taus(seed: int, n: int) -> np.ndarray:
    """Inputs of the reference's D3D12 random tests (values in [0, 1023])."""
    out = np.empty(n, np.uint32)
    if n:
        lib().orc_gen_d3d12_taus(seed & 0xFFFFFFFF, out, n)
    return out


def cpu_threads() -> int:
    return os.cpu_count() or 1
