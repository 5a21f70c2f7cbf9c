"""ctypes binding of libgps.so (the C ABI in include/gps.h).

Loading fails loudly when the library has not been built: there is no CPU or
PyTorch fallback for any entry point.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libgps.so"

GPS_OK = 0
GPS_ERR_INVALID = 1
GPS_ERR_TOO_LARGE = 2
GPS_ERR_HIP = 3
GPS_ERR_UNSUPPORTED = 4
GPS_ERR_DEVICE = 5

ALGO_CSDL = 0
ALGO_RTS = 1
ALGO_CSDLDF = 2

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_fp = ctypes.POINTER(ctypes.c_float)


class GpsTuning(ctypes.Structure):
    _fields_ = [
        ("block_threads", _u32),
        ("vec_per_thread", _u32),
        ("spin_limit", _u32),
        ("stall_period", _u32),
        ("schedule", _u32),
    ]


SCHEDULE_DEFAULT = 0
SCHEDULE_TILE = 1
SCHEDULE_PERSISTENT = 2
SCHEDULE_CONTROL = 3
SCHEDULE_SPLIT = 4


class GpsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gps error {code}: {msg}")
        self.code = code


# name -> (restype, argtypes). Every function declared in include/gps.h must
# be listed here; tests/test_abi.py checks both directions. (The vendor
# comparator's entry points, include/gps_vendor.h, are bound in vendor.py.)
SIGNATURES = {
    "gps_version": (ctypes.c_char_p, []),
    "gps_last_error": (ctypes.c_char_p, []),
    "gps_create": (_int, [_int, _u64, ctypes.POINTER(GpsTuning), ctypes.POINTER(_vp)]),
    "gps_destroy": (_int, [_vp]),
    "gps_status_memory": (_int, [ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "gps_get_tuning": (_int, [_vp, ctypes.POINTER(GpsTuning), ctypes.POINTER(_u32)]),
    "gps_inclusive_scan_u32": (_int, [_vp, _vp, _vp, _u64, _vp]),
    "gps_exclusive_scan_u32": (_int, [_vp, _vp, _vp, _u64, _vp]),
    "gps_scan_u32": (_int, [_vp, _vp, _vp, _u64, _int, _vp]),
    "gps_scan_u32_carry": (_int, [_vp, _vp, _vp, _u64, _int, _vp, _vp, _vp]),
    "gps_segmented_scan_u32": (_int, [_vp, _vp, _vp, _vp, _u64, _u64, _int, _vp]),
    "gps_segmented_scan_u32_per_segment": (_int, [_vp, _vp, _vp, ctypes.POINTER(_u64), _u64, _int, _vp]),
    "gps_segmented_graph_create": (_int, [_vp, _vp, _vp, _vp, _u64, _u64, _int, ctypes.POINTER(_vp)]),
    "gps_inclusive_scan_f32": (_int, [_vp, _vp, _vp, _u64, _int, _vp]),
    "gps_inclusive_scan_bf16_f32": (_int, [_vp, _vp, _vp, _u64, _int, _vp]),
    "gps_inclusive_scan_fp_per_segment": (_int, [_vp, _vp, _vp, ctypes.POINTER(_u64), _u64, _int, _int, _vp]),
    "gps_segmented_scan_f32": (_int, [_vp, _vp, _vp, _vp, _u64, _u64, _int, _vp]),
    "gps_inclusive_scan_u32x4": (_int, [_vp, _vp, _vp, _u64, _vp]),
    "gps_exclusive_scan_u32x4": (_int, [_vp, _vp, _vp, _u64, _vp]),
    "gps_survey_segment_u32": (_int, [_int, ctypes.POINTER(_u32)]),
    "gps_survey_scan_u32": (_int, [_int, _vp, _vp, _u64, _int, _vp]),
    "gps_estimate_occupancy": (_int, [_vp, _vp, _vp, _u64, _int, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "gps_segmented_scan_bf16_f32": (_int, [_vp, _vp, _vp, _vp, _u64, _u64, _int, _vp]),
    "gps_mlp_param_count": (_u64, [_u32, _u32]),
    "gps_mlp_create": (_int, [_u32, _u32, _u32, ctypes.POINTER(_vp)]),
    "gps_mlp_destroy": (_int, [_vp]),
    "gps_mlp_set_params": (_int, [_vp, _fp]),
    "gps_mlp_get_params": (_int, [_vp, _fp]),
    "gps_mlp_forward": (_int, [_vp, _fp, _u32, _fp]),
    "gps_mlp_loss_grad": (_int, [_vp, _fp, _fp, _u32, _fp, _fp]),
    "gps_mlp_train": (_int, [_vp, _fp, _fp, _u32, _u32, ctypes.c_float, ctypes.c_float, _fp]),
    "gps_graph_launch": (_int, [_vp, _vp]),
    "gps_graph_destroy": (_int, [_vp]),
    "gps_get_counters": (_int, [_vp, ctypes.POINTER(_u64)]),
    "gps_reset_counters": (_int, [_vp]),
    "gps_device_error": (_int, [_vp, ctypes.POINTER(_u32)]),
    "gps_reset": (_int, [_vp, _vp]),
    "gps_time_scan_u32": (_int, [_vp, _vp, _vp, _u64, _int, _vp, _u32, ctypes.POINTER(ctypes.c_float)]),
    "gps_scan_u32_noalloc": (_int, [_vp, _vp, _u64, _int, _vp]),
    "gps_fill_u32": (_int, [_vp, _u64, _u32, _vp]),
    "gps_copy_u32": (_int, [_vp, _vp, _u64, _vp]),
    "gps_fill_random_u32": (_int, [_vp, _u64, _u64, _u64, _u32, _vp]),
    "gps_validate_ones_u32": (_int, [_vp, _u64, _int, _vp, ctypes.POINTER(_u64)]),
    "gps_validate_scan_u32": (_int, [_vp, _vp, _u64, _int, _vp, ctypes.POINTER(_u64)]),
    "gps_dispatcher_test_all": (_int, [_int, _u32, _int, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "gps_wave_primitive_u32": (_int, [_vp, _vp, _u64, _int, _vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads libgps.so; raises if it is missing (build with
    `python -m gpuprefixsums_amd.build` or __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m gpuprefixsums_amd.build`")
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue  # reported by the ABI test; callers get AttributeError
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != GPS_OK:
        raise GpsError(rc, load().gps_last_error().decode())
