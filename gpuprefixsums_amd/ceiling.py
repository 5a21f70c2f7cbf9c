"""Memory-ceiling probes (ctypes binding of libgps_ceiling.so, the C entry
points in include/gps_ceiling.h).

Benchmark only: copies and one-direction streams with the access patterns a
single-pass scan can take, and the runtime's own copy and fill, timed beside
the scan so its roofline fraction can be read against what HBM delivers to
each pattern. The library links libgps.so and reports errors through
gps_last_error().
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import PKG, GpsError

CEILING_LIB_PATH = PKG / "libgps_ceiling.so"

(COPY_64X1, COPY_256X4, COPY_1024X8, TICKET_1024X4, TICKET_1024X8, COPY_LDS_256X4, READ_1024X1, WRITE_256X1,
 MEMCPY, MEMSET, TICKET_1024X4_STATUS, TICKET_1024X8_STATUS, TICKET_1024X4_STATUS_UC,
 TICKET_1024X8_STATUS_UC) = range(14)
KINDS = 14
COPY_KINDS = (COPY_64X1, COPY_256X4, COPY_1024X8, TICKET_1024X4, TICKET_1024X8, COPY_LDS_256X4, MEMCPY,
              TICKET_1024X4_STATUS, TICKET_1024X8_STATUS, TICKET_1024X4_STATUS_UC, TICKET_1024X8_STATUS_UC)
# kinds that use the caller's scratch (ticket word, and status slots for the status kinds)
TICKETED = (TICKET_1024X4, TICKET_1024X8, TICKET_1024X4_STATUS, TICKET_1024X8_STATUS)

_vp, _u64, _int = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int

# name -> (restype, argtypes) for every function in include/gps_ceiling.h.
SIGNATURES = {
    "gps_ceiling_granule": (_int, [_int, ctypes.POINTER(_u64)]),
    "gps_ceiling_name": (ctypes.c_char_p, [_int]),
    "gps_ceiling_scratch_bytes": (_int, [_int, _u64, ctypes.POINTER(_u64)]),
    "gps_ceiling_run": (_int, [_int, _vp, _vp, _u64, _vp, _vp]),
}

_clib = None


def load() -> ctypes.CDLL:
    """Loads libgps_ceiling.so (after libgps.so, which it links); raises if it
    has not been built."""
    global _clib
    if _clib is not None:
        return _clib
    _lib.load()
    if not CEILING_LIB_PATH.exists():
        raise ImportError(f"{CEILING_LIB_PATH} is not built; run `python -m gpuprefixsums_amd.build`")
    L = ctypes.CDLL(str(CEILING_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _clib = L
    return L


def _check(rc: int) -> None:
    if rc != _lib.GPS_OK:
        raise GpsError(rc, _lib.load().gps_last_error().decode())


def _ptr(x) -> int:
    from . import _ptr as p
    return p(x)


def name(kind: int) -> str:
    return load().gps_ceiling_name(kind).decode()


def granule(kind: int) -> int:
    """Bytes per tile of `kind`; a run moves a whole number of them."""
    g = _u64()
    _check(load().gps_ceiling_granule(kind, ctypes.byref(g)))
    return g.value


def scratch_bytes(kind: int, nbytes: int) -> int:
    """Device scratch the probe needs for `nbytes` (0 for the one-shot kinds)."""
    b = _u64()
    _check(load().gps_ceiling_scratch_bytes(kind, nbytes, ctypes.byref(b)))
    return b.value


def run(kind: int, src, dst, nbytes: int, scratch=None, stream=None) -> None:
    """One launch of probe `kind` over `nbytes` (src -> dst for the copies;
    src only for the read stream, dst only for the write stream and fill).
    `scratch` (zeroed, 256-byte aligned, scratch_bytes(kind, nbytes) long) is
    needed by the ticketed copies; each launch leaves its ticket word zero.
    Launches on torch's current stream when `stream` is None."""
    from . import _stream
    _check(load().gps_ceiling_run(kind, _ptr(src), _ptr(dst), nbytes, _ptr(scratch), _stream(stream, src, dst)))
