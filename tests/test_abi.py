"""The C-ABI library loads and exports everything include/*.h declares.

CPU only: no call here touches a GPU (only argument checks that fail before
any HIP call are exercised).
"""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


# The comparator library's header and the ceiling probes' header; every other
# include/*.h belongs to libgps.so.
VENDOR_HEADER = "gps_vendor.h"
CEILING_HEADER = "gps_ceiling.h"


def declared_in(h: Path) -> set:
    text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
    return {m.group(1) for m in re.finditer(r"\b(gps_[a-z0-9_]+)\s*\(", text)}


def declared_functions() -> set:
    """Entry points of the product library libgps.so."""
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        if h.name not in (VENDOR_HEADER, CEILING_HEADER):
            names |= declared_in(h)
    return names


def vendor_functions() -> set:
    return declared_in(ROOT / "include" / VENDOR_HEADER)


def ceiling_functions() -> set:
    return declared_in(ROOT / "include" / CEILING_HEADER)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("gps_create", "gps_destroy", "gps_inclusive_scan_u32", "gps_exclusive_scan_u32"):
        assert required in names


def test_library_exports_every_declared_symbol(gps_lib):
    missing = [n for n in sorted(declared_functions()) if not hasattr(gps_lib, n)]
    assert not missing, f"libgps.so lacks {missing}"


def test_python_binding_covers_header():
    from gpuprefixsums_amd import _lib, ceiling, vendor
    assert declared_functions() == set(_lib.SIGNATURES), (
        declared_functions() ^ set(_lib.SIGNATURES))
    assert vendor_functions() == set(vendor.SIGNATURES), vendor_functions() ^ set(vendor.SIGNATURES)
    assert ceiling_functions() == set(ceiling.SIGNATURES), ceiling_functions() ^ set(ceiling.SIGNATURES)


def test_comparator_library_exports_its_header(gps_lib):
    from gpuprefixsums_amd import build, vendor
    if not vendor.VENDOR_LIB_PATH.exists():
        build.build_lib()
    v = vendor.load()
    missing = [n for n in sorted(vendor_functions()) if not hasattr(v, n)]
    assert not missing, f"libgps_vendor.so lacks {missing}"


def test_ceiling_library_exports_its_header(gps_lib):
    from gpuprefixsums_amd import build, ceiling
    if not ceiling.CEILING_LIB_PATH.exists():
        build.build_lib()
    c = ceiling.load()
    missing = [n for n in sorted(ceiling_functions()) if not hasattr(c, n)]
    assert not missing, f"libgps_ceiling.so lacks {missing}"
    assert not any(hasattr(gps_lib, n) for n in ceiling_functions())


def test_ceiling_probe_metadata_and_argument_errors(gps_lib):
    """Names, granules and argument checks that fail before any HIP call."""
    from gpuprefixsums_amd import ceiling
    from gpuprefixsums_amd._lib import GPS_ERR_INVALID, GpsError
    header = (ROOT / "include" / CEILING_HEADER).read_text()
    assert f"GPS_CEIL_KINDS = {ceiling.KINDS}" in header
    names = [ceiling.name(k) for k in range(ceiling.KINDS)]
    assert len(set(names)) == ceiling.KINDS and ceiling.name(ceiling.KINDS) == "unknown"
    assert ceiling.granule(ceiling.COPY_64X1) == 1024
    assert ceiling.granule(ceiling.COPY_1024X8) == 128 * 1024
    assert ceiling.granule(ceiling.TICKET_1024X4) == 64 * 1024
    with pytest.raises(GpsError) as e:
        ceiling.granule(ceiling.KINDS)
    assert e.value.code == GPS_ERR_INVALID
    fake = 1 << 20  # never dereferenced: every call below fails its argument check first
    for kind, nbytes, src, dst in ((ceiling.COPY_1024X8, 1024, fake, fake),        # not a whole tile
                                   (ceiling.COPY_64X1, 0, fake, fake),             # empty
                                   (ceiling.COPY_64X1, 1024, 0, fake),             # null source
                                   (ceiling.WRITE_256X1, 4096, 0, 0),              # null destination
                                   (ceiling.COPY_64X1, 1024, fake + 4, fake)):     # misaligned
        with pytest.raises(GpsError) as e:
            ceiling.run(kind, src, dst, nbytes, stream=0)
        assert e.value.code == GPS_ERR_INVALID
    for kind in ceiling.TICKETED:  # no scratch, or scratch off its 256-byte alignment
        for scratch in (None, fake + 128):
            with pytest.raises(GpsError) as e:
                ceiling.run(kind, fake, fake, ceiling.granule(kind), scratch=scratch, stream=0)
            assert e.value.code == GPS_ERR_INVALID
    assert ceiling.scratch_bytes(ceiling.COPY_64X1, 1 << 20) == 0
    assert ceiling.scratch_bytes(ceiling.TICKET_1024X4, 1 << 20) == 256
    assert ceiling.scratch_bytes(ceiling.TICKET_1024X4_STATUS, 1 << 20) == 256 + 16 * 128


def test_product_library_carries_no_vendor_scan(gps_lib):
    import shutil
    import subprocess
    from gpuprefixsums_amd import _lib
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm is not installed")
    syms = subprocess.run([nm, "-D", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    assert "rocprim" not in syms.lower()
    assert not any(hasattr(gps_lib, n) for n in vendor_functions())


def test_argument_errors_before_any_device_call(gps_lib):
    from gpuprefixsums_amd._lib import GPS_ERR_INVALID, GpsTuning
    h = ctypes.c_void_p()
    assert gps_lib.gps_create(0, 1024, None, None) == GPS_ERR_INVALID
    assert gps_lib.gps_create(7, 1024, None, ctypes.byref(h)) == GPS_ERR_INVALID
    assert b"unknown algorithm" in gps_lib.gps_last_error()
    assert gps_lib.gps_create(0, 0, None, ctypes.byref(h)) == GPS_ERR_INVALID
    bad = GpsTuning(384, 4, 0, 0)
    assert gps_lib.gps_create(0, 1024, ctypes.byref(bad), ctypes.byref(h)) == GPS_ERR_INVALID
    assert b"unsupported tile shape" in gps_lib.gps_last_error()
    stall_on_csdl = GpsTuning(0, 0, 0, 8)
    assert gps_lib.gps_create(0, 1024, ctypes.byref(stall_on_csdl), ctypes.byref(h)) == GPS_ERR_INVALID
    period_one = GpsTuning(0, 0, 0, 1)
    assert gps_lib.gps_create(2, 1024, ctypes.byref(period_one), ctypes.byref(h)) == GPS_ERR_INVALID
    huge_spin = GpsTuning(0, 0, 65537, 8)
    assert gps_lib.gps_create(2, 1024, ctypes.byref(huge_spin), ctypes.byref(h)) == GPS_ERR_INVALID
    assert b"spin_limit" in gps_lib.gps_last_error()
    persistent, control = GpsTuning(0, 0, 0, 0, 2), GpsTuning(0, 0, 0, 0, 3)
    assert gps_lib.gps_create(1, 1024, ctypes.byref(persistent), ctypes.byref(h)) == GPS_ERR_INVALID  # RTS
    assert gps_lib.gps_create(2, 1024, ctypes.byref(control), ctypes.byref(h)) == GPS_ERR_INVALID  # CSDLDF
    assert gps_lib.gps_inclusive_scan_u32(None, None, None, 0, None) == GPS_ERR_INVALID
    assert gps_lib.gps_destroy(None) == GPS_ERR_INVALID
    assert gps_lib.gps_version().startswith(b"gpuprefixsums_amd")


def test_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    from gpuprefixsums_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(ImportError):
        _lib.load()


def test_survey_segment_lengths():
    # Host-only query: no device work.
    import gpuprefixsums_amd as gps
    want = {"serial": 64, "wave_kogge_stone_lds": 64, "wave_kogge_stone_shuffle": 64, "wave_dpp": 64,
            "wave_raking": 1024, "block_kogge_stone": 256, "block_sklansky": 256, "block_brent_kung": 256,
            "block_wave_raking": 256, "block_wave_sklansky": 256, "tile_registers": 4096}
    assert {k: gps.survey_segment(v) for k, v in gps.SURVEY_NETWORKS.items()} == want
    with pytest.raises(gps.GpsError):
        gps.survey_segment(len(want))


def test_sized_buffers_are_checked_against_n():
    # An explicit n is checked against tensors' sizes before any call.
    torch = pytest.importorskip("torch")
    import gpuprefixsums_amd as gps
    t = torch.zeros(100, dtype=torch.int32)
    gps._need("scan", 400, t)  # exactly enough
    with pytest.raises(ValueError, match="400 bytes"):
        gps._need("scan", 404, t)
    gps._need("scan", 1 << 40, 12345)  # raw addresses are not sized
