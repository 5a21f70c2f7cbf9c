"""Occupancy estimate of the persistent scans (the reference's occupancy test
variants, GPUPrefixSumsWebGPUapis/SharedShaders/TestVariants/csdldf_occ.wgsl:70-82
and csdldf_struct_occ.wgsl): the scan runs with one workgroup launched per
tile, and the workgroups that took a tile are counted. The output must be the
ordinary scan result, and the count must be a plausible residency.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


@pytest.mark.parametrize("log2n", [20, 24, 26])
def test_u32_occupancy_and_result(log2n):
    n = 1 << log2n
    x = oracle.gen_u32(log2n, n, 32)
    d_in, d_out = to_dev(x), sentinel_u32(n)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        measured, resident = h.estimate_occupancy(d_in, d_out, n)
        tiles = -(-n // h.tile)
        print(f"n=2^{log2n} tiles={tiles} measured={measured} resident={resident}")
        assert resident > 0
        assert 1 <= measured <= tiles
        assert measured >= min(tiles, resident) // 2
        got = from_dev_u32(d_out)
        want = oracle.scan_u32(x, True)
        assert np.array_equal(got, want), first_mismatch(got, want)
        assert h.device_error() == 0
        # The handle's ordinary scans still work after the oversubscribed launch.
        h.inclusive(d_in, d_out, n)
        sync()
        assert np.array_equal(from_dev_u32(d_out), want)


def test_struct_occupancy_and_result():
    n = 1 << 24
    x = oracle.gen_u32(9, 4 * n, 32)
    d_in, d_out = to_dev(x), sentinel_u32(4 * n)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        measured, resident = h.estimate_occupancy(d_in, d_out, n, struct_scan=True)
        tiles = -(-n // 8192)
        print(f"structs n=2^24 tiles={tiles} measured={measured} resident={resident}")
        assert resident > 0
        assert 1 <= measured <= tiles
        assert measured >= min(tiles, resident) // 2
        got = from_dev_u32(d_out).reshape(-1, 4)
        want = oracle.scan_u32x4(x, True)
        assert np.array_equal(got, want), first_mismatch(got, want)
        assert h.device_error() == 0


def test_occupancy_argument_errors():
    d = torch.zeros(4096, dtype=torch.int32, device="cuda")
    with gps.ScanHandle(gps.ALGO_RTS, 4096) as h:
        with pytest.raises(gps.GpsError, match="CSDL"):
            h.estimate_occupancy(d, d, 4096)
    with gps.ScanHandle(gps.ALGO_CSDL, 4096) as h:
        with pytest.raises(gps.GpsError, match="positive"):
            h.estimate_occupancy(d, d, 0)
        big = torch.zeros(8192, dtype=torch.int32, device="cuda")
        with pytest.raises(gps.GpsError, match="max_n"):
            h.estimate_occupancy(big, big, 4097)
        with pytest.raises(ValueError, match="buffer"):  # n past the tensors: refused on the host
            h.estimate_occupancy(d, d, 4097)
        with pytest.raises(gps.GpsError, match="16-byte"):
            h.estimate_occupancy(d.data_ptr() + 4, d, 100, struct_scan=True)
