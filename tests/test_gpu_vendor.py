"""The vendor comparator (rocPRIM's device scan, standing in for the
reference's CUB comparator, GPUPrefixSumsCUDA/CubDispatcher.cuh:32-76),
checked bit-exact against the CPU oracle. The reference never validates its
comparator; here its numbers are only reported once its results match.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from gpuprefixsums_amd import vendor  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def run_vendor(x: np.ndarray, inclusive: bool, in_place: bool = False) -> np.ndarray:
    n = x.size
    temp_bytes = vendor.vendor_scan_temp_bytes(n, inclusive)
    temp = torch.empty(max(temp_bytes, 4), dtype=torch.uint8, device="cuda")
    d_in = to_dev(x, pad=64)
    d_out = d_in if in_place else sentinel_u32(n + 64)
    vendor.vendor_scan_u32(d_in, d_out, n, temp, temp_bytes, inclusive)
    sync()
    out = from_dev_u32(d_out)
    assert np.all(out[n:n + 64] == 0xDEADBEEF), "wrote past n"
    return out[:n]


@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("n", [1, 3, 4097, 1 << 20, (1 << 22) + 5])
def test_vendor_full_range_matches_oracle(n, inclusive):
    x = oracle.gen_u32(seed=2, n=n, bits=32)
    got, want = run_vendor(x, inclusive), oracle.scan_u32(x, inclusive)
    assert np.array_equal(got, want), first_mismatch(got, want)


@pytest.mark.parametrize("inclusive", [True, False])
def test_vendor_c1_in_place(inclusive):
    # n = 2^20, uniform [0, 255], seed 1, scanned in place as the reference does
    x = oracle.gen_u32(seed=1, n=1 << 20, bits=8)
    got, want = run_vendor(x, inclusive, in_place=True), oracle.scan_u32(x, inclusive)
    assert np.array_equal(got, want), first_mismatch(got, want)


def test_vendor_temp_too_small_is_rejected():
    n = 1 << 16
    need = vendor.vendor_scan_temp_bytes(n, True)
    if need == 0:
        pytest.skip("rocPRIM needs no temporary storage at this size")
    d = to_dev(np.ones(n, np.uint32))
    temp = torch.empty(need, dtype=torch.uint8, device="cuda")
    with pytest.raises(gps.GpsError):
        vendor.vendor_scan_u32(d, d, n, temp, need - 1, True)


def test_cub_dispatcher_batch_timing():
    keys = vendor.CubDispatcher(1 << 20).batch_timing_cub_chained_scan(1 << 20, 4)
    assert keys > 0
    # oversize request: printed and 0, like the reference
    assert vendor.CubDispatcher(1 << 10).batch_timing_cub_chained_scan(1 << 11, 1) == 0.0
