"""gps_scan_u32_noalloc (the allocation-free scan over the library's static
device scratch) against the handle-based gps_scan_u32 and the oracle,
bit-exact: two calls in a row, the second smaller, inclusive and exclusive;
the largest size it accepts; and the size limit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from gpuprefixsums_amd import _lib  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

LIMIT = 1 << 28


@pytest.fixture(scope="module")
def noalloc():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    return _lib.load().gps_scan_u32_noalloc


def stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("sizes", [((1 << 26) + 5, (1 << 20) + 3), ((1 << 24) + 77, 4099), (3000, 1)])
def test_matches_the_handle_scan_on_two_calls(noalloc, sizes, inclusive):
    with gps.ScanHandle(gps.ALGO_CSDL, max(sizes)) as h:
        for i, n in enumerate(sizes):  # the second call is smaller than the first
            x = oracle.gen_u32(100 + i + 7 * n, n, 32)
            d_in = to_dev(x)
            a = sentinel_u32(n + 64)
            b = sentinel_u32(n + 64)
            assert noalloc(d_in.data_ptr(), a.data_ptr(), n, int(inclusive), stream()) == 0, _lib.load().gps_last_error()
            h.scan(d_in, b, n, inclusive)
            sync()
            got, ref = from_dev_u32(a), from_dev_u32(b)
            assert np.all(got[n:] == 0xDEADBEEF), "wrote past n"
            assert np.array_equal(got, ref), f"call {i} n={n}: {first_mismatch(got, ref)}"
            want = oracle.scan_u32(x, inclusive)
            assert np.array_equal(got[:n], want), f"call {i} n={n}: {first_mismatch(got[:n], want)}"


def test_largest_size_in_place(noalloc):
    n = LIMIT
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    gps.fill_random_u32(x, n, seed=5, bits=32)
    y = x.clone()
    assert noalloc(y.data_ptr(), y.data_ptr(), n, 1, stream()) == 0, _lib.load().gps_last_error()
    sync()
    assert gps.validate_scan_u32(x, y, n, True) == 0
    del x, y
    torch.cuda.empty_cache()


def test_size_limit_and_arguments(noalloc):
    L = _lib.load()
    assert noalloc(16, 16, LIMIT + 1, 1, None) == _lib.GPS_ERR_TOO_LARGE
    assert b"limit" in L.gps_last_error()
    assert noalloc(None, None, 5, 1, None) == _lib.GPS_ERR_INVALID
    assert noalloc(None, None, 0, 1, None) == 0
    d = torch.zeros(8, dtype=torch.int32, device="cuda")
    assert noalloc(d.data_ptr() + 2, d.data_ptr(), 4, 1, None) == _lib.GPS_ERR_INVALID
