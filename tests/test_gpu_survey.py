"""Survey of wave and block scan networks (Survey.hlsl:49-855 counterparts),
bit-exact against the oracle's segmented scan with equal-length segments.

Each network scans independent segments of its own length; sizes cover one
segment, a workgroup left partly empty (wave networks pack four segments per
workgroup, the serial network 256), several workgroups, and 2^20 elements.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

NETWORKS = gps.SURVEY_NETWORKS


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def want(x, seg, inclusive):
    off = np.arange(0, x.size + 1, seg, dtype=np.uint64)
    return oracle.segmented_scan_u32(x, off, inclusive)


@pytest.mark.parametrize("inclusive", [True, False], ids=["incl", "excl"])
@pytest.mark.parametrize("name", list(NETWORKS))
def test_network_matches_oracle(name, inclusive):
    alg = NETWORKS[name]
    seg = gps.survey_segment(alg)
    for k in (1, 3, 7, 257, (1 << 20) // seg):
        n = k * seg
        x = oracle.gen_u32(alg * 1000 + k, n, 32)
        d_in = to_dev(x)
        d_out = sentinel_u32(n + 64)
        gps.survey_scan_u32(alg, d_in, d_out, n, inclusive)
        sync()
        out = from_dev_u32(d_out)
        assert np.all(out[n:] == 0xDEADBEEF), f"{name}: wrote past n"
        w = want(x, seg, inclusive)
        assert np.array_equal(out[:n], w), f"{name} n={n}: {first_mismatch(out[:n], w)}"


def test_ones_closed_form_every_network():
    # all ones: inclusive i % seg + 1, exclusive i % seg
    for name, alg in NETWORKS.items():
        seg = gps.survey_segment(alg)
        n = 16 * seg
        d_in = torch.ones(n, dtype=torch.int32, device="cuda")
        d_out = torch.empty_like(d_in)
        for inclusive in (True, False):
            gps.survey_scan_u32(alg, d_in, d_out, n, inclusive)
            sync()
            got = from_dev_u32(d_out)
            expect = (np.arange(n, dtype=np.uint32) % seg) + (1 if inclusive else 0)
            assert np.array_equal(got, expect), f"{name} incl={inclusive}: {first_mismatch(got, expect)}"


def test_survey_argument_errors():
    d = torch.zeros(8192, dtype=torch.int32, device="cuda")
    with pytest.raises(gps.GpsError, match="multiple"):
        gps.survey_scan_u32(NETWORKS["block_sklansky"], d, d, 100)
    with pytest.raises(gps.GpsError, match="unknown"):
        gps.survey_scan_u32(99, d, d, 256)
    with pytest.raises(gps.GpsError, match="16-byte"):
        gps.survey_scan_u32(NETWORKS["tile_registers"], d.data_ptr() + 4, d, 4096)
    gps.survey_scan_u32(NETWORKS["wave_dpp"], d, d, 0)  # n = 0 is a no-op
