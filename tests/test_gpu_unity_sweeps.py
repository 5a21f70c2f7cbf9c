"""The reference's Unity test shapes (GPUPrefixSumsUnity/Runtime/Tests.cs:428-435),
bit-exact on the GPU:

  all ones over n in [2^25, 2^25 + 3072) in steps of 4 (one partition of
      the Unity kernels, k_partitionSize = 3072, GPUPrefixSumsBase.cs:15);
  random inputs over n in [3072, 6144) in steps of 4, from the Unity
      generator (Shaders/Utility.compute:50-72, restated in the oracle) with
      the seed the reference's tests run (0), and with seed n;
  all ones over n in [3072 * 65536 - 1024, 3072 * 65536) in steps of 4, the
      sizes around the Unity path's dispatch split (k_maxDispatch = 65535
      workgroups per dimension). This library launches one grid dimension of
      up to 2^32 - 1 workgroups, so nothing splits here; the sizes stay as
      tile-boundary cases at 2^27.6 elements.

Ones are filled and checked on the device (the reference's InitOne /
ValidateOne closed forms, gps_fill_u32 / gps_validate_ones_u32); random inputs
are checked against the oracle's serial scan.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import empty_u32, first_mismatch, from_dev_u32, sync, to_dev  # noqa: E402

PARTITION = 3072    # GPUPrefixSumsBase.cs:15
MAX_DISPATCH = 65535  # GPUPrefixSumsBase.cs:17

ALGOS = {
    "csdl": dict(algo=gps.ALGO_CSDL),
    "rts": dict(algo=gps.ALGO_RTS),
    "csdldf": dict(algo=gps.ALGO_CSDLDF),
    "csdldf_split8": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_SPLIT, stall_period=8),
}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def ones_sweep(h, start, stop, inclusive):
    buf = empty_u32(stop)
    failed = []
    for n in range(start, stop, 4):
        gps.fill_u32(buf, n, 1)
        h.scan(buf, buf, n, inclusive)  # in place, as the Unity tests run
        errors = gps.validate_ones_u32(buf, n, inclusive)
        if errors:
            failed.append((n, errors))
    sync()
    return failed


@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("name", list(ALGOS))
def test_unity_ones_one_partition_above_2p25(name, inclusive):
    start = 1 << 25
    with gps.ScanHandle(max_n=start + PARTITION, **ALGOS[name]) as h:
        failed = ones_sweep(h, start, start + PARTITION, inclusive)
        assert not failed, f"{len(failed)} of {PARTITION // 4} sizes failed, first {failed[0]}"
        assert h.device_error() == 0


@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("name", list(ALGOS))
def test_unity_ones_around_the_dispatch_split(name, inclusive):
    stop = PARTITION * (MAX_DISPATCH + 1)
    with gps.ScanHandle(max_n=stop, **ALGOS[name]) as h:
        failed = ones_sweep(h, stop - 1024, stop, inclusive)
        assert not failed, f"{len(failed)} of 256 sizes failed, first {failed[0]}"
        assert h.device_error() == 0
    torch.cuda.empty_cache()


@pytest.mark.parametrize("seed_kind", ["reference_seed0", "seed_n"])
@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("name", list(ALGOS))
def test_unity_random_one_to_two_partitions(name, inclusive, seed_kind):
    with gps.ScanHandle(max_n=2 * PARTITION, **ALGOS[name]) as h:
        d_out = empty_u32(2 * PARTITION)
        for n in range(PARTITION, 2 * PARTITION, 4):
            x = oracle.gen_unity_taus(0 if seed_kind == "reference_seed0" else n, n)
            d_in = to_dev(x)
            h.scan(d_in, d_out, n, inclusive)
            sync()
            got, want = from_dev_u32(d_out)[:n], oracle.scan_u32(x, inclusive)
            assert np.array_equal(got, want), f"n={n}: {first_mismatch(got, want)}"
        assert h.device_error() == 0
