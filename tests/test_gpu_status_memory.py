"""Status memory: handles keep their ticket and status words in uncached
device memory that the library never frees, recycling released blocks through
a per-device free list; GPS_STATUS_MEMORY=plain selects ordinary memory. Both
kinds give the oracle's bits, and the free list bounds what is held.
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import empty_u32, first_mismatch, from_dev_u32, sync, to_dev  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


@pytest.fixture(autouse=True)
def restore_env():
    old = os.environ.get("GPS_STATUS_MEMORY")
    yield
    if old is None:
        os.environ.pop("GPS_STATUS_MEMORY", None)
    else:
        os.environ["GPS_STATUS_MEMORY"] = old


def use_memory(kind):
    if kind == "plain":
        os.environ["GPS_STATUS_MEMORY"] = "plain"
    else:
        os.environ.pop("GPS_STATUS_MEMORY", None)


CASES = [
    ("csdl_persistent", dict(algo=gps.ALGO_CSDL, block_threads=1024, vec_per_thread=8,
                             schedule=gps.SCHEDULE_PERSISTENT)),
    ("csdl_tile", dict(algo=gps.ALGO_CSDL, block_threads=1024, vec_per_thread=8, schedule=gps.SCHEDULE_TILE)),
    ("csdldf_stall7", dict(algo=gps.ALGO_CSDLDF, stall_period=7)),
    ("csdldf_pp_stall3", dict(algo=gps.ALGO_CSDLDF, stall_period=3, schedule=gps.SCHEDULE_PERSISTENT)),
]


def scan_both_memories(kw, n, seed):
    x = oracle.gen_u32(seed, n, 32)
    d_in, d_out = to_dev(x), empty_u32(n)
    got = {}
    for kind in ("uncached", "plain"):
        use_memory(kind)
        kw2 = dict(kw)
        algo = kw2.pop("algo")
        with gps.ScanHandle(algo, n, **kw2) as h:
            for inclusive in (True, False):
                for _ in range(2):  # a second launch runs under the next epoch
                    h.scan(d_in, d_out, n, inclusive)
                sync()
                assert h.device_error() == 0
                got[kind, inclusive] = from_dev_u32(d_out, n).copy()
    return x, got


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_plain_and_uncached_status_give_the_oracle(name, kw):
    n = (1 << 22) + 12345
    x, got = scan_both_memories(kw, n, seed=len(name))
    for inclusive in (True, False):
        want = oracle.scan_u32(x, inclusive)
        for kind in ("uncached", "plain"):
            assert np.array_equal(got[kind, inclusive], want), \
                f"{name} {kind} inclusive={inclusive}: {first_mismatch(got[kind, inclusive], want)}"


def test_released_blocks_are_reused():
    use_memory("uncached")
    n = (1 << 21) + 7
    with gps.ScanHandle(gps.ALGO_CSDL, n):
        held0, _ = gps.status_memory()
    held1, free1 = gps.status_memory()
    assert held1 == held0 and free1 >= n // 8, "a destroyed handle's block is not on the free list"
    # the same size, then a smaller one, take blocks from the free list
    with gps.ScanHandle(gps.ALGO_CSDL, n):
        held2, free2 = gps.status_memory()
        assert held2 == held1 and free2 < free1
    with gps.ScanHandle(gps.ALGO_CSDL, n // 4):
        assert gps.status_memory()[0] == held1
    # a scan through a recycled block still matches the oracle
    x = oracle.gen_u32(9, n, 32)
    d_in, d_out = to_dev(x), empty_u32(n)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        assert gps.status_memory()[0] == held1
        h.scan(d_in, d_out, n, True)
        sync()
        assert np.array_equal(from_dev_u32(d_out, n), oracle.scan_u32(x, True))


def test_plain_memory_is_not_held():
    use_memory("plain")
    held0, free0 = gps.status_memory()
    with gps.ScanHandle(gps.ALGO_CSDL, 1 << 20):
        assert gps.status_memory() == (held0, free0)
    assert gps.status_memory() == (held0, free0)


@pytest.mark.parametrize("kind", ["uncached", "plain"])
def test_fp_pair_scan_in_either_memory(kind):
    # the FP scans publish both status words of a tile in one 16-byte access
    use_memory(kind)
    n = (1 << 22) + 12345
    x = oracle.gen_f32_normal(4, n)
    y64 = oracle.scan_f32_f64(x)
    S = oracle.abs_scan_f32_f64(x)
    d_in = torch.from_numpy(x).cuda()
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        for comp in (gps.FP_PAIR, gps.FP_PLAIN):
            out = torch.full((n + 8,), -7.0, dtype=torch.float32, device="cuda")
            for _ in range(2):
                h.inclusive_f32(d_in, out, n, comp)
            sync()
            assert h.device_error() == 0
            y = out.cpu().numpy()
            assert np.all(y[n:] == -7.0), "wrote past n"
            err = np.abs(y[:n].astype(np.float64) - y64)
            if comp == gps.FP_PAIR:
                bound = 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S
            else:
                # tile index at the smallest FP tile (4096): looser than the kernel's own
                bound = (64 + 2 * (np.arange(n) // 4096 + 1)) * 2.0 ** -24 * S
            assert np.all(err <= bound), f"{kind} comp={comp}: {int(np.sum(err > bound))} elements over the bound"
