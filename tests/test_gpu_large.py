"""n = 2^28 full-range uint32 scans (1 GiB in, 1 GiB out), bit-exact against
the CPU oracle over every element: CSDL, reduce-then-scan, and the decoupled
fallback under emulated deadlock with stall periods 7 (every 7th tile) and 8
(the reference's !(idx & 7), GPUPrefixSumsCUDA/EmulatedDeadlocking.cuh:16,65);
and scans of more than 2^32 elements, checked on the device.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import empty_u32, first_mismatch, from_dev_u32, sync, to_dev  # noqa: E402

N = 1 << 28


@pytest.fixture(scope="module")
def data():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    x = oracle.gen_u32(2, N, 32)
    d_in = to_dev(x)
    d_out = empty_u32(N)
    return x, d_in, d_out


@pytest.fixture(scope="module")
def want_incl(data):
    return oracle.scan_u32(data[0], True)


CASES = [
    ("csdl", dict(algo=gps.ALGO_CSDL)),
    ("rts", dict(algo=gps.ALGO_RTS)),
    ("csdldf_stall7", dict(algo=gps.ALGO_CSDLDF, stall_period=7)),
    ("csdldf_stall8", dict(algo=gps.ALGO_CSDLDF, stall_period=8)),
    ("csdldf_pp_stall7", dict(algo=gps.ALGO_CSDLDF, stall_period=7, schedule=gps.SCHEDULE_PERSISTENT)),
    ("csdldf_pp_stall8", dict(algo=gps.ALGO_CSDLDF, stall_period=8, schedule=gps.SCHEDULE_PERSISTENT)),
    ("csdldf_split7", dict(algo=gps.ALGO_CSDLDF, stall_period=7, schedule=gps.SCHEDULE_SPLIT)),
    ("csdldf_split8", dict(algo=gps.ALGO_CSDLDF, stall_period=8, schedule=gps.SCHEDULE_SPLIT)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_inclusive_2p28_full_range(data, want_incl, name, kw):
    x, d_in, d_out = data
    with gps.ScanHandle(max_n=N, **kw) as h:
        d_out.fill_(0)
        h.inclusive(d_in, d_out, N)
        sync()
        got = from_dev_u32(d_out)
        assert np.array_equal(got, want_incl), f"{name}: {first_mismatch(got, want_incl)}"
        # second launch on the same handle (no clearing in between)
        h.inclusive(d_in, d_out, N)
        sync()
        assert np.array_equal(from_dev_u32(d_out), want_incl)
        if kw.get("stall_period"):
            c = h.counters()
            assert c["stall_timeouts"] == 0 and c["insertions"] > 0
        assert h.device_error() == 0


def test_exclusive_2p28_full_range(data):
    x, d_in, d_out = data
    want = oracle.scan_u32(x, False)
    for kw in (dict(algo=gps.ALGO_CSDL), dict(algo=gps.ALGO_CSDLDF, stall_period=7),
               dict(algo=gps.ALGO_CSDLDF, stall_period=8, schedule=gps.SCHEDULE_PERSISTENT),
               dict(algo=gps.ALGO_CSDLDF, stall_period=8, schedule=gps.SCHEDULE_SPLIT)):
        with gps.ScanHandle(max_n=N, **kw) as h:
            h.exclusive(d_in, d_out, N)
            sync()
            got = from_dev_u32(d_out)
            assert np.array_equal(got, want), first_mismatch(got, want)


def test_beyond_2p32_elements():
    # n = 2^32 + 4099 uint32 (16 GiB in, 16 GiB out): every index, tile base
    # and status slot above 32 bits. Checked on the device: the validator
    # compares every difference scan[i] - scan[i-1] with the input and the
    # first element with its closed form, which fixes the whole scan; a
    # planted error above 2^32 shows that the check reaches that far.
    n = (1 << 32) + 4099
    d_in = torch.empty(n, dtype=torch.int32, device="cuda")
    gps.fill_random_u32(d_in, n, seed=32, bits=32)
    d_out = torch.empty(n, dtype=torch.int32, device="cuda")
    cases = [dict(algo=gps.ALGO_CSDL), dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_TILE), dict(algo=gps.ALGO_RTS),
             dict(algo=gps.ALGO_CSDLDF, stall_period=8)]
    for kw in cases:
        with gps.ScanHandle(max_n=n, **kw) as h:
            for inclusive in (True, False):
                h.scan(d_in, d_out, n, inclusive)
                sync()
                assert gps.validate_scan_u32(d_in, d_out, n, inclusive) == 0, (kw, inclusive)
                assert h.device_error() == 0
    # the batched segmented scan with heads below and above 2^32: each segment
    # checked as a plain scan of its own range
    off = np.array([0, 1 << 31, (1 << 32) + 100, n], dtype=np.uint64)
    d_off = torch.from_numpy(off.view(np.int64)).cuda()
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        for inclusive in (True, False):
            h.segmented(d_in, d_out, d_off, off.size - 1, n, inclusive)
            sync()
            for a, b in zip(off[:-1].tolist(), off[1:].tolist()):
                assert gps.validate_scan_u32(d_in[a:b], d_out[a:b], b - a, inclusive) == 0, (a, b, inclusive)
            assert h.device_error() == 0
    # d_out holds the segmented exclusive scan: plant the error inside the
    # segment [2^31, 2^32 + 100), above 2^32
    k = (1 << 32) + 5
    d_out[k] += 1
    sync()
    a, b = 1 << 31, (1 << 32) + 100
    assert gps.validate_scan_u32(d_in[a:b], d_out[a:b], b - a, False) == 2
    del d_in, d_out
    torch.cuda.empty_cache()


def test_fp_beyond_2p32_elements():
    # n = 2^32 + 4099 FP32 ones: with pair carries every output is the
    # exact prefix i + 1 rounded once to FP32 (round to nearest even), so it
    # must equal float32(i + 1) bit for bit, above 2^32 too.
    n = (1 << 32) + 4099
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        for comp in (gps.FP_PAIR,):
            y.fill_(-1.0)
            h.inclusive_f32(x, y, n, comp)
            sync()
            assert h.device_error() == 0
            step = 1 << 28
            for a in range(0, n, step):
                b = min(a + step, n)
                want = (torch.arange(a, b, dtype=torch.float64, device="cuda") + 1.0).to(torch.float32)
                bad = torch.nonzero(y[a:b] != want)
                assert bad.numel() == 0, (comp, a + int(bad[0]), float(y[a + int(bad[0])]), float(want[int(bad[0])]))
                del want
    del x, y
    torch.cuda.empty_cache()
