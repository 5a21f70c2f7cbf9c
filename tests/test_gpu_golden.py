"""The library against the committed golden fixtures (tests/golden/*.npz):
bit-exact for the integer scans, within the stated bounds for FP."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"
HANDLES = {
    "csdl_default": dict(algo=gps.ALGO_CSDL),
    "csdl_tile": dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_TILE),
    "csdl_persistent": dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_PERSISTENT),
    "csdl_control": dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_CONTROL),
    "rts": dict(algo=gps.ALGO_RTS),
    "csdldf_stall7": dict(algo=gps.ALGO_CSDLDF, stall_period=7),
    "csdldf_pp_stall8": dict(algo=gps.ALGO_CSDLDF, stall_period=8, schedule=gps.SCHEDULE_PERSISTENT),
}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


@pytest.mark.parametrize("handle", list(HANDLES))
@pytest.mark.parametrize("name", ["u32_small_values", "u32_full_range", "u32_ones"])
def test_u32_golden(name, handle):
    f = load(name)
    x = f["x"]
    n = x.size
    kw = dict(HANDLES[handle])
    algo = kw.pop("algo")
    with gps.ScanHandle(algo, n, **kw) as h:
        d_in = to_dev(x)
        for inclusive, key in ((True, "incl"), (False, "excl")):
            d_out = sentinel_u32(n)
            h.scan(d_in, d_out, n, inclusive)
            sync()
            got = from_dev_u32(d_out)
            assert np.array_equal(got, f[key]), f"{name} {handle} {key}: {first_mismatch(got, f[key])}"
        assert h.device_error() == 0


@pytest.mark.parametrize("schedule", [0, gps.SCHEDULE_TILE])
def test_segmented_golden(schedule):
    f = load("segmented_u32")
    x, off = f["x"], f["offsets"]
    n = x.size
    with gps.ScanHandle(gps.ALGO_CSDL, n, schedule=schedule) as h:
        d_in, d_off = to_dev(x), to_dev(off)
        for inclusive, key in ((True, "incl"), (False, "excl")):
            d_out = sentinel_u32(n)
            h.segmented(d_in, d_out, d_off, off.size - 1, n, inclusive)
            sync()
            got = from_dev_u32(d_out)
            assert np.array_equal(got, f[key]), f"{key}: {first_mismatch(got, f[key])}"


def test_struct_golden():
    f = load("struct_u32x4")
    x = f["x"]
    n = x.shape[0]
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_in = to_dev(x.reshape(-1))
        for inclusive, key in ((True, "incl"), (False, "excl")):
            d_out = sentinel_u32(4 * n)
            h.scan_u32x4(d_in, d_out, n, inclusive)
            sync()
            got = from_dev_u32(d_out).reshape(-1, 4)
            assert np.array_equal(got, f[key]), f"{key}: {first_mismatch(got, f[key])}"


@pytest.mark.parametrize("mode", [gps.FP_PAIR, gps.FP_PLAIN])
@pytest.mark.parametrize("name", ["f32_normal", "bf16_logmag"])
def test_fp_golden(name, mode):
    f = load(name)
    x, y64, s64 = f["x"], f["y64"], f["s64"]
    n = x.size
    bf16 = name == "bf16_logmag"
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_in = torch.from_numpy(x.view(np.int16) if bf16 else x).cuda()
        d_out = torch.empty(n, dtype=torch.float32, device="cuda")
        (h.inclusive_bf16_f32 if bf16 else h.inclusive_f32)(d_in, d_out, n, mode)
        sync()
        y = d_out.cpu().numpy().astype(np.float64)
    err = np.abs(y - y64)
    if mode == gps.FP_PLAIN:
        bound = (64 + 2 * 8) * 2.0 ** -24 * s64  # n < 8 tiles of any FP shape
    else:
        bound = 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * s64
    assert np.all(err <= bound), f"{name} mode {mode}: {np.count_nonzero(err > bound)} over the bound"
