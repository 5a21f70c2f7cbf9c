"""FP32 / BF16-input inclusive scans on the GPU against the FP64 oracle.

Error measures (stated with every result):
  rel_i = |y_i - y64_i| / max(|y64_i|, 1)   (relative, floored at 1: both
          inputs have unit median magnitude, so prefixes near a zero crossing
          are judged in absolute terms)
  S_i   = sum_{j <= i} |x_j|                 (for normwise bounds)
Modes: plain FP32 carries (0), two-float TwoSum pairs (1).
Bounds asserted:
  compensated (pairs):
                 |y_i - y64_i| <= 2^-23 |y64_i| + 2^-36 S_i
                 (one FP32 rounding of the exact prefix, plus the u^2-level
                 error of the pair arithmetic)
  uncompensated: |y_i - y64_i| <= (64 + 2 (t_i + 1)) 2^-24 S_i
                 (classical bound for a sum whose dependency path has at most
                 that many additions; t_i = tile index of element i)
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402

MODES = [gps.FP_PAIR, gps.FP_PLAIN]
MODE_IDS = ["pair", "plain"]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def run(h, x, bf16, comp):
    n = x.size
    d_in = torch.from_numpy(x.view(np.int16) if bf16 else x).cuda()
    out = torch.full((n + 8,), -7.0, dtype=torch.float32, device="cuda")
    (h.inclusive_bf16_f32 if bf16 else h.inclusive_f32)(d_in, out, n, comp)
    torch.cuda.synchronize()
    y = out.cpu().numpy()
    assert np.all(y[n:] == -7.0), "wrote past n"
    return y[:n]


def check_bounds(y, y64, S, comp, tile):
    err = np.abs(y.astype(np.float64) - y64)
    if comp:
        bound = 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S
    else:
        t = np.arange(y.size) // tile
        bound = (64 + 2 * (t + 1)) * 2.0 ** -24 * S
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, f"{bad.size} elements over the bound, first {bad[0]}: err {err[bad[0]]} bound {bound[bad[0]]}"
    return float(np.max(err / np.maximum(np.abs(y64), 1.0))), float(np.median(err / np.maximum(np.abs(y64), 1.0)))


@pytest.mark.parametrize("comp", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("bf16", [False, True])
def test_integer_valued_inputs_are_exact(bf16, comp):
    # |prefix| <= 8 * 2^20 = 2^23: every partial sum is exact in FP32.
    n = (1 << 20) + 5
    v = (oracle.gen_u32(9, n, 4).astype(np.int64) - 8).astype(np.float32)
    x = (v.view(np.uint32) >> 16).astype(np.uint16) if bf16 else v  # small integers are exact in bf16
    want = np.cumsum(v.astype(np.float64))
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        y = run(h, x, bf16, comp)
    assert np.array_equal(y.astype(np.float64), want)


@pytest.mark.parametrize("schedule", ["tile", "persistent"])
@pytest.mark.parametrize("shape", [(256, 4), (1024, 4), (1024, 8), (0, 0)])
@pytest.mark.parametrize("comp", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("n", [1, 3, 4095, 4097, 65537, (1 << 22) + 7])
def test_f32_normal_bounds(n, comp, shape, schedule):
    # One workgroup per tile, or the persistent pipelined kernel.
    x = oracle.gen_f32_normal(4, n)
    y64 = oracle.scan_f32_f64(x)
    S = oracle.abs_scan_f32_f64(x)
    sched = {"tile": gps.SCHEDULE_TILE, "persistent": gps.SCHEDULE_PERSISTENT}[schedule]
    with gps.ScanHandle(gps.ALGO_CSDL, n, block_threads=shape[0], vec_per_thread=shape[1], schedule=sched) as h:
        tile = h.tile  # the FP kernels run the handle's shape (all shapes here are FP shapes)
        y = run(h, x, False, comp)
        assert h.device_error() == 0
    check_bounds(y, y64, S, comp, tile)


@pytest.mark.parametrize("comp", MODES, ids=MODE_IDS)
def test_bf16_logmag_bounds(comp):
    n = (1 << 22) + 3
    b = oracle.gen_bf16_logmag(5, n)
    y64 = oracle.scan_bf16_f64(b)
    S = oracle.abs_scan_bf16_f64(b)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        y = run(h, b, True, comp)
        tile = h.tile
    check_bounds(y, y64, S, comp, tile)


def test_compensation_beats_plain_fp32():
    n = 1 << 24
    x = oracle.gen_f32_normal(4, n)
    y64 = oracle.scan_f32_f64(x)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        yc = run(h, x, False, gps.FP_PAIR)
        yp = run(h, x, False, gps.FP_PLAIN)
    ec = np.abs(yc - y64) / np.maximum(np.abs(y64), 1.0)
    ep = np.abs(yp - y64) / np.maximum(np.abs(y64), 1.0)
    assert np.median(ec) < np.median(ep) and ec.max() < ep.max()
    assert ec.max() <= 2.0 ** -23


@pytest.mark.parametrize("bf16", [False, True])
def test_config4_n2p27(bf16):
    # n = 2^27; FP32 ~ N(0,1) (seed 4) or BF16 +-10^U(-4,4) (seed 5).
    n = 1 << 27
    if bf16:
        x = oracle.gen_bf16_logmag(5, n)
        y64, S = oracle.scan_bf16_f64(x), oracle.abs_scan_bf16_f64(x)
    else:
        x = oracle.gen_f32_normal(4, n)
        y64, S = oracle.scan_f32_f64(x), oracle.abs_scan_f32_f64(x)
    report = {}
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        for comp in MODES:
            y = run(h, x, bf16, comp)
            report[comp] = check_bounds(y, y64, S, comp, h.tile)
            del y
        assert h.device_error() == 0
    print(f"n=2^27 {'bf16' if bf16 else 'f32'}: pair max/median rel {report[gps.FP_PAIR]}, "
          f"plain {report[gps.FP_PLAIN]}")
    assert report[gps.FP_PAIR][0] <= 2.0 ** -23 and report[gps.FP_PAIR][0] < report[gps.FP_PLAIN][0]


def test_carry_mode_errors():
    d = torch.zeros(1024, dtype=torch.float32, device="cuda")
    off = torch.tensor([0, 1024], dtype=torch.int64, device="cuda")
    with gps.ScanHandle(gps.ALGO_CSDL, 1024) as h:
        for bad in (2, 3, -1):  # only plain (0) and pairs (1) exist; no FP64 mode
            with pytest.raises(gps.GpsError, match="compensated"):
                h.inclusive_f32(d, d, 1024, bad)
            with pytest.raises(gps.GpsError, match="compensated"):
                h.segmented_fp(d, d, off, 1, 1024, compensated=bad)
            with pytest.raises(gps.GpsError, match="compensated"):
                h.fp_per_segment(d, d, [0, 1024], compensated=bad)
        torch.cuda.synchronize()
