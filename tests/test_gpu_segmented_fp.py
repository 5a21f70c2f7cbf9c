"""Batched segmented FP inclusive scan (FP32 / BF16 input, FP32 output) on the
GPU against the per-segment FP64 oracle.

Error measures as in test_gpu_fp.py, with S_i the running sum of |x| from
the start of element i's segment:
  compensated:   |y_i - y64_i| <= 2^-23 |y64_i| + 2^-36 S_i
  uncompensated: |y_i - y64_i| <= (64 + 2 (t_i + 1)) 2^-24 S_i   (t_i: tile index)
Integer-valued inputs with small prefixes must come out exact, which checks
where every carry restarts independently of rounding.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402

SHAPES = [(256, 4), (1024, 4), (1024, 8)]
# one workgroup per tile, or the persistent pipelined kernel
SCHEDULES = {"tile": gps.SCHEDULE_TILE, "persistent": gps.SCHEDULE_PERSISTENT}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def run(h, x, off, bf16, comp):
    n = x.size
    d_in = torch.from_numpy(x.view(np.int16) if bf16 else x).cuda()
    d_off = torch.from_numpy(off.astype(np.uint64).view(np.int64)).cuda()
    out = torch.full((n + 8,), -7.0, dtype=torch.float32, device="cuda")
    h.segmented_fp(d_in, out, d_off, off.size - 1, n, bf16=bf16, compensated=comp)
    torch.cuda.synchronize()
    y = out.cpu().numpy()
    assert np.all(y[n:] == -7.0), "wrote past n"
    return y[:n]


def check_bounds(y, y64, S, comp, tile):
    err = np.abs(y.astype(np.float64) - y64)
    if comp:
        bound = 2.0 ** -23 * np.abs(y64) + 2.0 ** -36 * S
    else:
        bound = (64 + 2 * (np.arange(y.size) // tile + 1)) * 2.0 ** -24 * S
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, f"{bad.size} over the bound, first {bad[0]}: err {err[bad[0]]} bound {bound[bad[0]]}"


def offsets_cases(tile):
    rng = np.random.default_rng(11)
    cases = {
        "one_segment": np.array([0, 3 * tile + 5], dtype=np.uint64),
        "lognormal": oracle.gen_lognormal_offsets(3, 300, median=700.0, sigma=1.5, lo=1, hi=1 << 16),
        "empty_and_unit": np.cumsum(np.r_[0, rng.choice([0, 1, 2, 7, 300], size=500)]).astype(np.uint64),
        "heads_on_tile_edges": np.array([0, tile, 2 * tile, 2 * tile + 1, 4 * tile - 1, 4 * tile, 5 * tile + 3],
                                        dtype=np.uint64),
    }
    # a long run of segments that all start inside one tile, then one spanning many tiles
    cases["dense_then_long"] = np.r_[np.arange(0, 4000, 4), [4000 + 9 * tile + 17]].astype(np.uint64)
    # heads at the edges of 32-element runs (and of 16-element ones), two heads in one run, and runs
    # without a head between them: the lane-blocked kernel's first pass sums a run from its last
    # head, and its output pass resets only in lanes whose run holds a head
    pos = []
    for r in range(3 * 64):
        b = 32 * r
        pos += [[b], [b + 1], [b + 31], [b + 30], [b + 3, b + 29], [b + 15, b + 16], []][r % 7]
    cases["lane_run_edges"] = np.r_[sorted(set([0] + pos)), [32 * 3 * 64 + 2 * tile + 7]].astype(np.uint64)
    return cases


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_integer_valued_inputs_are_exact(shape, bf16, schedule):
    with gps.ScanHandle(gps.ALGO_CSDL, 1 << 22, block_threads=shape[0], vec_per_thread=shape[1],
                        schedule=SCHEDULES[schedule]) as h:
        for name, off in offsets_cases(h.tile).items():
            n = int(off[-1])
            v = (oracle.gen_u32(21 + n, n, 4).astype(np.int64) - 8).astype(np.float32)  # |prefix| < 2^24
            x = (v.view(np.uint32) >> 16).astype(np.uint16) if bf16 else v
            want = oracle.segmented_scan_fp_f64(v, off)
            for comp in (True, False):
                y = run(h, x, off, bf16, comp)
                bad = np.flatnonzero(y.astype(np.float64) != want)
                assert bad.size == 0, f"{name} comp={comp}: {bad.size} wrong, first {bad[0]}: {y[bad[0]]} vs {want[bad[0]]}"
        assert h.device_error() == 0


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("comp", [True, False])
@pytest.mark.parametrize("shape", SHAPES)
def test_f32_normal_bounds(shape, comp, schedule):
    with gps.ScanHandle(gps.ALGO_CSDL, 1 << 22, block_threads=shape[0], vec_per_thread=shape[1],
                        schedule=SCHEDULES[schedule]) as h:
        for name, off in offsets_cases(h.tile).items():
            n = int(off[-1])
            x = oracle.gen_f32_normal(31 + n, n)
            y = run(h, x, off, False, comp)
            check_bounds(y, oracle.segmented_scan_fp_f64(x, off), oracle.segmented_scan_fp_f64(x, off, absolute=True),
                         comp, h.tile)
        assert h.device_error() == 0


@pytest.mark.parametrize("comp", [True, False])
def test_bf16_logmag_bounds(comp):
    off = oracle.gen_lognormal_offsets(3, 1024, median=2048.0, sigma=1.5, lo=1, hi=1 << 20)
    n = int(off[-1])
    b = oracle.gen_bf16_logmag(5, n)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        y = run(h, b, off, True, comp)
        check_bounds(y, oracle.segmented_scan_fp_f64(b, off, bf16=True),
                     oracle.segmented_scan_fp_f64(b, off, bf16=True, absolute=True), comp, h.tile)
        assert h.device_error() == 0


@pytest.mark.parametrize("bf16", [False, True])
def test_default_handle_on_both_sides_of_2p22(bf16):
    # A default handle switches to the persistent 1024 x 8 kernel above 2^22
    # elements.
    for n_target in ((1 << 22) - 1000, (1 << 22) + 40_000):
        off = oracle.gen_lognormal_offsets(13, 700, median=4096.0, sigma=1.0, lo=1, hi=1 << 20)
        off = np.minimum(off, n_target).astype(np.uint64)
        off[-1] = n_target
        n = n_target
        if bf16:
            x = oracle.gen_bf16_logmag(17, n)
        else:
            x = oracle.gen_f32_normal(17, n)
        with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
            y = run(h, x, off, bf16, True)
            assert h.device_error() == 0
        check_bounds(y, oracle.segmented_scan_fp_f64(x, off, bf16=bf16),
                     oracle.segmented_scan_fp_f64(x, off, bf16=bf16, absolute=True), True, 1)


def test_one_segment_matches_plain_fp_scan_bounds_and_per_segment_launches():
    x = oracle.gen_f32_normal(4, (1 << 20) + 3)
    n = x.size
    off = np.array([0, n], dtype=np.uint64)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        seg = run(h, x, off, False, True)
        d_in = torch.from_numpy(x).cuda()
        plain = torch.empty(n, dtype=torch.float32, device="cuda")
        h.inclusive_f32(d_in, plain, n, True)
        torch.cuda.synchronize()
        plain = plain.cpu().numpy()
    y64 = oracle.scan_f32_f64(x)
    e_seg = np.abs(seg - y64) / np.maximum(np.abs(y64), 1.0)
    e_plain = np.abs(plain - y64) / np.maximum(np.abs(y64), 1.0)
    assert e_seg.max() <= 2.0 ** -23 and e_plain.max() <= 2.0 ** -23


def test_batched_matches_per_segment_launches_within_bounds():
    off = oracle.gen_lognormal_offsets(7, 200, median=3000.0, sigma=1.0, lo=1, hi=1 << 16)
    n = int(off[-1])
    x = oracle.gen_f32_normal(12, n)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        batched = run(h, x, off, False, True)
        d_in = torch.from_numpy(x).cuda()
        loop = torch.empty(n, dtype=torch.float32, device="cuda")
        h.fp_per_segment(d_in, loop, off, bf16=False, compensated=True)
        torch.cuda.synchronize()
        loop = loop.cpu().numpy()
    y64 = oracle.segmented_scan_fp_f64(x, off)
    S = oracle.segmented_scan_fp_f64(x, off, absolute=True)
    for y in (batched, loop):
        check_bounds(y, y64, S, True, 1)
