"""Batched segmented uint32 scan on the GPU, bit-exact against the per-segment
serial oracle; the one-launch kernel, its hipGraph replays, and the
per-segment-launch baseline must all give the same output."""
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


def layouts():
    r = np.random.default_rng(5)
    yield "lognormal64", oracle.gen_lognormal_offsets(11, 3000, median=64.0, sigma=1.5, lo=1, hi=1 << 16)
    yield "tiny", np.concatenate([[0], np.cumsum(r.integers(1, 4, 20000))]).astype(np.uint64)
    lens = r.integers(0, 50, 5000)
    lens[::7] = 0  # empty segments
    yield "with_empty", np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    yield "single", np.array([0, 1_000_003], np.uint64)
    yield "tile_aligned", np.arange(0, 40 * 16384 + 1, 16384, dtype=np.uint64)
    yield "long_and_short", np.concatenate([[0], np.cumsum([1, 700_001, 3, 16383, 16385, 2, 1 << 20])]).astype(np.uint64)
    yield "lognormal4096", oracle.gen_lognormal_offsets(3, 512, median=4096.0, sigma=1.5, lo=1, hi=1 << 22)


LAYOUTS = list(layouts())


SCHEDULES = {"persistent": gps.SCHEDULE_PERSISTENT, "tile": gps.SCHEDULE_TILE}


@pytest.mark.parametrize("inclusive", [False, True])
@pytest.mark.parametrize("schedule", list(SCHEDULES))
@pytest.mark.parametrize("shape", [(0, 0), (256, 4), (512, 4), (1024, 2), (1024, 4), (1024, 8)])
@pytest.mark.parametrize("name,off", LAYOUTS, ids=[x[0] for x in LAYOUTS])
def test_segmented_matches_oracle(name, off, shape, schedule, inclusive):
    """Both segmented kernels: the persistent one (default schedule) and one
    workgroup per tile, at every shape (the default handle's is 1024x8)."""
    n = int(off[-1])
    x = oracle.gen_u32(17 + n, n, 32)
    want = oracle.segmented_scan_u32(x, off, inclusive)
    with gps.ScanHandle(gps.ALGO_CSDL, max(n, 1), block_threads=shape[0], vec_per_thread=shape[1],
                        schedule=SCHEDULES[schedule]) as h:
        d_in, d_off = to_dev(x), to_dev(off)
        d_out = sentinel_u32(n + 16)
        h.segmented(d_in, d_out, d_off, off.size - 1, n, inclusive)
        sync()
        out = from_dev_u32(d_out)
        assert np.array_equal(out[:n], want), f"{name} {shape}: {first_mismatch(out[:n], want)}"
        assert np.all(out[n:n + 16] == 0xDEADBEEF)
        assert h.device_error() == 0


@pytest.mark.parametrize("inclusive", [False, True])
def test_per_segment_baseline_equal_output(inclusive):
    off = oracle.gen_lognormal_offsets(3, 300, median=4096.0, sigma=1.5, lo=1, hi=1 << 20)
    n = int(off[-1])
    x = oracle.gen_u32(3, n, 10)
    want = oracle.segmented_scan_u32(x, off, inclusive)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_in, d_off = to_dev(x), to_dev(off)
        a, b = sentinel_u32(n), sentinel_u32(n)
        h.segmented(d_in, a, d_off, off.size - 1, n, inclusive)
        h.segmented_per_segment(d_in, b, off, inclusive)
        sync()
        assert np.array_equal(from_dev_u32(a, n), want)
        assert np.array_equal(from_dev_u32(b, n), want)


def test_graph_replays_match_eager():
    # Replay several times with eager work and synchronisations in between;
    # every replay must equal the eager result (no memset nodes are used).
    off = oracle.gen_lognormal_offsets(21, 800, median=2048.0, sigma=1.5, lo=1, hi=1 << 18)
    n = int(off[-1])
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_off = to_dev(off)
        d_out = sentinel_u32(n)
        for rep in range(4):
            x = oracle.gen_u32(100 + rep, n, 32)
            d_in = to_dev(x)
            want = oracle.segmented_scan_u32(x, off, False)
            with h.segmented_graph(d_in, d_out, d_off, off.size - 1, n, False) as g:
                for k in range(3):
                    d_out.fill_(0)
                    g.launch()
                    sync()
                    got = from_dev_u32(d_out, n)
                    assert np.array_equal(got, want), f"replay {rep}.{k}: {first_mismatch(got, want)}"
                    h.inclusive(d_in, d_out, n)  # eager work on the same handle between replays
                    sync()
        assert h.device_error() == 0


def test_config3_lognormal_4096_segments():
    # 4096 segments, log-normal lengths (median 4096, sigma 1.5, clipped to
    # [1, 2^22]), values uniform in [0, 1023], seed 3; exclusive.
    off = oracle.gen_lognormal_offsets(3, 4096, median=4096.0, sigma=1.5, lo=1, hi=1 << 22)
    n = int(off[-1])
    x = oracle.gen_u32(3, n, 10)
    want = oracle.segmented_scan_u32(x, off, False)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_in, d_off = to_dev(x), to_dev(off)
        d_out = sentinel_u32(n)
        with h.segmented_graph(d_in, d_out, d_off, 4096, n, False) as g:
            g.launch()
            sync()
        got = from_dev_u32(d_out, n)
        assert np.array_equal(got, want), first_mismatch(got, want)
        d_out.fill_(0)
        h.segmented_per_segment(d_in, d_out, off, False)
        sync()
        assert np.array_equal(from_dev_u32(d_out, n), want)


def test_segmented_argument_errors():
    with gps.ScanHandle(gps.ALGO_RTS, 1000) as h:
        d = sentinel_u32(1000)
        off = to_dev(np.array([0, 1000], np.uint64))
        with pytest.raises(gps.GpsError):
            h.segmented(d, d, off, 1, 1000)
    with gps.ScanHandle(gps.ALGO_CSDL, 1000) as h:
        big = sentinel_u32(1001)
        with pytest.raises(gps.GpsError):  # n above the handle's max_n
            h.segmented(big, big, off, 1, 1001)
        with pytest.raises(ValueError, match="buffer"):  # n past the tensors: refused on the host
            h.segmented(d, d, off, 1, 1001)
        with pytest.raises(gps.GpsError):
            h.segmented(d, d, 0, 1, 1000)
