"""Member-wise scan of 4-member uint32 structs on the GPU, bit-exact against
the per-member oracle (through the C ABI).

The reference's own check is all-ones structs against (i >> 2) + 1 per uint32
(GPUPrefixSumsWebGPUapis/SharedShaders/validate.wgsl:51-60, run after
csdldf_struct.wgsl). Added here: members with different data (a lookback that
mixed members up would show), full-range wraparound, exclusive output, tails
around one to three tiles of every shape, in-place buffers, the smallest
handle that admits n, repeated launches and argument errors.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import first_mismatch, from_dev_u32, sentinel_u32, sync, to_dev  # noqa: E402

SHAPES = [(512, 8), (1024, 4), (1024, 8)]
SCHEDULES = {"tile": gps.SCHEDULE_TILE, "persistent": gps.SCHEDULE_PERSISTENT}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def run(h, x: np.ndarray, inclusive: bool, in_place: bool = False) -> np.ndarray:
    """x: (n, 4) uint32. Returns the (n, 4) device result; checks that the
    64 uint32 after the output are untouched."""
    n = x.shape[0]
    flat = np.ascontiguousarray(x, np.uint32).reshape(-1)
    if in_place:
        d = to_dev(flat, pad=64) if n else sentinel_u32(64)
        h.scan_u32x4(d, d, n, inclusive)
        out_t = d
    else:
        d_in = to_dev(flat) if n else sentinel_u32(4)
        out_t = sentinel_u32(4 * n + 64)
        h.scan_u32x4(d_in, out_t, n, inclusive)
    sync()
    out = from_dev_u32(out_t)
    assert np.all(out[4 * n:4 * n + 64] == 0xDEADBEEF), "wrote past the last struct"
    return out[:4 * n].reshape(-1, 4)


def check(h, x, inclusive, label="", in_place=False):
    got = run(h, x, inclusive, in_place)
    want = oracle.scan_u32x4(x, inclusive)
    assert np.array_equal(got, want), f"{label} n={x.shape[0]} incl={inclusive}: {first_mismatch(got, want)}"


def structs(seed: int, n: int, bits: int = 32) -> np.ndarray:
    return oracle.gen_u32(seed, 4 * n, bits).reshape(-1, 4)


def sizes(tile: int):
    s = {1, 2, 3, 63, 64, 65, 1000}
    for k in (1, 2, 3):
        s.update({k * tile - 1, k * tile, k * tile + 1})
    return sorted(s)


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_all_ones_known_answer(shape, schedule):
    block, vec = shape
    tile = block * vec
    with gps.ScanHandle(gps.ALGO_CSDL, 4 * tile + 8, block_threads=block, vec_per_thread=vec,
                        schedule=SCHEDULES[schedule]) as h:
        for n in sizes(tile):
            got = run(h, np.ones((n, 4), np.uint32), True).reshape(-1)
            want = (np.arange(4 * n, dtype=np.uint32) >> 2) + 1
            assert np.array_equal(got, want), f"{shape} n={n}: {first_mismatch(got, want)}"


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("inclusive", [True, False], ids=["incl", "excl"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_random_members_full_range(shape, inclusive, schedule):
    block, vec = shape
    tile = block * vec
    with gps.ScanHandle(gps.ALGO_CSDL, 4 * tile + 8, block_threads=block, vec_per_thread=vec,
                        schedule=SCHEDULES[schedule]) as h:
        for n in sizes(tile):
            check(h, structs(1000 + n, n), inclusive, f"{shape}")


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("inclusive", [True, False], ids=["incl", "excl"])
def test_in_place(inclusive, schedule):
    with gps.ScanHandle(gps.ALGO_CSDL, 1 << 20, schedule=SCHEDULES[schedule]) as h:
        for n in (1, 4095, 4096, 40_000):
            check(h, structs(5 + n, n), inclusive, "in place", in_place=True)


def test_members_carry_different_streams():
    # One member all ones, one all 0xFFFFFFFF (wraps), one zeros, one random:
    # a lookback that took a word from the wrong member would show at once.
    n = 300_000
    x = np.empty((n, 4), np.uint32)
    x[:, 0] = 1
    x[:, 1] = 0xFFFFFFFF
    x[:, 2] = 0
    x[:, 3] = oracle.gen_u32(77, n, 32)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        for inclusive in (True, False):
            check(h, x, inclusive, "mixed members")


def test_smallest_handle_admits_every_n():
    # A handle sized for max_n uint32 holds the four status words per tile
    # for every n <= max_n structs, down to n = 1.
    for max_n in (1, 1000, 4096, 5000):
        with gps.ScanHandle(gps.ALGO_CSDL, max_n) as h:
            for n in sorted({1, max_n // 2 or 1, max_n}):
                check(h, structs(max_n + n, n), True, f"max_n={max_n}")


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_repeated_launches_reuse_status_words(schedule):
    # Alternating the plain and the struct scan on one handle: each launch's
    # epoch makes the other kernel's status words read as not ready.
    n = 1 << 18
    x = structs(3, n)
    want = oracle.scan_u32x4(x, True)
    d_in = to_dev(x.reshape(-1))
    d_out = sentinel_u32(4 * n)
    d_tmp = sentinel_u32(4 * n)
    with gps.ScanHandle(gps.ALGO_CSDL, 4 * n, schedule=SCHEDULES[schedule]) as h:
        for _ in range(20):
            h.scan_u32x4(d_in, d_out, n, True)
            h.inclusive(d_in, d_tmp, 4 * n)
        sync()
        got = from_dev_u32(d_out).reshape(-1, 4)
        assert np.array_equal(got, want), first_mismatch(got, want)
        assert h.device_error() == 0


def test_large_2p25_structs():
    n = 1 << 25  # 512 MiB in, 512 MiB out
    x = structs(25, n)
    want = oracle.scan_u32x4(x, True)
    d_in = to_dev(x.reshape(-1))
    d_out = torch.empty(4 * n, dtype=torch.int32, device="cuda")
    for name, schedule in SCHEDULES.items():
        with gps.ScanHandle(gps.ALGO_CSDL, n, schedule=schedule) as h:
            h.scan_u32x4(d_in, d_out, n, True)
            sync()
            got = from_dev_u32(d_out).reshape(-1, 4)
            assert np.array_equal(got, want), f"{name}: {first_mismatch(got, want)}"
            assert h.device_error() == 0


@pytest.mark.parametrize("inclusive", [True, False], ids=["incl", "excl"])
def test_default_handle_across_size_thresholds(inclusive):
    # A default handle runs one workgroup per tile at 1024x4 up to 2^19
    # structs, at 1024x8 below 2^22, and the persistent kernel from 2^22.
    sizes = [1, (1 << 19) - 1, 1 << 19, (1 << 19) + 1, (1 << 22) - 1, 1 << 22, (1 << 22) + 3]
    with gps.ScanHandle(gps.ALGO_CSDL, sizes[-1]) as h:
        for n in sizes:
            check(h, structs(n, n), inclusive, f"default handle n={n}")
        assert h.device_error() == 0


def test_n_zero_is_a_no_op():
    with gps.ScanHandle(gps.ALGO_CSDL, 1024) as h:
        d = sentinel_u32(64)
        h.scan_u32x4(d, d, 0, True)
        sync()
        assert np.all(from_dev_u32(d) == 0xDEADBEEF)


def test_argument_errors():
    d = torch.zeros(4 * 4096 + 4, dtype=torch.int32, device="cuda")
    with gps.ScanHandle(gps.ALGO_CSDL, 4096) as h:
        with pytest.raises(gps.GpsError, match="16-byte aligned"):
            h.scan_u32x4(d.data_ptr() + 4, d, 100, True)
        with pytest.raises(gps.GpsError, match="max_n"):
            h.scan_u32x4(d, d, 4097, True)
    with gps.ScanHandle(gps.ALGO_RTS, 4096) as h:
        with pytest.raises(gps.GpsError, match="CSDL"):
            h.scan_u32x4(d, d, 100, True)
