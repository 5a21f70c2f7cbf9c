"""uint32 scans on the GPU, bit-exact against the CPU oracle (through the C ABI).

Covers the reference's own tests (all-ones sweeps over one to two tiles,
GPUPrefixSumsCUDA/ChainedScanDecoupledLookbackDispatcher.cuh:38-114; random
[0,1023] inputs with seed = n over n in [3072, 6144),
GPUPrefixSumsD3D12/GPUPrefixSumBase.h:135-161) and what they leave out:
full-range wraparound, exclusive decoupled fallback, stall periods 7 and 8
(and 2, 3 on the persistent schedule), n % 4 tails, unaligned and in-place buffers, every tile shape, 2^28 elements.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuprefixsums_amd as gps  # noqa: E402
from tests.gpu_util import (empty_u32, first_mismatch, from_dev_u32, sentinel_u32, sync,  # noqa: E402
                            to_dev)

ALGOS = {
    "csdl": dict(algo=gps.ALGO_CSDL),  # persistent pipelined schedule (default)
    "csdl_tile": dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_TILE),  # one workgroup per tile
    "csdl_control": dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_CONTROL),  # persistent + control wave
    "rts": dict(algo=gps.ALGO_RTS),
    "csdldf": dict(algo=gps.ALGO_CSDLDF),
    "csdldf_stall7": dict(algo=gps.ALGO_CSDLDF, stall_period=7),
    "csdldf_stall8": dict(algo=gps.ALGO_CSDLDF, stall_period=8),
    # decoupled fallback on the persistent pipelined schedule
    "csdldf_pp": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_PERSISTENT),
    "csdldf_pp_stall7": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_PERSISTENT, stall_period=7),
    "csdldf_pp_stall8": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_PERSISTENT, stall_period=8),
    # period 2 keeps the plain order on the persistent schedule; from 3 a
    # stalled workgroup prepublishes its next tile (3: the next tile stalls
    # too more often, and the tile before the last one stalls at small sizes)
    "csdldf_pp_stall2": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_PERSISTENT, stall_period=2),
    "csdldf_pp_stall3": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_PERSISTENT, stall_period=3),
    # the D3D12 emulation: tiles % P == 0 deferred to a second launch, so
    # their successors' fallbacks hit tiles that have not started
    "csdldf_split2": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_SPLIT, stall_period=2),
    "csdldf_split8": dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_SPLIT, stall_period=8),
}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def run_scan(h, x: np.ndarray, inclusive: bool) -> np.ndarray:
    n = x.size
    d_in = to_dev(x) if n else empty_u32(1)
    d_out = sentinel_u32(n + 64)  # guard region after n must stay untouched
    h.scan(d_in, d_out, n, inclusive)
    sync()
    out = from_dev_u32(d_out)
    assert np.all(out[n:n + 64] == 0xDEADBEEF), "wrote past n"
    return out[:n]


def check(h, x, inclusive, label=""):
    got = run_scan(h, x, inclusive)
    want = oracle.scan_u32(x, inclusive)
    assert np.array_equal(got, want), f"{label} n={x.size} incl={inclusive}: {first_mismatch(got, want)}"


def sweep_sizes(tile: int):
    s = {1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025}
    for k in (1, 2, 3):
        s.update(range(k * tile - 5, k * tile + 6))
    s.update({10 * tile + 3, 64 * tile + 1, 100 * tile - 1, (1 << 20) - 1, 1 << 20, (1 << 20) + 3})
    return sorted(v for v in s if v > 0)


@pytest.mark.parametrize("name", list(ALGOS))
@pytest.mark.parametrize("inclusive", [True, False])
def test_size_sweep_full_range(name, inclusive):
    with gps.ScanHandle(max_n=1 << 23, **ALGOS[name]) as h:
        for n in sweep_sizes(h.tile):
            x = oracle.gen_u32(1000 + n, n, 32)
            check(h, x, inclusive, name)
        assert h.device_error() == 0


@pytest.mark.parametrize("inclusive", [True, False])
def test_config1_n2p20_values_0_255(inclusive):
    # n = 2^20, uniform [0, 255], seed 1; CSDL against the host scan.
    x = oracle.gen_u32(1, 1 << 20, 8)
    with gps.ScanHandle(gps.ALGO_CSDL, 1 << 20) as h:
        check(h, x, inclusive, "csdl")
        assert h.device_error() == 0


@pytest.mark.parametrize("name", list(ALGOS))
def test_reference_d3d12_random_sweep(name):
    # GPUPrefixSumBase.h:141-148: random [0,1023] (Utility.hlsl InitRandom) with
    # seed = n for n in [3072, 6144); every 7th size here to bound the time.
    with gps.ScanHandle(max_n=8192, **ALGOS[name]) as h:
        for n in range(3072, 6144, 7):
            x = oracle.gen_d3d12_taus(n, n)
            for inclusive in (True, False):
                check(h, x, inclusive, name)


@pytest.mark.parametrize("name", list(ALGOS))
@pytest.mark.parametrize("inclusive", [True, False])
def test_reference_ones_sweep_one_to_two_tiles(name, inclusive):
    # ...Dispatcher.cuh:48,88: all-ones input for every n in [T, 2T], checked
    # with the closed form on the device (UtilityKernels.cuh:22-38).
    with gps.ScanHandle(max_n=1 << 16, **ALGOS[name]) as h:
        T = h.tile
        buf = empty_u32(2 * T)
        fails = []
        for n in range(T, 2 * T + 1):
            gps.fill_u32(buf, n, 1)
            h.scan(buf, buf, n, inclusive)
            if gps.validate_ones_u32(buf, n, inclusive):
                fails.append(n)
        assert not fails, f"failed sizes {fails[:10]}"


@pytest.mark.parametrize("block,vec", [(b, v) for b in (256, 512, 1024) for v in (2, 4, 8)])
@pytest.mark.parametrize("name", ["csdl", "csdl_tile", "csdl_control", "rts", "csdldf_stall7", "csdldf_pp_stall7"])
def test_every_tile_shape(block, vec, name):
    with gps.ScanHandle(max_n=1 << 20, block_threads=block, vec_per_thread=vec, **ALGOS[name]) as h:
        T = h.tile
        assert T == (block - 64 if name == "csdl_control" else block) * vec * 4
        for n in (1, T - 1, T, T + 1, 3 * T + 5, (1 << 20) - 3):
            x = oracle.gen_u32(n * 3 + block + vec, n, 32)
            check(h, x, True, f"{block}x{vec}")
            check(h, x, False, f"{block}x{vec}")
        assert h.device_error() == 0


@pytest.mark.parametrize("name", list(ALGOS))
@pytest.mark.parametrize("inclusive", [True, False])
def test_in_place(name, inclusive):
    n = (1 << 20) + 13
    x = oracle.gen_u32(77, n, 32)
    with gps.ScanHandle(max_n=n, **ALGOS[name]) as h:
        d = to_dev(x)
        h.scan(d, d, n, inclusive)
        sync()
        got = from_dev_u32(d)
        want = oracle.scan_u32(x, inclusive)
        assert np.array_equal(got, want), first_mismatch(got, want)


@pytest.mark.parametrize("name", ["csdl", "rts", "csdldf_stall8"])
@pytest.mark.parametrize("shift_in,shift_out", [(1, 0), (0, 3), (2, 1), (3, 3)])
def test_unaligned_buffers(name, shift_in, shift_out):
    # 4-byte aligned, not 16-byte aligned: the element-wise path.
    n = 50_001
    x = oracle.gen_u32(5, n, 32)
    with gps.ScanHandle(max_n=n, **ALGOS[name]) as h:
        base_in = to_dev(np.concatenate([np.zeros(shift_in, np.uint32), x]))
        base_out = sentinel_u32(n + shift_out + 16)
        d_in = base_in.data_ptr() + 4 * shift_in
        d_out = base_out.data_ptr() + 4 * shift_out
        for inclusive in (True, False):
            h.scan(d_in, d_out, n, inclusive)
            sync()
            out = from_dev_u32(base_out)
            got = out[shift_out:shift_out + n]
            want = oracle.scan_u32(x, inclusive)
            assert np.array_equal(got, want), first_mismatch(got, want)
            assert np.all(out[:shift_out] == 0xDEADBEEF) and np.all(out[shift_out + n:] == 0xDEADBEEF)


@pytest.mark.parametrize("name", list(ALGOS))
def test_all_max_values_wrap(name):
    n = 300_007
    x = np.full(n, 0xFFFFFFFF, np.uint32)
    with gps.ScanHandle(max_n=n, **ALGOS[name]) as h:
        got = run_scan(h, x, True)
        k = np.arange(1, n + 1, dtype=np.uint64)
        assert np.array_equal(got, ((np.uint64(1 << 32) - k) & np.uint64(0xFFFFFFFF)).astype(np.uint32))


@pytest.mark.parametrize("inclusive", [True, False])
def test_default_handle_across_size_thresholds(inclusive):
    # A default CSDL handle picks shape and schedule from n: one workgroup per
    # tile at 1024x2 up to 2^18, 1024x4 up to 2^21, 1024x8 below 2^24, the
    # persistent 1024x8 kernel from 2^24. Sizes on both sides of each switch.
    sizes = [1, 5, (1 << 18) - 1, 1 << 18, (1 << 18) + 1, (1 << 21) - 3, 1 << 21, (1 << 21) + 5,
             (1 << 24) - 1, 1 << 24, (1 << 24) + 7]
    with gps.ScanHandle(gps.ALGO_CSDL, sizes[-1]) as h:
        assert (h.block_threads, h.vec_per_thread, h.schedule) == (1024, 8, gps.SCHEDULE_PERSISTENT)
        for n in sizes:
            check(h, oracle.gen_u32(n, n, 32), inclusive, f"default handle n={n}")
        assert h.device_error() == 0


@pytest.mark.parametrize("schedule", ["default", "tile", "persistent"])
@pytest.mark.parametrize("inclusive", [True, False])
def test_scan_continued_from_a_carry(schedule, inclusive):
    sched = {"default": 0, "tile": gps.SCHEDULE_TILE, "persistent": gps.SCHEDULE_PERSISTENT}[schedule]
    n = (1 << 22) + 123
    x = oracle.gen_u32(77, n, 32)
    want = oracle.scan_u32(x, inclusive).astype(np.uint64)
    total = int(x.astype(np.uint64).sum())
    C = 0x9E3779B9
    with gps.ScanHandle(gps.ALGO_CSDL, n, schedule=sched) as h:
        d_in = to_dev(x)
        # one call: outputs shifted by the carry, carry_out = total + carry
        d_out = sentinel_u32(n)
        c_in = to_dev(np.array([C], np.uint32))
        c_out = sentinel_u32(1)
        h.scan_carry(d_in, d_out, n, inclusive, c_in, c_out)
        sync()
        got = from_dev_u32(d_out)
        exp = ((want + C) % (1 << 32)).astype(np.uint32)
        assert np.array_equal(got, exp), first_mismatch(got, exp)
        assert int(from_dev_u32(c_out)[0]) == (total + C) % (1 << 32)
        # chunks chained through one carry word (carry_in == carry_out)
        word = to_dev(np.zeros(1, np.uint32))
        d_out = sentinel_u32(n)
        bounds = [0, 1, 4096, 4096 + 65536 + 3, (1 << 20) + 7, (1 << 21), n]
        for a, b in zip(bounds[:-1], bounds[1:]):
            h.scan_carry(d_in[a:b], d_out[a:b], b - a, inclusive, word, word)
        sync()
        got = from_dev_u32(d_out)
        exp = want.astype(np.uint32)  # each chunk starts from the running total
        assert np.array_equal(got, exp), first_mismatch(got, exp)
        assert int(from_dev_u32(word)[0]) == total % (1 << 32)
        assert h.device_error() == 0


def test_carry_edge_cases_and_errors():
    with gps.ScanHandle(gps.ALGO_CSDL, 4096) as h:
        d = sentinel_u32(16)
        c_in = to_dev(np.array([12345], np.uint32))
        c_out = sentinel_u32(1)
        h.scan_carry(d, d, 0, True, c_in, c_out)  # n = 0: carry passes through
        sync()
        assert int(from_dev_u32(c_out)[0]) == 12345
        h.scan_carry(d, d, 0, True, None, c_out)  # no carry-in: 0
        sync()
        assert int(from_dev_u32(c_out)[0]) == 0
    with gps.ScanHandle(gps.ALGO_RTS, 4096) as h:
        with pytest.raises(gps.GpsError, match="carries"):
            h.scan_carry(d, d, 16, True, c_in, None)
    with gps.ScanHandle(gps.ALGO_CSDL, 4096, schedule=gps.SCHEDULE_CONTROL) as h:
        with pytest.raises(gps.GpsError, match="carries"):
            h.scan_carry(d, d, 16, True, None, c_out)


def test_argument_errors():
    from gpuprefixsums_amd._lib import GPS_ERR_INVALID, GPS_ERR_TOO_LARGE
    with gps.ScanHandle(gps.ALGO_CSDL, 1000) as h:
        d = empty_u32(2000)
        h.inclusive(d, d, 0)  # n = 0 is a no-op
        with pytest.raises(gps.GpsError) as e:
            h.inclusive(d, d, 1001)
        assert e.value.code == GPS_ERR_TOO_LARGE
        with pytest.raises(gps.GpsError) as e:
            h.inclusive(d.data_ptr() + 2, d, 10)
        assert e.value.code == GPS_ERR_INVALID
        with pytest.raises(gps.GpsError) as e:
            h.inclusive(0, d, 10)
        assert e.value.code == GPS_ERR_INVALID


def test_many_launches_and_reset():
    # Epoch-tagged status words: thousands of back-to-back launches on one
    # handle with no clearing in between, with an explicit re-arm midway.
    n = 40_000
    x = oracle.gen_u32(8, n, 32)
    want = oracle.scan_u32(x, True)
    with gps.ScanHandle(gps.ALGO_CSDL, n) as h:
        d_in, d_out = to_dev(x), empty_u32(n)
        for i in range(3000):
            h.inclusive(d_in, d_out, n - (i % 5000))
            if i == 1500:
                h.reset()
        h.inclusive(d_in, d_out, n)
        sync()
        assert np.array_equal(from_dev_u32(d_out, n), want)
        assert h.device_error() == 0


def test_fill_random_matches_host_generator():
    n = 100_000
    d = empty_u32(n)
    for bits in (8, 10, 32):
        gps.fill_random_u32(d, n, seed=2, bits=bits)
        sync()
        assert np.array_equal(from_dev_u32(d, n), oracle.gen_u32(2, n, bits))


def test_device_validators():
    n = 70_001
    x = oracle.gen_u32(3, n, 32)
    d_in = to_dev(x)
    for inclusive in (True, False):
        s = oracle.scan_u32(x, inclusive)
        d_s = to_dev(s)
        assert gps.validate_scan_u32(d_in, d_s, n, inclusive) == 0
        s2 = s.copy()
        s2[n // 2] += 1  # one wrong element breaks two differences
        assert gps.validate_scan_u32(d_in, to_dev(s2), n, inclusive) == 2
    ones = to_dev(np.arange(1, n + 1, dtype=np.uint32))
    assert gps.validate_ones_u32(ones, n, True) == 0
    assert gps.validate_ones_u32(ones, n, False) == n


@pytest.mark.parametrize("schedule", ["tile", "persistent"])
@pytest.mark.parametrize("period", [7, 8])
@pytest.mark.parametrize("inclusive", [True, False])
def test_emulated_deadlock_counters(period, inclusive, schedule):
    n = 1 << 22
    x = oracle.gen_u32(21, n, 32)
    sched = {"tile": gps.SCHEDULE_TILE, "persistent": gps.SCHEDULE_PERSISTENT}[schedule]
    with gps.ScanHandle(gps.ALGO_CSDLDF, n, stall_period=period, spin_limit=4, schedule=sched) as h:
        h.reset_counters()
        check(h, x, inclusive, f"stall{period}")
        c = h.counters()
        tiles = -(-n // h.tile)
        stalled = len([t for t in range(0, tiles - 1, period)])
        # every stalled tile is released only by a successor's insertion (on
        # the persistent schedule the tile before the last skips its stall
        # when its own workgroup's next tile is the last one)
        skippable = schedule == "persistent" and (tiles - 2) % period == 0
        assert c["stall_timeouts"] == 0
        assert c["insertions"] >= stalled - int(skippable)
        assert c["fallbacks"] >= c["insertions"]
        assert h.device_error() == 0


@pytest.mark.parametrize("period", [2, 7, 8])
@pytest.mark.parametrize("inclusive", [True, False])
def test_split_schedule_inserts_every_deferred_tile(period, inclusive):
    # GPUPrefixSumsD3D12/EmulatedDeadlock.cpp:97-118: the first launch runs
    # every tile except index % P == 0 and the second launch the deferred
    # ones. A deferred tile with a successor has not started when that
    # successor looks back, so a fallback must reduce it from the input and
    # insert it: insertions >= deferred tiles with a successor, and a fallback
    # onto tile 0 happens in every launch.
    n = (1 << 22) + 9
    x = oracle.gen_u32(31 + period, n, 32)
    with gps.ScanHandle(gps.ALGO_CSDLDF, n, stall_period=period, spin_limit=4, schedule=gps.SCHEDULE_SPLIT) as h:
        for rep in range(2):  # the second scan reuses the status words (next epoch)
            h.reset_counters()
            check(h, x, inclusive, f"split{period} rep{rep}")
            c = h.counters()
            tiles = -(-n // h.tile)
            deferred = len(range(0, tiles - 1, period))
            assert c["insertions"] >= deferred, c
            assert c["stall_timeouts"] == 0
        assert h.device_error() == 0


def test_split_schedule_needs_decoupled_fallback_and_a_period():
    for kw in (dict(algo=gps.ALGO_CSDL, schedule=gps.SCHEDULE_SPLIT),
               dict(algo=gps.ALGO_CSDLDF, schedule=gps.SCHEDULE_SPLIT),
               dict(algo=gps.ALGO_RTS, schedule=gps.SCHEDULE_SPLIT)):
        with pytest.raises(gps.GpsError, match="split"):
            gps.ScanHandle(max_n=4096, **kw)


@pytest.mark.parametrize("spin_limit", [1, 32])
@pytest.mark.parametrize("period", [2, 3])
def test_persistent_fallback_dense_stalls(period, spin_limit):
    # Every second or third tile stalls on the persistent schedule. At period
    # 2 each stalled workgroup holds its next ticket back until a successor
    # has inserted its aggregate; at period 3 it first takes its next tile and
    # publishes that tile's aggregate (unless it stalls too). Either way the
    # grid keeps going only through fallbacks.
    n = (1 << 22) + 77
    x = oracle.gen_u32(period * 100 + spin_limit, n, 32)
    with gps.ScanHandle(gps.ALGO_CSDLDF, n, stall_period=period, spin_limit=spin_limit,
                        schedule=gps.SCHEDULE_PERSISTENT) as h:
        for inclusive in (True, False):
            h.reset_counters()
            check(h, x, inclusive, f"period {period} spin {spin_limit}")
            c = h.counters()
            tiles = -(-n // h.tile)
            stalled = len(range(0, tiles - 1, period))
            skippable = period >= 3 and (tiles - 2) % period == 0  # see test_emulated_deadlock_counters
            assert c["stall_timeouts"] == 0 and c["insertions"] >= stalled - int(skippable)
        assert h.device_error() == 0


@pytest.mark.parametrize("algo", [gps.ALGO_CSDL, gps.ALGO_RTS, gps.ALGO_CSDLDF])
@pytest.mark.parametrize("inclusive", [True, False])
def test_dispatcher_test_all(algo, inclusive):
    # The reference's TestAll* methods, run by the C++ dispatcher classes.
    d = {gps.ALGO_CSDL: gps.ChainedScanDecoupledLookbackDispatcher,
         gps.ALGO_RTS: gps.ReduceThenScanDispatcher,
         gps.ALGO_CSDLDF: gps.EmulatedDeadlockingDispatcher}[algo](1 << 22)
    passed, total = d.test_all_inclusive() if inclusive else d.test_all_exclusive()
    assert total > 4000 and passed == total


@pytest.mark.parametrize("op", ["inclusive", "sum", "exclusive"])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 4096 + 17, (1 << 22) + 5])
def test_wave64_primitives_against_serial_per_wave_scan(n, op):
    """Device unit test of the wave64 DPP primitives (Utils.cuh:22-71
    counterparts): each run of 64 consecutive elements is one wave."""
    x = oracle.gen_u32(77 + n, n, 32)
    d_in, d_out = to_dev(x), sentinel_u32(n + 64)
    code = {"inclusive": gps.WAVE_INCLUSIVE_SCAN, "sum": gps.WAVE_SUM, "exclusive": gps.WAVE_EXCLUSIVE_SCAN}[op]
    gps.wave_primitive_u32(d_in, d_out, n, code)
    sync()
    got = from_dev_u32(d_out)
    assert np.all(got[n:n + 64] == 0xDEADBEEF), "wrote past n"
    runs = np.zeros(-(-n // 64) * 64, dtype=np.uint64)
    runs[:n] = x
    runs = runs.reshape(-1, 64)
    incl = np.cumsum(runs, axis=1) & 0xFFFFFFFF  # serial per-wave scan in uint64, wrapped
    want = {"inclusive": incl, "sum": np.repeat(incl[:, -1:], 64, axis=1),
            "exclusive": (incl - runs) & 0xFFFFFFFF}[op].reshape(-1)[:n].astype(np.uint32)
    assert np.array_equal(got[:n], want), first_mismatch(got[:n], want)
