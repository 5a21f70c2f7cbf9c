"""Pins the CPU oracle (CPU only; no GPU).

Each check is independent of the oracle's own loop: closed forms from the
reference's tests, numpy uint64 cumsum brute force, algebraic invariants,
published generator test vectors, and exhaustive/rounding-rule checks.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = []


def brute_scan(x: np.ndarray, inclusive: bool) -> np.ndarray:
    c = np.cumsum(x.astype(np.uint64), dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    if inclusive:
        return c.astype(np.uint32)
    out = np.zeros_like(x, dtype=np.uint32)
    out[1:] = c[:-1].astype(np.uint32)
    return out


# ---- uint32 scan ---------------------------------------------------------

@pytest.mark.parametrize("inclusive", [True, False])
def test_ones_closed_form(inclusive):
    # GPUPrefixSumsCUDA/UtilityKernels.cuh:22-38: ones -> i+1 / i.
    n = 10_000
    s = oracle.scan_u32(np.ones(n, np.uint32), inclusive)
    expect = np.arange(n, dtype=np.uint64) + (1 if inclusive else 0)
    assert np.array_equal(s, expect.astype(np.uint32))
    assert oracle.count_ones_errors(s, inclusive) == 0
    s[1234] += 1
    assert oracle.count_ones_errors(s, inclusive) == 1


@pytest.mark.parametrize("inclusive", [True, False])
def test_all_max_wraps(inclusive):
    # all 0xFFFFFFFF: inclusive[i] = (i+1)(2^32-1) = -(i+1) mod 2^32.
    n = 5000
    s = oracle.scan_u32(np.full(n, 0xFFFFFFFF, np.uint32), inclusive)
    k = np.arange(n, dtype=np.uint64) + (1 if inclusive else 0)
    expect = ((np.uint64(1 << 32) - k) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(s, expect)


@pytest.mark.parametrize("inclusive", [True, False])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 17, 64, 4095, 4096, 4097, 100_003])
def test_brute_force_random(n, inclusive):
    x = oracle.gen_u32(7 + n, n, 32)
    assert np.array_equal(oracle.scan_u32(x, inclusive), brute_scan(x, inclusive))


def test_struct_scan_known_answer_and_members():
    # validate.wgsl:51-60: all-ones structs scan to (i >> 2) + 1 per uint32.
    n = 3000
    s = oracle.scan_u32x4(np.ones(4 * n, np.uint32), True).reshape(-1)
    assert np.array_equal(s, (np.arange(4 * n, dtype=np.uint32) >> 2) + 1)
    # Members stay independent: a struct stream whose members differ matches a
    # brute-force scan of each column, and a transposed (row-major) reading does not.
    x = oracle.gen_u32(11, 4 * 513, 32).reshape(-1, 4)
    got = oracle.scan_u32x4(x, False)
    for m in range(4):
        assert np.array_equal(got[:, m], brute_scan(x[:, m], False))
    assert not np.array_equal(got.reshape(-1), brute_scan(x.reshape(-1), False))


def test_invariants_full_range():
    n = 1 << 16
    x = oracle.gen_u32(2, n, 32)
    inc = oracle.scan_u32(x, True)
    exc = oracle.scan_u32(x, False)
    assert exc[0] == 0
    assert np.array_equal(exc[1:], inc[:-1])
    assert int(inc[-1]) == int(x.astype(np.uint64).sum()) % (1 << 32)
    # linearity mod 2^32: scan(a + b) = scan(a) + scan(b), scan(3a) = 3 scan(a)
    y = oracle.gen_u32(3, n, 32)
    assert np.array_equal(oracle.scan_u32(x + y, True), inc + oracle.scan_u32(y, True))
    assert np.array_equal(oracle.scan_u32(x * np.uint32(3), True), inc * np.uint32(3))


def test_validator_counts_mismatches():
    # Utility.hlsl:99-125 restated: every wrong element counts once.
    x = oracle.gen_u32(5, 3000, 10)
    for inclusive in (True, False):
        s = oracle.scan_u32(x, inclusive)
        assert oracle.validate_u32(x, s, inclusive) == 0
        s[0] ^= 1
        s[2999] += 7
        assert oracle.validate_u32(x, s, inclusive) == 2


def test_validator_catches_planted_off_by_one():
    # A plausible bug (exclusive output where inclusive is expected) fails.
    x = oracle.gen_u32(9, 2048, 8)
    assert oracle.validate_u32(x, oracle.scan_u32(x, False), True) > 0


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_multithreaded_baseline_matches(threads):
    x = oracle.gen_u32(11, 1_000_003, 32)
    for inclusive in (True, False):
        assert np.array_equal(oracle.scan_u32_mt(x, inclusive, threads), oracle.scan_u32(x, inclusive))


# ---- segmented -------------------------------------------------------------

def test_segmented_matches_per_segment_brute_force():
    off = oracle.gen_lognormal_offsets(3, 200, median=64.0, sigma=1.5, lo=1, hi=4096)
    n = int(off[-1])
    x = oracle.gen_u32(3, n, 10)
    for inclusive in (True, False):
        s = oracle.segmented_scan_u32(x, off, inclusive)
        for a, b in zip(off[:-1], off[1:]):
            a, b = int(a), int(b)
            assert np.array_equal(s[a:b], brute_scan(x[a:b], inclusive))
        if not inclusive:
            assert np.all(s[off[:-1].astype(np.int64)] == 0)


def test_segmented_empty_segments_and_single_segment():
    x = oracle.gen_u32(4, 100, 32)
    off = np.array([0, 0, 10, 10, 10, 55, 100, 100], np.uint64)
    s = oracle.segmented_scan_u32(x, off, False)
    assert np.array_equal(s[:10], brute_scan(x[:10], False))
    assert np.array_equal(s[10:55], brute_scan(x[10:55], False))
    assert np.array_equal(s[55:], brute_scan(x[55:], False))
    one = oracle.segmented_scan_u32(x, np.array([0, 100], np.uint64), True)
    assert np.array_equal(one, oracle.scan_u32(x, True))
    with pytest.raises(ValueError):
        oracle.segmented_scan_u32(x, np.array([0, 60, 50, 100], np.uint64), True)


# ---- generators ------------------------------------------------------------

def test_splitmix64_published_vectors():
    # Vigna's splitmix64.c: seed 1234567 yields this sequence.
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423,
              4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64_at(1234567, i) for i in range(5)] == expect
    assert oracle.splitmix64_at(0, 0) == 0xE220A8397B1DCDAF


def test_gen_u32_bits_ranges():
    for bits, hi in ((8, 255), (10, 1023), (32, 0xFFFFFFFF)):
        x = oracle.gen_u32(1, 1 << 16, bits)
        assert int(x.max()) <= hi
        assert int(x.max()) > hi * 0.99  # covers the range
    x = oracle.gen_u32(1, 8, 32)
    assert [int(v) for v in x] == [oracle.splitmix64_at(1, i) >> 32 for i in range(8)]
    # offset streams are windows of the same stream
    assert np.array_equal(oracle.gen_u32(1, 5, 32, first=3), x[3:8])


def test_normal_generator_moments():
    z = np.array([oracle.normal_at(4, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    f = oracle.gen_f32_normal(4, 1000)
    assert np.array_equal(f, np.array([oracle.normal_at(4, i) for i in range(1000)], np.float32))


def test_lognormal_offsets_distribution():
    off = oracle.gen_lognormal_offsets(3, 4096)
    lengths = np.diff(off.astype(np.int64))
    assert lengths.min() >= 1 and lengths.max() <= 1 << 22
    # median 4096 (log-normal median = e^mu); sample median within 10%
    assert 0.9 * 4096 < np.median(lengths) < 1.1 * 4096
    # expected total ~ 4096 * 4096 * e^(1.5^2/2) ~ 5.2e7 (not 2^27)
    assert 3e7 < int(off[-1]) < 9e7


def test_d3d12_taus_lcg_part():
    # With seed 0 every Tausworthe component starts at 0 and stays 0; the LCG
    # component of thread t starts at 0 as well, so the value is
    # (0 * 1664525 + 1013904223) & 1023 for every element (Numerical Recipes
    # LCG first output from 0 is 1013904223).
    x = oracle.gen_d3d12_taus(0, 70000)
    assert np.all(x == (1013904223 & 1023))
    y = oracle.gen_d3d12_taus(4000, 4000)
    assert y.max() <= 1023 and len(np.unique(y)) > 900
    # stride-65536 quirk: element i and i + 65536 come from the same thread
    z = oracle.gen_d3d12_taus(77, 65536 + 10)
    assert np.array_equal(z[:10], z[65536:])


def test_unity_taus_known_values():
    # Seed 0 (what the reference's Unity tests run): the Tausworthe components
    # stay 0 and every thread's LCG starts at 0, so uint4 element v holds the
    # bytes, low first, of the (v // 65536 + 1)-th output of the Numerical
    # Recipes LCG from 0: 1013904223 = 0x3C6EF35F, then 1196435762 = 0x47502932.
    x = oracle.gen_unity_taus(0, 4 * 65536 + 8)
    assert list(x[:4]) == [0x5F, 0xF3, 0x6E, 0x3C]
    assert np.all(x[:4 * 65536].reshape(-1, 4) == x[:4])  # every thread alike in the first round
    assert list(x[4 * 65536:4 * 65536 + 4]) == [0x32, 0x29, 0x50, 0x47]
    # Seed 1, thread 0: states (0, 1, 2, 3); one Tausworthe step maps 0, 1
    # and 2 to 0 (the shifted-out bits), so the output is the LCG of 3:
    # 3 * 1664525 + 1013904223 = 1018897798 = 0x3CBB2586 (bytes low first).
    y = oracle.gen_unity_taus(1, 8)
    assert list(y[:4]) == [0x86, 0x25, 0xBB, 0x3C]
    # per-thread state persists: thread 0's second element differs from its first
    z = oracle.gen_unity_taus(12345, 4 * 65536 + 4)
    assert not np.array_equal(z[:4], z[4 * 65536:]) and z.max() <= 255
    # partial last uint4: the first n values of the longer stream
    assert np.array_equal(oracle.gen_unity_taus(9, 4099), oracle.gen_unity_taus(9, 4100)[:4099])


# ---- floating point --------------------------------------------------------

def test_bf16_roundtrip_and_rne():
    # Every bf16 pattern converts exactly and back.
    for b in list(range(0, 0x7F80, 97)) + [0x0000, 0x8000, 0x3F80, 0xBF80, 0x7F7F, 0x0001]:
        f = oracle.bf16_to_f32(b)
        assert oracle.f32_to_bf16(f) == b
    # Ties to even: 1 + 2^-8 is halfway between bf16 1.0 and 1.0078125.
    assert oracle.f32_to_bf16(1.0 + 2.0 ** -8) == 0x3F80
    assert oracle.f32_to_bf16(1.0 + 3 * 2.0 ** -8) == 0x3F82  # 1.0117 -> even 1.015625
    assert oracle.f32_to_bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 0x3F81  # above the tie
    assert oracle.f32_to_bf16(float("inf")) == 0x7F80
    assert oracle.f32_to_bf16(3.4e38) == 0x7F80  # rounds past the max finite bf16
    nan = oracle.f32_to_bf16(float("nan"))
    assert (nan & 0x7F80) == 0x7F80 and (nan & 0x7F)


def test_bf16_rne_matches_torch():
    torch = pytest.importorskip("torch")
    f = np.concatenate([oracle.gen_f32_normal(8, 5000) * 1000, oracle.gen_f32_normal(9, 5000) * 1e-3])
    ours = np.array([oracle.f32_to_bf16(float(v)) for v in f], np.uint16)
    theirs = torch.from_numpy(f).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, theirs)


def test_fp64_scan_against_fsum():
    x = oracle.gen_f32_normal(4, 3000)
    y = oracle.scan_f32_f64(x)
    for i in (0, 1, 17, 999, 2999):
        assert y[i] == math.fsum(float(v) for v in x[: i + 1])
    bits = oracle.gen_bf16_logmag(5, 2000)
    yb = oracle.scan_bf16_f64(bits)
    vals = oracle.bf16_bits_to_f32(bits).astype(np.float64)
    for i in (0, 5, 1999):
        assert abs(yb[i] - math.fsum(vals[: i + 1])) <= 1e-12 * math.fsum(abs(vals[: i + 1]))
    assert np.allclose(oracle.abs_scan_f32_f64(x), np.cumsum(np.abs(x.astype(np.float64))))


def test_segmented_fp64_scan_against_fsum_per_segment():
    x = oracle.gen_f32_normal(8, 500)
    off = np.array([0, 0, 1, 1, 40, 41, 300, 500, 500], dtype=np.uint64)  # empty and length-1 segments
    y = oracle.segmented_scan_fp_f64(x, off)
    a_ = oracle.segmented_scan_fp_f64(x, off, absolute=True)
    for s in range(off.size - 1):
        a, b = int(off[s]), int(off[s + 1])
        for i in range(a, b):
            assert y[i] == math.fsum(float(v) for v in x[a:i + 1])
            assert a_[i] == pytest.approx(math.fsum(abs(float(v)) for v in x[a:i + 1]), rel=1e-15)
    bits = oracle.gen_bf16_logmag(5, 300)
    yb = oracle.segmented_scan_fp_f64(bits, np.array([0, 150, 300], dtype=np.uint64), bf16=True)
    assert yb[150] == float(oracle.bf16_bits_to_f32(bits[150:151])[0])  # restarts at the head
    with pytest.raises(ValueError):
        oracle.segmented_scan_fp_f64(x, np.array([0, 300, 200, 500], dtype=np.uint64))


def test_bf16_generator_magnitudes():
    bits = oracle.gen_bf16_logmag(5, 20000)
    v = np.abs(oracle.bf16_bits_to_f32(bits).astype(np.float64))
    assert v.min() >= 0.99e-4 and v.max() <= 1.01e4
    lg = np.log10(v)
    assert abs(lg.mean()) < 0.1  # U(-4, 4) in the exponent
    s = np.sign(oracle.bf16_bits_to_f32(bits))
    assert abs(s.mean()) < 0.05


# ---- MLP cost-model oracle (FP64) -----------------------------------------

def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp
    D, H, B = 7, 9, 13
    rng = np.random.default_rng(0)
    p = mlp.init_params(D, H, 1) + rng.normal(0, 0.05, mlp.param_count(D, H))
    x = rng.normal(size=(B, D))
    t = rng.normal(size=B)
    L, g = mlp.loss_grad(p, x, t, D, H)
    assert abs(L - mlp.loss(p, x, t, D, H)) < 1e-12
    h = 1e-6
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, x, t, D, H) - mlp.loss(p - e, x, t, D, H)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_oracle_layout():
    from oracle import mlp
    D, H = 3, 4
    assert mlp.param_count(D, H) == 4 * 3 + 4 + 16 + 4 + 4 + 1
    p = np.arange(mlp.param_count(D, H), dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, D, H)
    assert W1[1, 2] == 5 and b1[0] == 12 and W2[0, 0] == 16 and b2[3] == 35 and W3[0] == 36 and b3 == 40
