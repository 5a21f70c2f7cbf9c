"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle) are reproduced by the oracle and
agree with independent arithmetic (numpy uint64 cumsum, closed forms,
math.fsum-checked FP64), so neither the fixtures nor the oracle can drift
unnoticed."""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
U32 = ["u32_small_values", "u32_full_range", "u32_ones"]


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def cumsum_u32(x, inclusive):
    c = np.cumsum(x.astype(np.uint64)) & 0xFFFFFFFF
    return c.astype(np.uint32) if inclusive else np.concatenate([[0], c[:-1]]).astype(np.uint32)


@pytest.mark.parametrize("name", U32)
def test_u32_fixtures(name):
    f = load(name)
    for inclusive, key in ((True, "incl"), (False, "excl")):
        assert np.array_equal(oracle.scan_u32(f["x"], inclusive), f[key])
        assert np.array_equal(cumsum_u32(f["x"], inclusive), f[key])
    if name == "u32_ones":
        assert np.array_equal(f["incl"], np.arange(1, f["x"].size + 1, dtype=np.uint32))


def test_segmented_fixture():
    f = load("segmented_u32")
    off = f["offsets"]
    for inclusive, key in ((True, "incl"), (False, "excl")):
        assert np.array_equal(oracle.segmented_scan_u32(f["x"], off, inclusive), f[key])
        want = np.concatenate([cumsum_u32(f["x"][a:b], inclusive) for a, b in zip(off[:-1], off[1:])])
        assert np.array_equal(want, f[key])


def test_struct_fixture():
    f = load("struct_u32x4")
    for inclusive, key in ((True, "incl"), (False, "excl")):
        assert np.array_equal(oracle.scan_u32x4(f["x"], inclusive), f[key])
        for m in range(4):
            assert np.array_equal(cumsum_u32(f["x"][:, m], inclusive), f[key][:, m])


@pytest.mark.parametrize("name", ["f32_normal", "bf16_logmag"])
def test_fp_fixtures(name):
    f = load(name)
    x = f["x"]
    if name == "f32_normal":
        y, s = oracle.scan_f32_f64(x), oracle.abs_scan_f32_f64(x)
        xv = x.astype(np.float64)
    else:
        y, s = oracle.scan_bf16_f64(x), oracle.abs_scan_bf16_f64(x)
        xv = oracle.bf16_bits_to_f32(x).astype(np.float64)
    assert np.array_equal(y, f["y64"]) and np.array_equal(s, f["s64"])
    for i in (0, 1, x.size // 3, x.size - 1):
        assert f["y64"][i] == pytest.approx(math.fsum(xv[:i + 1]), rel=1e-15, abs=1e-12)
