"""Writes the committed golden fixtures (small inputs and expected outputs)
from the oracle. Run from the repository root:

    python tests/golden/make_golden.py

Each fixture is an .npz (no pickles) holding the input, the expected output
and the parameters. tests/test_golden.py checks that the oracle still
reproduces them; tests/test_gpu_golden.py checks the library against them.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
import oracle  # noqa: E402


def main() -> None:
    out = {}
    # uint32 scans: values in [0, 255], full-range (wrapping) values,
    # and the all-ones known answer, across one to a few tiles.
    for name, seed, n, bits in (("u32_small_values", 11, 9_001, 8), ("u32_full_range", 12, 20_003, 32)):
        x = oracle.gen_u32(seed, n, bits)
        out[name] = dict(x=x, incl=oracle.scan_u32(x, True), excl=oracle.scan_u32(x, False))
    ones = np.ones(17_000, np.uint32)
    out["u32_ones"] = dict(x=ones, incl=oracle.scan_u32(ones, True), excl=oracle.scan_u32(ones, False))
    # batched segmented u32 (exclusive and inclusive), with empty segments
    off = oracle.gen_lognormal_offsets(5, 97, median=100.0, sigma=1.5, lo=0, hi=1 << 12)
    x = oracle.gen_u32(13, int(off[-1]), 10)
    out["segmented_u32"] = dict(x=x, offsets=off, excl=oracle.segmented_scan_u32(x, off, False),
                                incl=oracle.segmented_scan_u32(x, off, True))
    # 4-member structs: random members (inclusive, exclusive)
    xs = oracle.gen_u32(14, 4 * 5_003, 32).reshape(-1, 4)
    out["struct_u32x4"] = dict(x=xs, incl=oracle.scan_u32x4(xs, True), excl=oracle.scan_u32x4(xs, False))
    # FP32 N(0,1) and BF16 log-magnitude inputs with their FP64 prefixes
    xf = oracle.gen_f32_normal(15, 9_000)
    out["f32_normal"] = dict(x=xf, y64=oracle.scan_f32_f64(xf), s64=oracle.abs_scan_f32_f64(xf))
    xb = oracle.gen_bf16_logmag(16, 9_000)
    out["bf16_logmag"] = dict(x=xb, y64=oracle.scan_bf16_f64(xb), s64=oracle.abs_scan_bf16_f64(xb))
    for name, arrays in out.items():
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
