"""CPU model of the multi-GPU sharding scheme for the u32 scan (a design, not
built on several GPUs): shard g scans [g*n/G, (g+1)*n/G) with an initial
carry equal to the exclusive prefix of the shard totals, which the shards
exchange as G 32-bit values. The concatenated shard scans must equal the
single-device scan, bit for bit, mod 2^32; the FP variant combines the
exchanged totals with TwoSum and must stay within one FP32 rounding.
"""
import numpy as np
import pytest

import oracle


def sharded_scan_u32(x, G, inclusive):
    shards = np.array_split(x, G)
    totals = np.array([int(s.astype(np.uint64).sum() % (1 << 32)) for s in shards], dtype=np.uint64)
    carries = (np.cumsum(totals) - totals) % (1 << 32)  # exclusive prefix of the totals ("all-gather")
    outs = [(oracle.scan_u32(s, inclusive).astype(np.uint64) + c) % (1 << 32) for s, c in zip(shards, carries)]
    return np.concatenate(outs).astype(np.uint32)


@pytest.mark.parametrize("G", [1, 2, 3, 8])
@pytest.mark.parametrize("inclusive", [True, False])
def test_sharded_u32_scan_equals_single_device(G, inclusive):
    for n in (1, 7, 4096, 100_003):
        x = oracle.gen_u32(40 + n, n, 32)
        assert np.array_equal(sharded_scan_u32(x, G, inclusive), oracle.scan_u32(x, inclusive))


def test_sharded_fp_scan_with_twosum_exchange():
    def two_sum(a, b):
        s = np.float32(a + b)
        bb = np.float32(s - a)
        return s, np.float32((a - np.float32(s - bb)) + (b - bb))

    n, G = 1 << 16, 8
    x = oracle.gen_f32_normal(4, n)
    y64 = oracle.scan_f32_f64(x)
    shards = np.array_split(x, G)
    out, cs, cc = [], np.float32(0), np.float32(0)
    for s in shards:
        local = oracle.scan_f32_f64(s)  # each shard's compensated scan, modelled in FP64
        out.append(np.float32(cs) + (np.float64(cc) + local))
        t = np.float32(local[-1])
        cs, e = two_sum(cs, t)
        cc = np.float32(cc + e + np.float32(local[-1] - np.float64(t)))
    y = np.concatenate(out).astype(np.float32).astype(np.float64)
    rel = np.abs(y - y64) / np.maximum(np.abs(y64), 1.0)
    assert rel.max() <= 2.0 ** -23
