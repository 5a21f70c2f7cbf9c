"""Host-side logic of the tuner (no GPU): the workload sweep, candidate
space, feature encoding, and the regret arithmetic used to judge the model.
"""
import math

import numpy as np
import pytest

from gpuprefixsums_amd.tuner.evaluate import best_static, regrets
from gpuprefixsums_amd.tuner.space import (FEATURES, KINDS, STALL_CONFIGS, Config, Workload, c5_workloads, candidates,
                                           default_config, features, gated_choice, lognormal_partition, split,
                                           stall_workloads)


def test_sweep_shape_and_split():
    ws = c5_workloads()
    assert len(ws) == 512
    train, held = split(ws)
    assert len(train) == 384 and len(held) == 128
    assert [w.wid for w in ws] == list(range(512))
    for w in ws:
        assert 2 ** 10 - 1 <= w.n <= 2 ** 30 + 1
        assert 1 <= w.segments <= 2 ** 14
        assert w.dtype in ("u32", "f32", "bf16")
        assert w.offsets[0] == 0 and int(w.offsets[-1]) == w.n
        assert np.all(np.diff(w.offsets.astype(np.int64)) >= 0)
    # seeded: the same sweep every time
    again = c5_workloads()
    assert all(a.n == b.n and a.dtype == b.dtype and np.array_equal(a.offsets, b.offsets) for a, b in zip(ws, again))


def test_lognormal_partition_sums_exactly():
    rng = np.random.default_rng(1)
    for n, s in [(1, 1), (5, 9), (1000, 7), (12345, 4096)]:
        off = lognormal_partition(n, s, rng)
        assert off.size == s + 1 and int(off[-1]) == n


def test_config_names_round_trip_and_candidates():
    for dt in ("u32", "f32", "bf16"):
        cs = candidates(dt)
        assert len({c.name for c in cs}) == len(cs)
        for c in cs:
            assert Config.parse(c.name) == c
        for n in (1 << 10, 1 << 18, (1 << 18) + 1, 1 << 20, 1 << 22, (1 << 22) + 1, 1 << 30):
            assert default_config(dt, 1, n).name in {c.name for c in cs}
            assert default_config(dt, 100, n).name in {c.name for c in cs}
    assert {c.kind for c in candidates("f32")} == {"loop_fp", "batched_fp"}
    assert default_config("bf16", 100).kind == "batched_fp" and default_config("u32", 100).kind == "batched"
    # the library's size-adaptive defaults
    assert [default_config("u32", 9, n).name for n in (1 << 19, 1 << 21, 1 << 22)] == [
        "batched:1024x2", "batched:1024x4", "batched:1024x8"]
    assert [default_config("u32", 1, n).name for n in (1 << 18, (1 << 18) + 1, 1 << 21, (1 << 24) - 1, 1 << 24)] == [
        "loop_csdl:1024x2:tile", "loop_csdl:1024x4:tile", "loop_csdl:1024x4:tile", "loop_csdl:1024x8:tile",
        "loop_csdl:1024x8"]
    assert [default_config("f32", 1, n).name for n in (1 << 18, 1 << 22, (1 << 22) + 1, 1 << 24, (1 << 24) + 1)] == [
        "loop_fp:256x4", "loop_fp:1024x4", "loop_fp:1024x8", "loop_fp:1024x8", "loop_fp:1024x8:pp"]
    assert [default_config("bf16", 9, n).name for n in ((1 << 20) - 1, 1 << 20, 1 << 23)] == [
        "batched_fp:256x4", "batched_fp:1024x4", "batched_fp:1024x8:pp"]
    assert [default_config("f32", 9, n).name for n in (1 << 22, (1 << 22) + 1)] == [
        "batched_fp:1024x4", "batched_fp:1024x8:pp"]
    assert Config.parse("loop_fp:1024x8:pp") == Config("loop_fp", 1024, 8, persistent=True)


def test_features_dimension_and_values():
    w = Workload(0, 1 << 20, "bf16", 4, np.array([0, 1, 2, 3, 1 << 20], dtype=np.uint64))
    f = features(w, Config("loop_csdldf", 1024, 4, 32))
    assert len(f) == len(FEATURES) == 8 + len(KINDS) + 4
    d = dict(zip(FEATURES, f))
    assert d["log2_n"] == 20 and d["log2_segments"] == 2
    assert d["is_bf16"] == 1 and d["is_u32"] == 0 and d["log2_stall"] == 0
    assert d["kind_loop_csdldf"] == 1 and d["log2_block"] == 10 and d["log2_spin"] == math.log2(33)
    assert d["persistent"] == 0 and dict(zip(FEATURES, features(w, Config("loop_fp", 1024, 8, persistent=True))))[
        "persistent"] == 1
    # the u32 kernels run persistent unless marked ":tile"
    assert dict(zip(FEATURES, features(w, Config("loop_csdl", 1024, 8))))["persistent"] == 1
    assert dict(zip(FEATURES, features(w, Config("loop_csdl", 1024, 8, tile=True))))["persistent"] == 0
    ws = stall_workloads()[0]
    assert dict(zip(FEATURES, features(ws, STALL_CONFIGS[0])))["log2_stall"] == math.log2(ws.stall_period + 1)


def test_stall_sweep_and_its_candidates():
    ws = stall_workloads()
    assert len(ws) == 96 and ws[0].wid == 10000
    assert {w.stall_period for w in ws} == {7, 8}
    assert all(2 ** 16 - 1 <= w.n <= 2 ** 26 + 1 and w.segments == 1 and w.dtype == "u32" for w in ws)
    train, held = split(ws)
    assert len(train) == 72 and len(held) == 24
    cs = candidates("u32", 7)
    assert cs == STALL_CONFIGS and {c.kind for c in cs} == {"loop_csdldf"}
    assert {c.spin for c in cs} == {1, 2, 4, 8, 16, 32}
    assert default_config("u32", 1, 1 << 20, stall_period=8) == Config("loop_csdldf", 1024, 8, 4)
    assert default_config("u32", 1, 1 << 20, stall_period=8) in cs
    import gpuprefixsums_amd as gps
    assert Config("loop_csdl", 1024, 2, tile=True).handle_args() == (
        gps.ALGO_CSDL, dict(block_threads=1024, vec_per_thread=2, schedule=gps.SCHEDULE_TILE))


class _StubModel:
    """Predicts a fixed log-time per configuration name."""

    def __init__(self, table):
        self.table = table
        self.names = []

    def predict(self, X):
        # recover each row's configuration from its kind one-hot and shape features
        col = {name: i for i, name in enumerate(FEATURES)}
        out = []
        for row in np.asarray(X):
            kind = [k for k in KINDS if row[col[f"kind_{k}"]] == 1][0]
            c = Config(kind, int(round(2 ** row[col["log2_block"]])), int(round(2 ** row[col["log2_vec"]])),
                       int(round(2 ** row[col["log2_spin"]] - 1)) if kind == "loop_csdldf" else 0,
                       bool(row[col["persistent"]]) and kind in ("loop_fp", "batched_fp"),
                       kind == "loop_csdl" and not row[col["persistent"]])
            out.append(self.table.get(c.name, 100.0))
        return np.array(out)


def test_regret_arithmetic_against_hand_computed_values():
    # single-segment f32 workloads of 2^24 elements (the default handle's
    # choice there is 1024x8) timed on the three one-launch loop_fp shapes
    n = 1 << 24
    ws = [Workload(i, n, "f32", 1, np.array([0, n], dtype=np.uint64)) for i in range(4)]
    names = [c.name for c in candidates("f32") if c.kind == "loop_fp" and not c.persistent]  # 1024x8, 1024x4, 256x4
    assert default_config("f32", 1, n).name == names[0]
    times = {
        0: {names[0]: 2.0, names[1]: 1.0, names[2]: 4.0},
        1: {names[0]: 3.0, names[1]: 1.5, names[2]: 1.0},
        2: {names[0]: 1.0, names[1]: 2.0, names[2]: 2.0},
        3: {names[0]: 5.0, names[1]: 1.0, names[2]: 1.0},
    }
    # best static on workloads 0..1: mean regrets 1024x8: (1 + 2)/2 = 1.5; 1024x4: (0 + .5)/2 = .25; 256x4: (3+0)/2
    static = best_static(ws[:2], times)
    assert static == {"f32": names[1]}
    model = _StubModel({names[2]: 0.0, names[1]: 1.0, names[0]: 2.0})  # always picks 256x4
    r_mlp, r_def, r_stat, rows = regrets(model, ws[2:], times, static)
    # held out = workloads 2, 3. MLP picks 256x4: regrets 2/1-1 = 1, 1/1-1 = 0
    assert r_mlp["mean"] == pytest.approx(0.5) and r_mlp["max"] == pytest.approx(1.0)
    # default 1024x8: 0 and 4
    assert r_def["mean"] == pytest.approx(2.0)
    # static 1024x4: 1 and 0
    assert r_stat["mean"] == pytest.approx(0.5)
    assert [r["mlp"] for r in rows] == [names[2], names[2]]


def test_gated_choice_policy():
    cs = [Config("batched", 1024, 8), Config("batched", 1024, 4), Config("batched", 256, 4)]
    pred = np.log2(np.array([10.0, 9.0, 12.0]))  # 1024x4 predicted 10% faster than the default 1024x8
    d = cs[0]
    assert gated_choice(cs, pred, d, 0.0) == cs[1]
    assert gated_choice(cs, pred, d, 0.05) == cs[1]   # 11% predicted gain > 5%
    assert gated_choice(cs, pred, d, 0.2) == d        # < 20%: keep the default
    assert gated_choice(cs[1:], pred[1:], d, 0.2) == cs[1]  # default not a candidate: plain argmin


def test_config_handle_args():
    import gpuprefixsums_amd as gps
    assert Config("loop_fp", 1024, 8, persistent=True).handle_args() == (
        gps.ALGO_CSDL, dict(block_threads=1024, vec_per_thread=8, schedule=gps.SCHEDULE_PERSISTENT))
    assert Config("batched_fp", 256, 4).handle_args()[1]["schedule"] == gps.SCHEDULE_TILE
    assert Config("loop_csdldf", 1024, 4, 32).handle_args() == (
        gps.ALGO_CSDLDF, dict(block_threads=1024, vec_per_thread=4, spin_limit=32))
    assert Config("loop_rts", 512, 2).handle_args() == (gps.ALGO_RTS, dict(block_threads=512, vec_per_thread=2))
