"""The tuner's MLP kernels on the GPU against the FP64 CPU oracle (oracle/mlp.py,
itself pinned by finite differences), and the trained tuner's regret."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from gpuprefixsums_amd.tuner.model import CostModel, GpuMlp  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")


def problem(D, H, B, seed):
    rng = np.random.default_rng(seed)
    p = omlp.init_params(D, H, seed) + rng.normal(0, 0.05, omlp.param_count(D, H))
    x = rng.normal(size=(B, D))
    t = rng.normal(size=B)
    return p.astype(np.float32), x.astype(np.float32), t.astype(np.float32)


@pytest.mark.parametrize("D,H,B", [(15, 64, 1), (15, 64, 300), (7, 9, 1000), (3, 128, 5000)])
def test_forward_and_gradients_match_fp64_oracle(D, H, B):
    p, x, t = problem(D, H, B, seed=D + H + B)
    m = GpuMlp(D, H, B)
    try:
        assert m.count == omlp.param_count(D, H)
        m.set_params(p)
        assert np.array_equal(m.get_params(), p)
        y = m.forward(x)
        y64 = omlp.forward(p.astype(np.float64), x.astype(np.float64), D, H)
        assert np.allclose(y, y64, rtol=1e-5, atol=1e-5 * (1 + np.abs(y64).max()))
        loss, g = m.loss_grad(x, t)
        L64, g64 = omlp.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), D, H)
        assert abs(loss - L64) <= 1e-5 * max(1.0, L64)
        scale = np.abs(g64).max()
        err = np.abs(g - g64)
        assert err.max() <= 2e-5 * scale + 1e-7, (err.max(), scale, int(err.argmax()))
    finally:
        m.close()


def test_training_reduces_loss_and_fits():
    rng = np.random.default_rng(3)
    X = rng.uniform(-2, 2, size=(4000, 6))
    y = np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2 - X[:, 2] * X[:, 3]
    model = CostModel(hidden=64).fit(X, y, epochs=1500, lr=3e-3)
    assert model.curve[-1] < 0.05 * model.curve[0]
    pred = model.predict(X)
    assert np.sqrt(np.mean((pred - y) ** 2)) < 0.2


def test_trained_tuner_regret_on_held_out_workloads():
    """The shipped model's held-out regret against exhaustive search, and what
    it gains over the methods without the model: the library's per-size
    default, its fixed default and the best static rule, on the 128 held-out
    workloads of the C5 sweep and the 24 of the decoupled-fallback (stall)
    sweep, where the choice includes the lookback spin limit."""
    from gpuprefixsums_amd.tuner import DATA, DEFAULT_MODEL
    from gpuprefixsums_amd.tuner.evaluate import best_static, load_sweeps, regrets, summarize
    times_path, stall_path = DATA / "c5_times_mi355x.jsonl", DATA / "stall_times_mi355x.jsonl"
    if not times_path.exists() or not stall_path.exists() or not DEFAULT_MODEL.exists():
        pytest.skip("tuner data not generated yet")
    times, train, held, s_train, s_held = load_sweeps(times_path, stall_path)
    assert len(held) == 128 and len(s_held) == 24
    model = CostModel.load(DEFAULT_MODEL)
    static = best_static(train + s_train, times)
    r_mlp, r_def, r_stat, rows = regrets(model, held, times, static)
    r_fixed = summarize([r["fixed_regret"] for r in rows])
    s_mlp, s_def, s_stat, _ = regrets(model, s_held, times, static)
    print(json.dumps({"mlp": r_mlp, "per_size_default": r_def, "fixed_default": r_fixed, "best_static": r_stat,
                      "stall_mlp": s_mlp, "stall_default": s_def, "stall_static": s_stat}))
    assert r_mlp["mean"] <= 0.05
    assert r_mlp["mean"] < r_def["mean"]
    assert r_mlp["mean"] < r_fixed["mean"] and r_mlp["mean"] < r_stat["mean"]
    assert s_mlp["mean"] <= 0.05
    assert s_mlp["mean"] < s_def["mean"] and s_mlp["mean"] < s_stat["mean"]


def test_tuner_choose_with_the_model_and_by_exhaustive_search():
    """Tuner.choose: the shipped model's pick, and the exhaustive fallback
    (no model) that times every candidate on the GPU and keeps the fastest."""
    from gpuprefixsums_amd.tuner import DEFAULT_MODEL, Tuner
    from gpuprefixsums_amd.tuner.space import Workload, candidates, lognormal_partition
    rng = np.random.default_rng(3)
    w_u32 = Workload(0, 100_000, "u32", 12, lognormal_partition(100_000, 12, rng))
    w_f32 = Workload(1, 50_000, "f32", 3, lognormal_partition(50_000, 3, rng))
    if DEFAULT_MODEL.exists():
        for w in (w_u32, w_f32):
            assert Tuner().choose(w) in candidates(w.dtype)
    exhaustive = Tuner(model_path=None)
    for w in (w_u32, w_f32):
        c = exhaustive.choose(w)
        assert c in candidates(w.dtype)
        times = exhaustive._runner.time_all(w)
        # the pick is among the fastest (timings repeat within noise)
        assert times[c.name] <= 1.5 * min(times.values())
