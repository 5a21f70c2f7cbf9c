"""Learned tuner: picks the scan configuration (algorithm, tile shape,
schedule, lookback spin limit) for a workload from its size, dtype, segment
statistics and, for the decoupled-fallback scan under emulated deadlock, its
stall period.

    from gpuprefixsums_amd.tuner import Tuner
    cfg = Tuner().choose(workload)          # MLP cost model (trained on the GPU)
    algo, kw = cfg.handle_args()            # the handle that runs it
    h = gpuprefixsums_amd.ScanHandle(algo, max_n, **kw)

The cost model is a 3-layer MLP (hidden 64, FP32) trained on timings the
exhaustive harness (tuner.harness) measured on MI355X; the shipped weights and
data live in tuner/data/. Without a model, or for a workload outside the
trained range, `choose` falls back to exhaustive autotuning: it times every
candidate and keeps the fastest.
"""
from __future__ import annotations

from pathlib import Path
from typing import Optional

import numpy as np

from .space import Config, Workload, candidates, default_config, features, gated_choice

DATA = Path(__file__).resolve().parent / "data"
DEFAULT_MODEL = DATA / "mlp_mi355x.npz"


class Tuner:
    def __init__(self, model_path: Optional[Path] = DEFAULT_MODEL, runner=None):
        from .model import CostModel
        self.model = CostModel.load(model_path) if model_path and Path(model_path).exists() else None
        self._runner = runner

    def predict_ms(self, w: Workload) -> dict:
        cands = candidates(w.dtype, w.stall_period)
        X = np.array([features(w, c) for c in cands])
        pred = self.model.predict(X)
        return {c: float(2.0 ** p / 1000.0) for c, p in zip(cands, pred)}

    def in_range(self, w: Workload) -> bool:
        return 2 ** 10 <= w.n <= 2 ** 30 and 1 <= w.segments <= 2 ** 14

    def choose(self, w: Workload, fallback: str = "exhaustive") -> Config:
        if self.model is not None and self.in_range(w):
            cands = candidates(w.dtype, w.stall_period)
            pred = self.model.predict(np.array([features(w, c) for c in cands]))
            return gated_choice(cands, pred, default_config(w.dtype, w.segments, w.n, w.stall_period),
                                self.model.margin)
        if fallback == "default":
            return default_config(w.dtype, w.segments, w.n, w.stall_period)
        if self._runner is None:
            from .harness import Runner
            self._runner = Runner(max_n=max(w.n, 1 << 10))
        times = self._runner.time_all(w)
        return Config.parse(min(times, key=times.get))


__all__ = ["Tuner", "Config", "Workload", "candidates", "default_config"]
