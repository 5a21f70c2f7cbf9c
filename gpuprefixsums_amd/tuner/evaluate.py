"""Trains the cost model on the harness timings and measures its regret.

    python -m gpuprefixsums_amd.tuner.evaluate --times times.jsonl [--save]

Training set: the first 384 workloads of the sweep; held out: the last 128.
For a held-out workload with exhaustively measured times T(c):
    regret(choice) = T(choice) / min_c T(c) - 1
reported as mean / median / max over the held-out set, for the MLP's choice
and for two methods without the model: the library default, and the best
static rule (per dtype, the one configuration with the lowest mean regret on
the training workloads).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path
from typing import Dict, List, Tuple

import numpy as np

from . import DATA, DEFAULT_MODEL
from .model import CostModel
from .space import Config, Workload, c5_workloads, candidates, default_config, features, split


def load_times(path: Path) -> Dict[int, Dict[str, float]]:
    out: Dict[int, Dict[str, float]] = {}
    for line in Path(path).read_text().splitlines():
        if line.strip():
            r = json.loads(line)
            out[int(r["wid"])] = r["times_ms"]
    return out


def dataset(ws: List[Workload], times: Dict[int, Dict[str, float]]) -> Tuple[np.ndarray, np.ndarray]:
    X, y = [], []
    for w in ws:
        for name, ms in times[w.wid].items():
            X.append(features(w, Config.parse(name)))
            y.append(math.log2(ms * 1000.0))  # log2 microseconds
    return np.array(X), np.array(y)


def best_static(ws: List[Workload], times: Dict[int, Dict[str, float]]) -> Dict[str, str]:
    """Per dtype, the one configuration with the lowest mean regret over `ws`
    (the strongest model-free rule: a fixed choice tuned on the training set)."""
    out = {}
    for dt in sorted({w.dtype for w in ws}):
        sub = [w for w in ws if w.dtype == dt]
        names = [c.name for c in candidates(dt) if all(c.name in times[w.wid] for w in sub)]
        mean = {c: np.mean([times[w.wid][c] / min(times[w.wid].values()) - 1.0 for w in sub]) for c in names}
        out[dt] = min(mean, key=mean.get)
    return out


def regrets(model: CostModel, ws: List[Workload], times: Dict[int, Dict[str, float]], static: Dict[str, str]):
    mlp, dflt, stat, rows = [], [], [], []
    for w in ws:
        t = times[w.wid]
        best = min(t.values())
        cands = [c for c in candidates(w.dtype) if c.name in t]
        pred = model.predict(np.array([features(w, c) for c in cands]))
        pick = cands[int(np.argmin(pred))]
        d = default_config(w.dtype, w.segments)
        s = static.get(w.dtype, d.name)
        mlp.append(t[pick.name] / best - 1.0)
        dflt.append(t[d.name] / best - 1.0)
        stat.append(t[s] / best - 1.0)
        rows.append({"wid": w.wid, "n": w.n, "dtype": w.dtype, "segments": w.segments, "best": min(t, key=t.get),
                     "mlp": pick.name, "default": d.name, "static": s, "mlp_regret": mlp[-1],
                     "default_regret": dflt[-1], "static_regret": stat[-1]})
    summ = lambda r: {"mean": float(np.mean(r)), "median": float(np.median(r)), "max": float(np.max(r)),  # noqa: E731
                      "within_5pct": float(np.mean(np.array(r) <= 0.05))}
    return summ(mlp), summ(dflt), summ(stat), rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--times", default=str(DATA / "c5_times_mi355x.jsonl"))
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save", action="store_true", help="write the model and the regret table")
    ap.add_argument("--out-dir", default=str(DATA), help="where --save writes (default tuner/data/)")
    args = ap.parse_args(argv)
    times = load_times(Path(args.times))
    ws = [w for w in c5_workloads() if w.wid in times]
    train, held = split(ws)
    if not held:
        print("no held-out workloads measured", file=sys.stderr)
        return 1
    X, y = dataset(train, times)
    model = CostModel().fit(X, y, epochs=args.epochs, lr=args.lr, seed=args.seed)
    Xh, yh = dataset(held, times)
    rmse_train = float(np.sqrt(np.mean((model.predict(X) - y) ** 2)))
    rmse_held = float(np.sqrt(np.mean((model.predict(Xh) - yh) ** 2)))
    static = best_static(train, times)
    r_mlp, r_def, r_stat, rows = regrets(model, held, times, static)
    res = {"train_workloads": len(train), "held_out_workloads": len(held), "train_samples": int(len(X)),
           "rmse_log2_train": rmse_train, "rmse_log2_held": rmse_held, "final_loss": float(model.curve[-1]),
           "regret_mlp": r_mlp, "regret_default": r_def, "regret_best_static": r_stat,
           "best_static": static}
    print(json.dumps(res, indent=1))
    if args.save:
        out = Path(args.out_dir)
        out.mkdir(parents=True, exist_ok=True)
        model.save(out / DEFAULT_MODEL.name)
        (out / "regret_mi355x.json").write_text(json.dumps({**res, "held_out": rows}, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
