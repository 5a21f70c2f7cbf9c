"""Trains the cost model on the harness timings and measures its regret.

    python -m gpuprefixsums_amd.tuner.evaluate --times times.jsonl [--save]

Training set: the first 384 workloads of the C5 sweep, the 1024 workloads of
the extra training sweep (same distribution, another seed; optional) and the
first 72 of the decoupled-fallback (stall) sweep; held out: the last 128 of
the C5 sweep and the last 24 of the stall sweep. One model is trained on all
of it and judged on each held-out set separately.
For a held-out workload with exhaustively measured times T(c):
    regret(choice) = T(choice) / min_c T(c) - 1
reported as mean / median / max over the held-out set, for the tuner's
choice and for two methods without the model: the library default (a default
handle's per-size choice), and the best static rule (per dtype, the one
configuration with the lowest mean regret on the training workloads).

The model is an ensemble of ENSEMBLE MLPs trained from different seeds, its
prediction their mean. The tuner takes the model's fastest predicted
configuration only when it beats the default's predicted time by more than a
margin. The margin and the training hyper-parameters (weight decay, epochs)
are chosen by 4-fold cross-validation on the training workloads alone; the
held-out workloads judge the result once.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import DATA, DEFAULT_MODEL
from .model import CostModel
from .space import (Config, Workload, c5_workloads, candidates, default_config, extra_workloads, features,
                    fixed_default_config, gated_choice, split, stall_workloads)

CV_FOLDS = 4
ENSEMBLE = 5  # MLPs per model (seeds 0..4), predictions averaged
CV_WEIGHT_DECAY = (1e-5, 1e-4, 1e-3)
CV_EPOCHS = (2000, 4000)
CV_MARGINS = (0.0, 0.02, 0.05, 0.1, 0.2)


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


def group(w: Workload) -> str:
    """Key of the model-free static rule: the dtype, or "u32_stall" for the
    decoupled-fallback sweep."""
    return "u32_stall" if w.stall_period else w.dtype


def best_static(ws: List[Workload], times: Dict[int, Dict[str, float]]) -> Dict[str, str]:
    """Per group, the one configuration with the lowest mean regret over `ws`
    (the strongest model-free rule: a fixed choice tuned on the training set)."""
    out = {}
    for g in sorted({group(w) for w in ws}):
        sub = [w for w in ws if group(w) == g]
        names = [c.name for c in candidates(sub[0].dtype, sub[0].stall_period)
                 if all(c.name in times[w.wid] for w in sub)]
        mean = {c: np.mean([times[w.wid][c] / min(times[w.wid].values()) - 1.0 for w in sub]) for c in names}
        out[g] = min(mean, key=mean.get)
    return out


def regrets(model: CostModel, ws: List[Workload], times: Dict[int, Dict[str, float]], static: Dict[str, str]):
    """Regrets of the tuner's policy (the gated choice at the model's margin),
    of the library default and of the static rule over `ws`."""
    margin = float(getattr(model, "margin", 0.0))
    mlp, dflt, stat, rows = [], [], [], []
    for w in ws:
        t = times[w.wid]
        best = min(t.values())
        cands = [c for c in candidates(w.dtype, w.stall_period) if c.name in t]
        pred = model.predict(np.array([features(w, c) for c in cands]))
        d = default_config(w.dtype, w.segments, w.n, w.stall_period)
        pick = gated_choice(cands, pred, d, margin)
        s = static.get(group(w), d.name)
        mlp.append(t[pick.name] / best - 1.0)
        dflt.append(t[d.name] / best - 1.0)
        stat.append(t[s] / best - 1.0)
        fx = fixed_default_config(w.dtype, w.segments, w.stall_period).name
        rows.append({"wid": w.wid, "n": w.n, "dtype": w.dtype, "segments": w.segments, "best": min(t, key=t.get),
                     "mlp": pick.name, "default": d.name, "static": s, "mlp_regret": mlp[-1],
                     "default_regret": dflt[-1], "static_regret": stat[-1], "fixed_regret": t[fx] / best - 1.0})
    return summarize(mlp), summarize(dflt), summarize(stat), rows


def summarize(r) -> dict:
    return {"mean": float(np.mean(r)), "median": float(np.median(r)), "max": float(np.max(r)),
            "within_5pct": float(np.mean(np.array(r) <= 0.05))}


def cross_validate(train: List[Workload], times: Dict[int, Dict[str, float]], seed: int = 0) -> dict:
    """Chooses (weight decay, epochs, margin) by CV_FOLDS-fold cross-validation
    on the training workloads: mean regret of the gated choice on each
    validation fold, averaged over folds; ties go to the smaller margin."""
    folds = [[w for i, w in enumerate(train) if i % CV_FOLDS == f] for f in range(CV_FOLDS)]
    scores = {}
    for wd in CV_WEIGHT_DECAY:
        for ep in CV_EPOCHS:
            per_margin = {m: [] for m in CV_MARGINS}
            for f in range(CV_FOLDS):
                fit_ws = [w for g in range(CV_FOLDS) if g != f for w in folds[g]]
                X, y = dataset(fit_ws, times)
                model = CostModel(members=ENSEMBLE).fit(X, y, epochs=ep, weight_decay=wd, seed=seed)
                print(f"[cv] weight decay {wd:g}, epochs {ep}, fold {f + 1}/{CV_FOLDS} fitted on {len(X)} samples",
                      file=sys.stderr, flush=True)
                for w in folds[f]:
                    t = times[w.wid]
                    best = min(t.values())
                    cands = [c for c in candidates(w.dtype, w.stall_period) if c.name in t]
                    pred = model.predict(np.array([features(w, c) for c in cands]))
                    d = default_config(w.dtype, w.segments, w.n, w.stall_period)
                    for m in CV_MARGINS:
                        per_margin[m].append(t[gated_choice(cands, pred, d, m).name] / best - 1.0)
            for m in CV_MARGINS:
                scores[(wd, ep, m)] = float(np.mean(per_margin[m]))
    wd, ep, m = min(scores, key=lambda k: (round(scores[k], 6), k[2]))
    return {"weight_decay": wd, "epochs": ep, "margin": m, "cv_regret": scores[(wd, ep, m)],
            "grid": {f"wd={k[0]:g} epochs={k[1]} margin={k[2]:g}": v for k, v in scores.items()}}


EXTRA_TIMES = "c5x_times_mi355x.jsonl"


def load_sweeps(c5_path: Path, stall_path: Path, extra_path: Optional[Path] = None):
    """Timings of the sweeps and their train / held-out splits:
    (times, c5 train, c5 held, stall train, stall held). The stall file is
    optional; so is the extra training sweep (by default the file of that
    name beside c5_path), whose workloads all join the c5 training set."""
    times = load_times(c5_path)
    if stall_path and Path(stall_path).exists():
        times.update(load_times(stall_path))
    if extra_path is None:
        extra_path = Path(c5_path).with_name(EXTRA_TIMES)
    if Path(extra_path).exists():
        times.update(load_times(extra_path))
    train, held = split([w for w in c5_workloads() if w.wid in times])
    train = train + [w for w in extra_workloads() if w.wid in times]
    s_train, s_held = split([w for w in stall_workloads() if w.wid in times])
    return times, train, held, s_train, s_held


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--times", default=str(DATA / "c5_times_mi355x.jsonl"))
    ap.add_argument("--stall-times", default=str(DATA / "stall_times_mi355x.jsonl"))
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save", action="store_true", help="write the model and the regret table")
    ap.add_argument("--out-dir", default=str(DATA), help="where --save writes (default tuner/data/)")
    args = ap.parse_args(argv)
    times, c5_train, held, s_train, s_held = load_sweeps(Path(args.times), Path(args.stall_times))
    train = c5_train + s_train
    if not held:
        print("no held-out workloads measured", file=sys.stderr)
        return 1
    cv = cross_validate(train, times, seed=args.seed)
    print(f"cross-validation: weight decay {cv['weight_decay']:g}, epochs {cv['epochs']}, "
          f"margin {cv['margin']:g}, regret {cv['cv_regret']:.4f}", file=sys.stderr)
    X, y = dataset(train, times)
    model = CostModel(members=ENSEMBLE).fit(X, y, epochs=cv["epochs"], lr=args.lr, weight_decay=cv["weight_decay"],
                                            seed=args.seed)
    model.margin = cv["margin"]
    Xh, yh = dataset(held, times)
    rmse_train = float(np.sqrt(np.mean((model.predict(X) - y) ** 2)))
    rmse_held = float(np.sqrt(np.mean((model.predict(Xh) - yh) ** 2)))
    static = best_static(train, times)
    r_mlp, r_def, r_stat, rows = regrets(model, held, times, static)
    model.margin, margin = 0.0, model.margin
    r_argmin = regrets(model, held, times, static)[0]
    model.margin = margin
    res = {"train_workloads": len(train), "held_out_workloads": len(held), "train_samples": int(len(X)),
           "cross_validation": cv, "regret_mlp_argmin": r_argmin,
           "rmse_log2_train": rmse_train, "rmse_log2_held": rmse_held, "final_loss": float(model.curve[-1]),
           "regret_mlp": r_mlp, "regret_default": r_def, "regret_best_static": r_stat,
           "regret_fixed_default": summarize([r["fixed_regret"] for r in rows]),
           "best_static": static}
    stall_rows = []
    if s_held:
        s_mlp, s_def, s_stat, stall_rows = regrets(model, s_held, times, static)
        res["stall_sweep"] = {"train_workloads": len(s_train), "held_out_workloads": len(s_held),
                              "regret_mlp": s_mlp, "regret_default": s_def, "regret_best_static": s_stat}
    print(json.dumps(res, indent=1))
    if args.save:
        out = Path(args.out_dir)
        out.mkdir(parents=True, exist_ok=True)
        model.save(out / DEFAULT_MODEL.name)
        (out / "regret_mi355x.json").write_text(
            json.dumps({**res, "held_out": rows, "stall_held_out": stall_rows}, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
