"""The learned cost model: an ensemble of 3-layer MLPs (hidden 64, FP32) over
workload and configuration features that predicts log2 of the run time in
microseconds.
Training and inference run on the GPU through the library's own HIP kernels
(gps_mlp_* in include/gps.h); this module only standardises features and
moves small host arrays.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

import numpy as np

from .._lib import check, load


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class GpuMlp:
    """Thin owner of a gps_mlp object (parameters, Adam state, activations)."""

    def __init__(self, in_dim: int, hidden: int, max_batch: int):
        self.in_dim, self.hidden, self.max_batch = in_dim, hidden, max_batch
        h = ctypes.c_void_p()
        check(load().gps_mlp_create(in_dim, hidden, max_batch, ctypes.byref(h)))
        self._h = h
        self.count = int(load().gps_mlp_param_count(in_dim, hidden))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            check(load().gps_mlp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.count
        check(load().gps_mlp_set_params(self._h, _fp(p)))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.count, np.float32)
        check(load().gps_mlp_get_params(self._h, _fp(p)))
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], np.float32)
        check(load().gps_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        g = np.empty(self.count, np.float32)
        check(load().gps_mlp_loss_grad(self._h, _fp(x), _fp(t), x.shape[0], _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int, lr: float, weight_decay: float = 0.0) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        curve = np.empty(epochs, np.float32)
        check(load().gps_mlp_train(self._h, _fp(x), _fp(t), x.shape[0], epochs, lr, weight_decay, _fp(curve)))
        return curve


class CostModel:
    """Standardised features -> an ensemble of GpuMlp -> predicted log2(us).

    `members` MLPs are trained from different seeds on the same data and
    their predictions averaged: one network's fit varies with its
    initialisation by more than many configurations differ in time, and the
    mean of several is steadier where the tuner must choose between near ties.
    """

    def __init__(self, hidden: int = 64, members: int = 1):
        self.hidden = hidden
        self.members = members
        self.margin = 0.0  # the tuner's gate: deviate from the default only above this predicted gain
        self.mean: Optional[np.ndarray] = None
        self.std: Optional[np.ndarray] = None
        self.params: Optional[np.ndarray] = None  # (members, count)
        self.curve: Optional[np.ndarray] = None  # the first member's training loss per epoch

    def _norm(self, X: np.ndarray) -> np.ndarray:
        return (np.asarray(X, np.float64) - self.mean) / self.std

    @staticmethod
    def init_params(D: int, H: int, seed: int) -> np.ndarray:
        """He-uniform weights, zero biases, in the flat layout of gps_mlp."""
        rng = np.random.default_rng(seed)
        W1 = rng.uniform(-1, 1, (H, D)) * np.sqrt(6.0 / D)
        W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
        W3 = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
        return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3, [0.0]]).astype(np.float32)

    def fit(self, X: np.ndarray, y: np.ndarray, epochs: int = 4000, lr: float = 3e-3, weight_decay: float = 1e-5,
            seed: int = 0) -> "CostModel":
        """Full-batch AdamW on the GPU (the library's forward/backward/Adam
        kernels), once per member with seeds seed, seed + 1, ..."""
        X = np.asarray(X, np.float64)
        self.mean = X.mean(axis=0)
        self.std = np.where(X.std(axis=0) > 1e-12, X.std(axis=0), 1.0)
        m = GpuMlp(X.shape[1], self.hidden, len(X))
        params = []
        try:
            for k in range(self.members):
                m.set_params(self.init_params(X.shape[1], self.hidden, seed + k))
                curve = m.train(self._norm(X), np.asarray(y, np.float32), epochs, lr, weight_decay)
                if k == 0:
                    self.curve = curve
                params.append(m.get_params())
        finally:
            m.close()
        self.params = np.stack(params)
        return self

    def predict(self, X: np.ndarray) -> np.ndarray:
        assert self.params is not None, "model not trained or loaded"
        m = GpuMlp(self.mean.size, self.hidden, max(1, len(X)))
        try:
            out = np.zeros(len(X), np.float64)
            for p in self.params:
                m.set_params(p)
                out += m.forward(self._norm(X)).astype(np.float64)
            return out / len(self.params)
        finally:
            m.close()

    def save(self, path: Path) -> None:
        np.savez(path, params=self.params, mean=self.mean, std=self.std, hidden=self.hidden,
                 curve=self.curve if self.curve is not None else np.zeros(0), margin=self.margin)

    @classmethod
    def load(cls, path: Path) -> "CostModel":
        d = np.load(path, allow_pickle=False)
        params = np.atleast_2d(d["params"])
        m = cls(int(d["hidden"]), members=params.shape[0])
        m.params, m.mean, m.std = params, d["mean"], d["std"]
        m.curve = d["curve"]
        m.margin = float(d["margin"]) if "margin" in d.files else 0.0
        return m
