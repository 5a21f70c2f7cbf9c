"""Workloads, candidate configurations and feature encoding of the tuner.

A workload is a batch of `segments` scans over `n` elements of one dtype
(u32, f32 or bf16), stored back to back; one segment is a plain scan. A
configuration fixes how it runs:

  batched      the one-launch segmented kernel (u32)           tile shape
  loop_csdl    one chained-scan launch per segment (u32)       tile shape
  loop_rts     one reduce-then-scan launch per segment (u32)   tile shape
  loop_csdldf  one decoupled-fallback launch per segment (u32) spin limit
  loop_fp      one compensated FP scan launch per segment      tile shape, schedule
  batched_fp   the one-launch segmented FP kernel              tile shape, schedule

The FP kinds run one workgroup per tile, or (":pp") the persistent pipelined
kernel. The u32 kinds run as a handle with that shape runs them: batched and
loop_csdl persistent (loop_csdl ":tile": one workgroup per tile, what a
default handle runs below 2^24 elements), loop_rts and loop_csdldf one
workgroup per tile.

A second sweep covers the decoupled-fallback scan under emulated deadlock
(every stall_period-th tile stalled until a successor inserts its
aggregate): there the choice is the tile shape and the lookback spin limit of
loop_csdldf, and a workload's stall period is one of its features. Without
stalls the spin limit never matters (no lookback waits long enough to fall
back), which is why the stall sweep exists.

Reference knobs these generalise: CSDL_WARPS / UINT4_PER_THREAD
(GPUPrefixSumsCUDA/ChainedScanDecoupledLookbackDispatcher.cuh:13-14) and
MAX_SPIN_COUNT (GPUPrefixSumsCUDA/EmulatedDeadlocking.cuh:15).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

DTYPES = ("u32", "f32", "bf16")
KINDS = ("batched", "loop_csdl", "loop_rts", "loop_csdldf", "loop_fp", "batched_fp")


@dataclass(frozen=True)
class Config:
    kind: str
    block: int
    vec: int
    spin: int = 0
    persistent: bool = False  # FP kinds: the persistent pipelined kernel
    tile: bool = False        # loop_csdl: one workgroup per tile instead of the persistent kernel

    @property
    def name(self) -> str:
        s = f"{self.kind}:{self.block}x{self.vec}"
        s += f":spin{self.spin}" if self.kind == "loop_csdldf" else ""
        return s + (":pp" if self.persistent else "") + (":tile" if self.tile else "")

    def handle_args(self):
        """(algo, keyword arguments) of the gpuprefixsums_amd.ScanHandle that
        runs this configuration: ScanHandle(algo, max_n, **kwargs)."""
        import gpuprefixsums_amd as gps
        algo = {"batched": gps.ALGO_CSDL, "loop_csdl": gps.ALGO_CSDL, "loop_rts": gps.ALGO_RTS,
                "loop_csdldf": gps.ALGO_CSDLDF, "loop_fp": gps.ALGO_CSDL, "batched_fp": gps.ALGO_CSDL}[self.kind]
        kw = dict(block_threads=self.block, vec_per_thread=self.vec)
        if self.kind == "loop_csdldf":
            kw["spin_limit"] = self.spin
        if self.kind in ("loop_fp", "batched_fp"):
            kw["schedule"] = gps.SCHEDULE_PERSISTENT if self.persistent else gps.SCHEDULE_TILE
        if self.kind == "loop_csdl" and self.tile:
            kw["schedule"] = gps.SCHEDULE_TILE
        return algo, kw

    @property
    def runs_persistent(self) -> bool:
        """True when the configuration runs the persistent pipelined kernel."""
        return self.persistent or (self.kind in ("batched", "loop_csdl") and not self.tile)

    @staticmethod
    def parse(name: str) -> "Config":
        parts = name.split(":")
        b, v = (int(t) for t in parts[1].split("x"))
        spin = next((int(t[4:]) for t in parts[2:] if t.startswith("spin")), 0)
        return Config(parts[0], b, v, spin, "pp" in parts[2:], "tile" in parts[2:])


U32_CONFIGS = (
    [Config("batched", b, v) for b, v in ((256, 4), (512, 4), (1024, 2), (1024, 4), (1024, 8))]
    + [Config("loop_csdl", b, v) for b, v in ((1024, 8), (1024, 4), (512, 8), (256, 4))]
    + [Config("loop_csdl", b, v, tile=True) for b, v in ((1024, 2), (1024, 4), (1024, 8))]
    + [Config("loop_rts", b, v) for b, v in ((512, 2), (256, 4), (1024, 2))]
    + [Config("loop_csdldf", 1024, 4, s) for s in (1, 4, 32)]
)
STALL_SPINS = (1, 2, 4, 8, 16, 32)
STALL_CONFIGS = [Config("loop_csdldf", b, v, s) for b, v in ((1024, 4), (1024, 8)) for s in STALL_SPINS]
FP_SHAPES = ((1024, 8), (1024, 4), (256, 4))
FP_PP_SHAPES = ((1024, 8), (1024, 4))
FP_CONFIGS = ([Config(k, b, v) for k in ("loop_fp", "batched_fp") for b, v in FP_SHAPES]
              + [Config(k, b, v, persistent=True) for k in ("loop_fp", "batched_fp") for b, v in FP_PP_SHAPES])


def candidates(dtype: str, stall_period: int = 0) -> List[Config]:
    if stall_period:
        return list(STALL_CONFIGS)
    return list(U32_CONFIGS) if dtype == "u32" else list(FP_CONFIGS)


def default_config(dtype: str, segments: int, n: int = 1 << 30, stall_period: int = 0) -> Config:
    """What the library does without the tuner: a default handle's choice for
    n elements (the same rules as libgps's size-adaptive defaults). Under
    emulated deadlock it is the default decoupled-fallback handle (1024 x 8,
    spin limit 4)."""
    if stall_period:
        return Config("loop_csdldf", 1024, 8, 4)
    if dtype != "u32":
        if segments == 1:
            b, v = (256, 4) if n <= 1 << 18 else (1024, 4) if n <= 1 << 22 else (1024, 8)
            return Config("loop_fp", b, v, persistent=n > 1 << 24)
        if n > 1 << 22:
            return Config("batched_fp", 1024, 8, persistent=True)
        b, v = (256, 4) if n < 1 << 20 else (1024, 4)
        return Config("batched_fp", b, v)
    if segments == 1:
        if n <= 1 << 18:
            return Config("loop_csdl", 1024, 2, tile=True)
        if n <= 1 << 21:
            return Config("loop_csdl", 1024, 4, tile=True)
        if n < 1 << 24:
            return Config("loop_csdl", 1024, 8, tile=True)
        return Config("loop_csdl", 1024, 8)
    b, v = (1024, 2) if n < 1 << 20 else (1024, 4) if n < 1 << 22 else (1024, 8)
    return Config("batched", b, v)


def fixed_default_config(dtype: str, segments: int, stall_period: int = 0) -> Config:
    """The library's fixed configuration for every size (what a handle did
    before its defaults depended on n): 1024x8 throughout, FP kernels one
    workgroup per tile."""
    if stall_period:
        return Config("loop_csdldf", 1024, 8, 4)
    if dtype != "u32":
        return Config("loop_fp" if segments == 1 else "batched_fp", 1024, 8)
    return Config("loop_csdl" if segments == 1 else "batched", 1024, 8)


def gated_choice(cands: List[Config], pred_log2: "np.ndarray", default: Config, margin: float) -> Config:
    """The tuner's policy: the model's fastest predicted configuration if it
    beats the predicted time of the library default by more than `margin`
    (relative), else the default. margin = 0 is the plain argmin."""
    names = [c.name for c in cands]
    best = int(np.argmin(pred_log2))
    if default.name not in names:
        return cands[best]
    if pred_log2[best] < pred_log2[names.index(default.name)] - math.log2(1.0 + margin):
        return cands[best]
    return default


@dataclass
class Workload:
    wid: int
    n: int
    dtype: str
    segments: int
    offsets: np.ndarray  # uint64[segments + 1], offsets[0] = 0, offsets[-1] = n
    stall_period: int = 0  # emulated deadlock (decoupled-fallback sweep); 0 = none

    @property
    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))


def lognormal_partition(n: int, segments: int, rng: np.random.Generator, sigma: float = 1.5) -> np.ndarray:
    """Splits n into `segments` log-normal lengths (sigma 1.5) that sum to n
    exactly (largest-remainder rounding; segments may be empty when
    segments > n)."""
    w = np.exp(sigma * rng.standard_normal(segments))
    q = w / w.sum() * n
    lengths = np.floor(q).astype(np.int64)
    rem = n - int(lengths.sum())
    if rem:
        lengths[np.argsort(-(q - lengths), kind="stable")[:rem]] += 1
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)


def c5_workloads(seed: int = 6, count: int = 512, wid_base: int = 0) -> List[Workload]:
    """The tuner sweep: n log-uniform in [2^10, 2^30], dtype uniform over
    {u32, f32, bf16}, segment count log-uniform in [1, 2^14]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(2.0 ** rng.uniform(10, 30)))
        dtype = DTYPES[int(rng.integers(3))]
        segments = int(min(max(round(2.0 ** rng.uniform(0, 14)), 1), 1 << 14))
        out.append(Workload(wid_base + i, n, dtype, segments, lognormal_partition(n, segments, rng)))
    return out


EXTRA_WID_BASE = 20000


def extra_workloads(seed: int = 8, count: int = 1024) -> List[Workload]:
    """More training workloads from the same distribution as the tuner sweep
    (another seed; workload ids start at 20000). All of them train; the tuner
    sweep's held-out workloads stay the only judge."""
    return c5_workloads(seed, count, EXTRA_WID_BASE)


STALL_WID_BASE = 10000


def stall_workloads(seed: int = 7, count: int = 96) -> List[Workload]:
    """The decoupled-fallback sweep: single uint32 scans with n log-uniform in
    [2^16, 2^26] under emulated deadlock with stall period 7 or 8 (every 7th
    tile, or the reference's !(idx & 7)). Workload ids start at 10000."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(2.0 ** rng.uniform(16, 26)))
        period = (7, 8)[int(rng.integers(2))]
        out.append(Workload(STALL_WID_BASE + i, n, "u32", 1, np.array([0, n], dtype=np.uint64), period))
    return out


def split(workloads: List[Workload], train: int = 384):
    """Workloads are i.i.d., so the first `train` form the training set and the
    rest are held out (the stall sweep: the first 72 of 96)."""
    if workloads and workloads[0].stall_period:
        train = len(workloads) * 3 // 4
    return workloads[:train], workloads[train:]


FEATURES = ["log2_n", "log2_segments", "log2_mean_len", "log2_max_len", "is_u32", "is_f32", "is_bf16",
            "log2_stall"] + [f"kind_{k}" for k in KINDS] + ["log2_block", "log2_vec", "log2_spin", "persistent"]


def workload_features(w: Workload) -> List[float]:
    L = w.lengths
    return [math.log2(w.n), math.log2(w.segments), math.log2(w.n / w.segments + 1.0), math.log2(float(L.max()) + 1.0),
            float(w.dtype == "u32"), float(w.dtype == "f32"), float(w.dtype == "bf16"),
            math.log2(w.stall_period + 1.0)]


def config_features(c: Config) -> List[float]:
    return [float(c.kind == k) for k in KINDS] + [math.log2(c.block), math.log2(c.vec), math.log2(c.spin + 1.0),
                                                  float(c.runs_persistent)]


def features(w: Workload, c: Config) -> List[float]:
    return workload_features(w) + config_features(c)
