"""Device-buffer plumbing for the GPU tests (torch tensors on the ROCm device).

uint32 data lives in int32 tensors (same bits); numpy uint32 arrays go in and
come out through .view().
"""
from __future__ import annotations

import numpy as np
import torch

DEV = "cuda"


def to_dev(x: np.ndarray, pad: int = 0) -> torch.Tensor:
    """Copies a numpy array to the device (optionally with `pad` extra
    elements after it, filled with a sentinel pattern)."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.uint32:
        t = torch.from_numpy(x.view(np.int32))
    elif x.dtype == np.uint16:
        t = torch.from_numpy(x.view(np.int16))
    elif x.dtype == np.uint64:
        t = torch.from_numpy(x.view(np.int64))
    else:
        t = torch.from_numpy(x)
    if pad:
        t = torch.cat([t, torch.full((pad,), -0x21524111, dtype=t.dtype)])  # 0xDEADBEEF
    return t.to(DEV)


def empty_u32(n: int) -> torch.Tensor:
    return torch.empty(max(n, 1), dtype=torch.int32, device=DEV)


def sentinel_u32(n: int) -> torch.Tensor:
    return torch.full((max(n, 1),), -0x21524111, dtype=torch.int32, device=DEV)


def from_dev_u32(t: torch.Tensor, n: int | None = None) -> np.ndarray:
    a = t.cpu().numpy().view(np.uint32)
    return a if n is None else a[:n]


def sync() -> None:
    torch.cuda.synchronize()


def first_mismatch(a: np.ndarray, b: np.ndarray) -> str:
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = np.flatnonzero(a != b)
    if bad.size == 0:
        return "equal"
    i = int(bad[0])
    return f"{bad.size} mismatches, first at {i}: got {int(a[i])} expected {int(b[i])}"
