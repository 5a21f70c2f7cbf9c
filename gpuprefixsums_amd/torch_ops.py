"""PyTorch operators over the C ABI, for framework callers.

    import gpuprefixsums_amd.torch_ops  # registers torch.ops.gps.*
    y = torch.ops.gps.cumsum(x)                       # uint32 / int32 (bits, mod 2^32), float32, bfloat16
    y = torch.ops.gps.cumsum(x, inclusive=False)      # exclusive (integer inputs)
    y = torch.ops.gps.segmented_cumsum(x, offsets)    # one launch over concatenated segments (inclusive;
                                                      # inclusive=False for exclusive integer scans)

Each operator runs the library's HIP kernels on the tensor's device and
torch's current stream; there is no CPU path (a CPU tensor is an error).
Integer scans wrap mod 2^32 and return the input's dtype; floating-point
scans accumulate in FP32 with two-sum compensated carries (compensated=True)
and return float32. Fake (meta) implementations give output shapes and
dtypes for tracing.

Graph capture: the ops launch one kernel and allocate nothing once a handle
of the needed size exists for the (device, stream) pair, so they can be
captured after one eager call at that size on the capturing stream; a call
that would need a new or larger handle during capture raises. Replaced
handles are kept open, so earlier captures stay valid.
"""
from __future__ import annotations

import threading
from typing import Dict, List, Tuple

import torch

from . import ALGO_CSDL, ScanHandle

_INT = (torch.int32, torch.uint32)
_FLOAT = (torch.float32, torch.bfloat16)
_handles: Dict[Tuple[int, int], ScanHandle] = {}
# Handles replaced by a larger one stay open until the process ends: a graph
# captured earlier (torch.cuda.graph, torch.compile's CUDA graphs) still holds
# their status and ticket addresses and may be replayed at any time.
_retired: List[ScanHandle] = []
_lock = threading.Lock()


def _handle(x: torch.Tensor, n: int) -> ScanHandle:
    """One handle per (device, stream), grown when a larger scan arrives: a
    handle's status words serve one scan at a time. Creating or growing a
    handle allocates device memory and synchronises, which stream capture
    forbids, so under capture the handle must already exist at size n (run
    the op once eagerly on that stream first)."""
    stream = torch.cuda.current_stream(x.device)
    key = (x.device.index or 0, int(stream.cuda_stream))
    with _lock:
        h = _handles.get(key)
        if h is None or h.max_n < n:
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError(
                    f"gps: no scan handle of {n} elements exists for this stream and one cannot be created "
                    "during graph capture (it allocates device memory); run the op once on this stream at "
                    "this size or larger before capturing")
            if h is not None:
                _retired.append(h)
            with torch.cuda.device(x.device):
                h = ScanHandle(ALGO_CSDL, max(n, 1 << 20))
            _handles[key] = h
        return h


def _check(x: torch.Tensor, what: str) -> None:
    if not x.is_cuda:
        raise RuntimeError(f"gps::{what} runs on a ROCm device; got a {x.device} tensor")
    if x.dim() != 1:
        raise RuntimeError(f"gps::{what} takes a 1-D tensor, got shape {tuple(x.shape)}")
    if x.dtype not in _INT + _FLOAT:
        raise RuntimeError(f"gps::{what}: unsupported dtype {x.dtype} (uint32, int32, float32, bfloat16)")


@torch.library.custom_op("gps::cumsum", mutates_args=())
def cumsum(x: torch.Tensor, inclusive: bool = True, compensated: bool = True) -> torch.Tensor:
    """Device-wide prefix sum of a 1-D tensor (see the module docstring)."""
    _check(x, "cumsum")
    x = x.contiguous()
    n = x.numel()
    if x.dtype not in _INT and not inclusive:
        raise RuntimeError("gps::cumsum: floating-point scans are inclusive")
    with torch.cuda.device(x.device):  # the library launches on the current device
        stream = torch.cuda.current_stream(x.device)
        if x.dtype in _INT:
            out = torch.empty_like(x)
            if n:
                _handle(x, n).scan(x, out, n, inclusive, stream=stream)
            return out
        out = torch.empty(n, dtype=torch.float32, device=x.device)
        if n:
            h = _handle(x, n)
            if x.dtype == torch.bfloat16:
                h.inclusive_bf16_f32(x.view(torch.int16), out, n, compensated, stream=stream)
            else:
                h.inclusive_f32(x, out, n, compensated, stream=stream)
        return out


@cumsum.register_fake
def _(x, inclusive=True, compensated=True):
    return torch.empty_like(x) if x.dtype in _INT else x.new_empty(x.shape, dtype=torch.float32)


@torch.library.custom_op("gps::segmented_cumsum", mutates_args=())
def segmented_cumsum(x: torch.Tensor, offsets: torch.Tensor, inclusive: bool = True,
                     compensated: bool = True) -> torch.Tensor:
    """Prefix sums restarting at every segment: segment s is
    x[offsets[s]:offsets[s+1]] (int64 offsets on the same device, offsets[0] = 0,
    offsets[-1] = x.numel()). One kernel launch for all segments."""
    _check(x, "segmented_cumsum")
    if offsets.dtype != torch.int64 or offsets.dim() != 1 or offsets.device != x.device:
        raise RuntimeError("gps::segmented_cumsum: offsets must be a 1-D int64 tensor on x's device")
    x, offsets = x.contiguous(), offsets.contiguous()
    n, segments = x.numel(), offsets.numel() - 1
    if segments < 0:
        raise RuntimeError("gps::segmented_cumsum: offsets needs at least one entry")
    if x.dtype not in _INT and not inclusive:
        raise RuntimeError("gps::segmented_cumsum: floating-point scans are inclusive")
    with torch.cuda.device(x.device):  # the library launches on the current device
        stream = torch.cuda.current_stream(x.device)
        if x.dtype in _INT:
            out = torch.empty_like(x)
            if n:
                _handle(x, n).segmented(x, out, offsets, segments, n, inclusive, stream=stream)
            return out
        out = torch.empty(n, dtype=torch.float32, device=x.device)
        if n:
            h = _handle(x, n)
            src = x.view(torch.int16) if x.dtype == torch.bfloat16 else x
            h.segmented_fp(src, out, offsets, segments, n, bf16=x.dtype == torch.bfloat16,
                           compensated=compensated, stream=stream)
        return out


@segmented_cumsum.register_fake
def _(x, offsets, inclusive=True, compensated=True):
    return torch.empty_like(x) if x.dtype in _INT else x.new_empty(x.shape, dtype=torch.float32)


__all__ = ["cumsum", "segmented_cumsum"]
