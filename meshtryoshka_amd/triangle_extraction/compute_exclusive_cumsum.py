"""In-place exclusive prefix sum (compute_exclusive_cumsum.py:19-21, .cu:8-39)."""

import torch
from torch import Tensor

from .. import _native
from .common import record_function


@record_function("compute_exclusive_cumsum")
def compute_exclusive_cumsum(x: Tensor) -> None:
    if x.dtype != torch.int32:
        raise RuntimeError("compute_exclusive_cumsum expects int32")
    if not x.is_contiguous():
        raise RuntimeError("compute_exclusive_cumsum expects a contiguous tensor")
    _native.require_device(x)
    n = x.numel()
    ws = _native.workspace(_native.query("mt_scan_workspace_bytes", n), x.device)
    _native.call(
        "mt_exclusive_scan_i32", x.data_ptr(), n, ws.data_ptr(), ws.numel(), _native.stream(x)
    )


def exclusive_offsets_i64(counts: Tensor) -> Tensor:
    """Exclusive int64 offsets of an int32 count array whose last entry is a zero pad
    (the reference's ``[n + 1]`` layout): out[i] = sum(counts[:i]), out[n] = the total.
    One pass of ``mt_inclusive_scan_i32_to_i64`` over counts[:n] into out[1:]."""
    if counts.dtype != torch.int32 or not counts.is_contiguous():
        raise RuntimeError("exclusive_offsets_i64 expects a contiguous int32 tensor")
    _native.require_device(counts)
    n = counts.numel() - 1
    out = torch.zeros(counts.numel(), dtype=torch.int64, device=counts.device)
    if n > 0:
        ws = _native.workspace(_native.query("mt_scan_workspace_bytes", n), counts.device)
        _native.call("mt_inclusive_scan_i32_to_i64", counts.data_ptr(), out[1:].data_ptr(), n,
                     ws.data_ptr(), ws.numel(), _native.stream(counts))
    return out
