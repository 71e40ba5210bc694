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
