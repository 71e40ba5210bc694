"""Stable signed-int64 radix sort of the paired keys (sort_keys.py:18-29)."""

import torch

from .. import _native
from .common import TileGrid, record_function
from .generate_keys import PairedKeys


@record_function("sort_keys")
def sort_keys(paired_keys: PairedKeys, grid: TileGrid) -> PairedKeys:
    """Returns new sorted tensors. Unlike the reference (CUB on the legacy default stream
    followed by a device-wide synchronize), the sort is enqueued on the current stream and
    nothing waits for it."""
    keys = paired_keys.keys.contiguous()
    values = paired_keys.triangle_indices.contiguous()
    if keys.dtype != torch.int64 or values.dtype != torch.int32 or keys.shape != values.shape:
        raise RuntimeError("sort_keys expects int64 keys and int32 values of the same shape")
    _native.require_device(keys, values)
    n = keys.numel()
    out_k = torch.empty_like(keys)
    out_v = torch.empty_like(values)
    if n > 0:
        ws = _native.workspace(_native.query("mt_sort_workspace_bytes", n), keys.device)
        _native.call("mt_sort_keys", keys.data_ptr(), values.data_ptr(), n, out_k.data_ptr(),
                     out_v.data_ptr(), ws.data_ptr(), ws.numel(), None, _native.stream(keys))
    return PairedKeys(out_k, out_v)
