"""Per-word voxel counts (count_occupancy.py:25-44)."""

from math import prod

import torch
from torch import Tensor

from .. import _native
from .common import record_function


@record_function("count_occupancy")
def count_occupancy(occupancy: Tensor) -> Tensor:
    _native.require_device(occupancy)
    occupancy = occupancy.contiguous()
    i, j, kp = occupancy.shape
    shape = (i + 1, j + 1, kp + 1)
    voxel_counts = torch.empty((prod(shape) + 1,), dtype=torch.int32, device=occupancy.device)
    _native.call(
        "mt_count_occupancy", occupancy.data_ptr(), i, j, kp, voxel_counts.data_ptr(),
        _native.stream(occupancy),
    )
    voxel_counts[-1] = 0
    return voxel_counts
