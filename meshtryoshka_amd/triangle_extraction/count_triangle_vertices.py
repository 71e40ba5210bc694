"""Per-column edge-crossing counts (count_triangle_vertices.py:22-50)."""

import torch
from torch import Tensor

from .. import _native
from .common import record_function


@record_function("count_triangle_vertices")
def count_triangle_vertices(signed_distances: Tensor, lower_corners: Tensor, level_sets: Tensor) -> Tensor:
    """Returns int32[S * voxel_and_subvoxel + 1], level-set major, last entry 0."""
    _native.require_device(signed_distances, lower_corners, level_sets)
    (num_level_sets,) = level_sets.shape
    _, num_columns = lower_corners.shape
    counts = torch.empty((num_level_sets * num_columns + 1,), dtype=torch.int32, device=signed_distances.device)
    counts[-1] = 0
    _native.call(
        "mt_count_triangle_vertices", signed_distances.contiguous().data_ptr(),
        lower_corners.contiguous().data_ptr(), num_columns, level_sets.contiguous().data_ptr(),
        num_level_sets, counts.data_ptr(), _native.stream(signed_distances),
    )
    return counts
