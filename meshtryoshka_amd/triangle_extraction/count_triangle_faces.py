"""Cube codes and face-vertex counts (count_triangle_faces.py:23-63)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class CountTriangleFacesResult(NamedTuple):
    # Field names follow count_triangle_faces.py:23-25 while the values are returned in the
    # order (face_counts, cell_codes) (count_triangle_faces.py:63): callers unpack
    # positionally, so the pairing is kept as the reference has it.
    cell_codes: Tensor
    face_counts: Tensor


@record_function("count_triangle_faces")
def count_triangle_faces(signed_distances: Tensor, lower_corners: Tensor, upper_corners: Tensor,
                         level_sets: Tensor) -> CountTriangleFacesResult:
    _native.require_device(signed_distances, lower_corners, upper_corners, level_sets)
    device = lower_corners.device
    _, num_voxels = upper_corners.shape
    _, num_columns = lower_corners.shape
    (num_level_sets,) = level_sets.shape
    codes = torch.empty((num_level_sets, num_voxels), dtype=torch.uint8, device=device)
    counts = torch.empty((num_level_sets * num_voxels + 1,), dtype=torch.int32, device=device)
    counts[-1] = 0
    _native.call(
        "mt_count_triangle_faces", signed_distances.contiguous().data_ptr(),
        lower_corners.contiguous().data_ptr(), num_columns, upper_corners.contiguous().data_ptr(),
        num_voxels, level_sets.contiguous().data_ptr(), num_level_sets, counts.data_ptr(),
        codes.data_ptr(), _native.stream(lower_corners),
    )
    return CountTriangleFacesResult(counts, codes)
