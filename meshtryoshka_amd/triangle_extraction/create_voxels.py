"""Voxel structure (create_voxels.py:25-96)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class Voxels(NamedTuple):
    vertices: Tensor  # f32[vertex, 3]
    neighbors: Tensor  # int32[7, voxel]
    lower_corners: Tensor  # int32[4, voxel_and_subvoxel]
    upper_corners: Tensor  # int32[4, voxel]
    indices: Tensor  # int32[voxel]
    grid_shape: tuple[int, int, int]

    @property
    def num_vertices(self) -> int:
        return self.vertices.shape[0]

    @property
    def num_voxels(self) -> int:
        return self.upper_corners.shape[1]

    @property
    def num_subvoxels(self) -> int:
        return self.lower_corners.shape[1] - self.upper_corners.shape[1]


@record_function("create_voxels")
def create_voxels(
    occupancy: Tensor,
    voxel_offsets: Tensor,
    num_voxels: int,
    vertex_occupancy: Tensor,
    vertex_offsets: Tensor,
    num_vertices: int,
) -> Voxels:
    _native.require_device(occupancy, voxel_offsets, vertex_occupancy, vertex_offsets)
    device = occupancy.device
    i, j, kp = occupancy.shape
    vertices = torch.empty((num_vertices, 3), dtype=torch.float32, device=device)
    neighbors = torch.empty((7, num_voxels), dtype=torch.int32, device=device)
    lower_corners = torch.empty((4, num_vertices), dtype=torch.int32, device=device)
    upper_corners = torch.empty((4, num_voxels), dtype=torch.int32, device=device)
    indices = torch.empty((num_voxels,), dtype=torch.int32, device=device)
    if num_voxels > 0:
        _native.call(
            "mt_create_voxels",
            occupancy.contiguous().data_ptr(), voxel_offsets.contiguous().data_ptr(),
            vertex_occupancy.contiguous().data_ptr(), vertex_offsets.contiguous().data_ptr(),
            i, j, kp, num_voxels, num_vertices, vertices.data_ptr(), neighbors.data_ptr(),
            lower_corners.data_ptr(), upper_corners.data_ptr(), indices.data_ptr(),
            _native.stream(occupancy),
        )
    return Voxels(vertices, neighbors, lower_corners, upper_corners, indices, (i, j, kp * 32))
