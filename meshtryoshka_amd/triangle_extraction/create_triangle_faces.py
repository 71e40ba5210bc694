"""Marching-cubes faces (create_triangle_faces.py:23-65).

``face_offsets`` may be the reference's int32 exclusive offsets or int64 ones (meshes
with 2^31 face corners or more, where an int32 scan wraps)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class CreateTriangleFacesResult(NamedTuple):
    faces: Tensor  # int32[face, 3]
    voxel_indices: Tensor | None  # int32[face]


@record_function("create_triangle_faces")
def create_triangle_faces(signed_distances: Tensor, neighbors: Tensor, indices: Tensor,
                          vertex_offsets: Tensor, triangle_vertex_types: Tensor,
                          voxel_cell_codes: Tensor, level_sets: Tensor,
                          face_offsets: Tensor) -> CreateTriangleFacesResult:
    _native.require_device(neighbors, indices, vertex_offsets, triangle_vertex_types,
                           voxel_cell_codes, face_offsets)
    device = signed_distances.device
    _, num_voxels = neighbors.shape
    num_faces = int(face_offsets[-1].item()) // 3
    (num_level_sets,) = level_sets.shape
    num_columns = (vertex_offsets.shape[0] - 1) // num_level_sets
    faces = torch.empty((num_faces, 3), dtype=torch.int32, device=device)
    voxel_indices = torch.empty((num_faces,), dtype=torch.int32, device=device)
    if face_offsets.dtype not in (torch.int32, torch.int64):
        raise RuntimeError("face_offsets must be int32 (the reference's layout) or int64")
    entry = "mt_create_triangle_faces" if face_offsets.dtype == torch.int32 else \
        "mt_create_triangle_faces_i64"
    _native.call(
        entry, neighbors.contiguous().data_ptr(), indices.contiguous().data_ptr(),
        num_voxels, num_columns, vertex_offsets.contiguous().data_ptr(),
        triangle_vertex_types.contiguous().data_ptr(), voxel_cell_codes.contiguous().data_ptr(),
        num_level_sets, face_offsets.contiguous().data_ptr(), faces.data_ptr(),
        voxel_indices.data_ptr(), num_faces, triangle_vertex_types.shape[0], None,
        _native.stream(neighbors),
    )
    return CreateTriangleFacesResult(faces, voxel_indices)
