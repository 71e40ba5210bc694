"""Drop-in replacement for the reference ``triangle_extraction`` package
(triangle_extraction/__init__.py:1-157), running on hand-written gfx950 HIP kernels.

Implemented: the voxel-structure builder (``extract_voxels`` and its three kernels), the
multi-level-set marching cubes (``tessellate_voxels`` and its kernels), the SDF
regularizers (``compute_sdf_regularizers``), the coarse-to-fine subdivision ops
(``write_occupancy``, ``dilate_occupancy``, ``upscale_occupancy``, ``upscale_parameters``,
``index_vertices``), the exclusive scan and occupancy packing. ``get_occupancy_mesh`` (a
visualisation helper) is exported under its name but raises ``NotImplementedError``.
"""

import logging
from typing import NamedTuple

import torch
from torch import Tensor

from .common import pack_occupancy as pack_occupancy
from .common import record_function
from .common import unpack_occupancy as unpack_occupancy
from .compute_exclusive_cumsum import compute_exclusive_cumsum, exclusive_offsets_i64
from .compute_sdf_regularizers import SDFRegularizerLosses as SDFRegularizerLosses
from .compute_sdf_regularizers import compute_sdf_regularizers as compute_sdf_regularizers
from .compute_vertex_occupancy import compute_vertex_occupancy
from .count_occupancy import count_occupancy
from .count_triangle_faces import count_triangle_faces
from .count_triangle_vertices import count_triangle_vertices
from .create_triangle_faces import create_triangle_faces
from .create_triangle_vertices import CreateTriangleVertices, create_triangle_vertices
from .create_voxels import Voxels, create_voxels
from .subdivision import UpscaledParameters as UpscaledParameters
from .subdivision import dilate_occupancy as dilate_occupancy
from .subdivision import index_vertices as index_vertices
from .subdivision import index_keys as index_keys
from .subdivision import upscale_occupancy as upscale_occupancy
from .subdivision import upscale_parameters as upscale_parameters
from .subdivision import write_occupancy as write_occupancy


def _not_in_this_build(name: str):
    def fn(*args, **kwargs):
        raise NotImplementedError(
            f"triangle_extraction.{name} is outside the extract -> rasterize -> composite path "
            "implemented by meshtryoshka_amd"
        )

    fn.__name__ = name
    return fn


get_occupancy_mesh = _not_in_this_build("get_occupancy_mesh")


@record_function("extract_voxels")
def extract_voxels(occupancy: Tensor) -> Voxels:
    """Packed occupancy int32[i, j, k/32] -> Voxels (triangle_extraction/__init__.py:31-62)."""
    i, j, k_packed = occupancy.shape
    logging.debug("extract_voxels: %d x %d x %d cell grid", i, j, k_packed * 32)
    vertex_occupancy, vertex_offsets = compute_vertex_occupancy(occupancy)
    compute_exclusive_cumsum(vertex_offsets)
    num_vertices = vertex_offsets[-1].item()
    vertex_offsets = vertex_offsets[:-1].view(vertex_occupancy.shape)
    voxel_offsets = count_occupancy(occupancy)
    compute_exclusive_cumsum(voxel_offsets)
    num_voxels = voxel_offsets[-1].item()
    voxel_offsets = voxel_offsets[:-1].view(vertex_occupancy.shape)
    logging.debug("extract_voxels: %d corners, %d occupied cells", num_vertices, num_voxels)
    return create_voxels(occupancy, voxel_offsets, num_voxels, vertex_occupancy, vertex_offsets,
                         num_vertices)


class Triangles(NamedTuple):
    vertices: Tensor  # f32[vertex, 3]
    signed_distances: Tensor  # f32[vertex]
    spherical_harmonics: Tensor  # f32[sh, vertex, 3]
    faces: Tensor  # int32[face, 3]
    voxel_indices: Tensor  # int32[face]
    vertex_boundaries: Tensor  # int32[level_set]
    face_boundaries: Tensor  # int32[level_set]


@record_function("tessellate_voxels")
def tessellate_voxels(signed_distances: Tensor, vertices: Tensor, spherical_harmonics: Tensor,
                      neighbors: Tensor, lower_corners: Tensor, upper_corners: Tensor,
                      indices: Tensor, level_sets: Tensor) -> Triangles:
    """Nested level-set meshes of the sparse grid (triangle_extraction/__init__.py:75-157).

    Level sets are expected in descending order (shell 0 outermost); each shell's
    vertices and faces occupy one contiguous block, delimited by the cumulative
    ``vertex_boundaries`` / ``face_boundaries``."""
    (num_level_sets,) = level_sets.shape
    logging.debug("tessellate_voxels: %d cells, %d level sets", neighbors.shape[1], num_level_sets)
    with torch.no_grad():
        vertex_offsets = count_triangle_vertices(signed_distances, lower_corners, level_sets)
        compute_exclusive_cumsum(vertex_offsets)
    (
        triangle_vertices,
        triangle_signed_distances,
        triangle_spherical_harmonics,
        triangle_vertex_types,
    ) = create_triangle_vertices(vertices, signed_distances, spherical_harmonics, lower_corners,
                                 level_sets, vertex_offsets)
    with torch.no_grad():
        face_counts, voxel_cell_codes = count_triangle_faces(signed_distances, lower_corners,
                                                             upper_corners, level_sets)
        # Face offsets in int64: the reference scans them in int32 in place
        # (compute_exclusive_cumsum), which wraps once a mesh has 2^31 face corners
        # (716M faces); below that the offsets, and so every output, are the same.
        face_offsets = exclusive_offsets_i64(face_counts)
        faces, voxel_indices = create_triangle_faces(signed_distances, neighbors, indices,
                                                     vertex_offsets, triangle_vertex_types,
                                                     voxel_cell_codes, level_sets, face_offsets)
    vertex_boundaries = vertex_offsets[1:].view(num_level_sets, -1)[:, -1].clone()
    face_boundaries = (face_offsets[1:].view(num_level_sets, -1)[:, -1] // 3).to(torch.int32)
    return Triangles(triangle_vertices, triangle_signed_distances, triangle_spherical_harmonics,
                     faces, voxel_indices, vertex_boundaries, face_boundaries)


__all__ = [
    "Triangles", "Voxels", "CreateTriangleVertices", "UpscaledParameters",
    "compute_exclusive_cumsum", "compute_sdf_regularizers", "compute_vertex_occupancy",
    "count_occupancy", "count_triangle_faces", "count_triangle_vertices",
    "create_triangle_faces", "create_triangle_vertices", "create_voxels", "dilate_occupancy",
    "extract_voxels", "get_occupancy_mesh", "index_keys", "index_vertices", "pack_occupancy",
    "tessellate_voxels", "unpack_occupancy", "upscale_occupancy", "upscale_parameters",
    "write_occupancy",
]
