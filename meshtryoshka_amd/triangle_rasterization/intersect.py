"""Explicit-mode first-hit rasterization (intersect.py:25-70)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import TileGrid, record_function


class RenderedImages(NamedTuple):
    # barycentric coordinates (u, v) of the triangle that was hit
    uv: Tensor  # f32[batch, shell, H, W, 2]
    # the index of the triangle that was hit (-1 = none)
    index: Tensor  # int32[batch, shell, H, W]


@record_function("intersect")
def intersect(vertices: Tensor, faces: Tensor, sorted_triangle_indices: Tensor,
              tile_boundaries: Tensor, grid: TileGrid, key_defines) -> RenderedImages:
    _native.require_device(vertices, faces, sorted_triangle_indices, tile_boundaries)
    device = vertices.device
    h, w = grid.image_shape
    b, s = grid.num_cameras, grid.num_shells
    out_uv = torch.empty((b * s, h, w, 2), dtype=torch.float32, device=device)
    out_index = torch.empty((b * s, h, w), dtype=torch.int32, device=device)
    th, tw = grid.tile_shape
    rows, cols = grid.grid_shape
    _native.call(
        "mt_intersect", sorted_triangle_indices.contiguous().data_ptr(),
        tile_boundaries.contiguous().data_ptr(), vertices.contiguous().data_ptr(), vertices.shape[1],
        faces.contiguous().data_ptr(), b, s, rows, cols, th, tw, h, w, out_uv.data_ptr(),
        out_index.data_ptr(), _native.stream(vertices),
    )
    return RenderedImages(uv=out_uv.view(b, s, h, w, 2), index=out_index.view(b, s, h, w))
