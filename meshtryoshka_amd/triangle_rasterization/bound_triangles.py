"""Culling and tile bounds (bound_triangles.py:23-71)."""

from dataclasses import dataclass

import torch
from torch import Tensor

from .. import _native
from .common import TileGrid, record_function


@dataclass
class TriangleBounds:
    # Minimum tile indices (inclusive) that each triangle overlaps, as (row, col).
    tile_minima: Tensor  # int32[batch, triangle, 2]
    # Maximum tile indices (exclusive) that each triangle overlaps.
    tile_maxima: Tensor  # int32[batch, triangle, 2]
    # The number of tiles each triangle overlaps.
    num_tiles_overlapped: Tensor  # int32[batch, triangle]


@record_function("bound_triangles")
def bound_triangles(vertices: Tensor, depths: Tensor, faces: Tensor, grid: TileGrid,
                    near_plane: float) -> TriangleBounds:
    assert near_plane >= 0
    _native.require_device(vertices, depths, faces)
    c, v, _ = vertices.shape
    t, _ = faces.shape
    kw = dict(device=vertices.device, dtype=torch.int32)
    tmin = torch.empty((c, t, 2), **kw)
    tmax = torch.empty((c, t, 2), **kw)
    counts = torch.empty((c, t), **kw)
    rows, cols = grid.grid_shape
    _native.call(
        "mt_bound_triangles", vertices.contiguous().data_ptr(), depths.contiguous().data_ptr(), v,
        faces.contiguous().data_ptr(), t, c, grid.tile_shape[0], grid.tile_shape[1], 0, 0, rows,
        cols, float(near_plane), tmin.data_ptr(), tmax.data_ptr(), counts.data_ptr(), None,
        _native.stream(vertices),
    )
    return TriangleBounds(tmin, tmax, counts)
