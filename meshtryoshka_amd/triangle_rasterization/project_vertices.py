"""Vertex projection (project_vertices.py:23-56)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class ProjectedVertices(NamedTuple):
    # Projected 2D vertex locations in pixel space (unnormalized).
    positions: Tensor  # f32[batch, vertex, 2]
    # Camera-space depths for each vertex.
    depths: Tensor  # f32[batch, vertex]


@record_function("project_vertices")
def project_vertices(vertices: Tensor, extrinsics: Tensor, intrinsics: Tensor) -> ProjectedVertices:
    _native.require_device(vertices, extrinsics, intrinsics)
    b, _, _ = extrinsics.shape
    v, _ = vertices.shape
    kw = dict(device=vertices.device, dtype=torch.float32)
    out_xy = torch.empty((b, v, 2), **kw)
    out_d = torch.empty((b, v), **kw)
    _native.call(
        "mt_project_vertices", vertices.contiguous().data_ptr(), v,
        extrinsics.contiguous().float().data_ptr(), intrinsics.contiguous().float().data_ptr(), b,
        out_xy.data_ptr(), out_d.data_ptr(), _native.stream(vertices),
    )
    return ProjectedVertices(out_xy, out_d)
