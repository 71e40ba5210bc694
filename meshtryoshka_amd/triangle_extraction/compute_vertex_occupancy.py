"""Vertex occupancy of a packed grid (compute_vertex_occupancy.py:26-61)."""

from math import prod
from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class VertexOccupancyFnResult(NamedTuple):
    vertex_occupancy: Tensor  # int32[i+1, j+1, k_packed+1]
    vertex_counts: Tensor  # int32[(i+1)(j+1)(k_packed+1) + 1]


@record_function("compute_vertex_occupancy")
def compute_vertex_occupancy(occupancy: Tensor) -> VertexOccupancyFnResult:
    _native.require_device(occupancy)
    occupancy = occupancy.contiguous()
    i, j, kp = occupancy.shape
    shape = (i + 1, j + 1, kp + 1)
    kwargs = dict(dtype=torch.int32, device=occupancy.device)
    vertex_occupancy = torch.empty(shape, **kwargs)
    vertex_counts = torch.empty((prod(shape) + 1,), **kwargs)
    _native.call(
        "mt_compute_vertex_occupancy", occupancy.data_ptr(), i, j, kp,
        vertex_occupancy.data_ptr(), vertex_counts.data_ptr(), _native.stream(occupancy),
    )
    vertex_counts[-1] = 0
    return VertexOccupancyFnResult(vertex_occupancy, vertex_counts)
