"""Occupancy bit packing shared by the extraction ops (triangle_extraction/common.py:17-35).
``ceildiv`` and the profiler-range decorator are the rasterizer's."""

import torch
from torch import Tensor

from ..triangle_rasterization.common import ceildiv as ceildiv
from ..triangle_rasterization.common import record_function as record_function


def pack_occupancy(occupancy: Tensor) -> Tensor:
    """bool[i, j, k] -> int32[i, j, k // 32]; bit b of word (i, j, kp) is cell 32 kp + b
    (triangle_extraction/common.py:17-26)."""
    assert occupancy.shape[-1] % 32 == 0
    i, j, k = occupancy.shape
    bits = occupancy.reshape(i, j, k // 32, 32).to(torch.int64)
    powers = 1 << torch.arange(32, device=occupancy.device, dtype=torch.int64)
    words = (bits * powers).sum(dim=-1)
    # Reinterpret the low 32 bits as int32 (two's complement).
    return torch.where(words >= 2**31, words - 2**32, words).to(torch.int32)


def unpack_occupancy(packed_occupancy: Tensor) -> Tensor:
    """Inverse of :func:`pack_occupancy` (triangle_extraction/common.py:29-35)."""
    occupancy = packed_occupancy.view(torch.int32)
    powers = 1 << torch.arange(32, device=occupancy.device, dtype=torch.int32)
    bits = (occupancy[..., None] & powers) != 0
    i, j, kp = packed_occupancy.shape
    return bits.reshape(i, j, kp * 32)
