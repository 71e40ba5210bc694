"""Tile ranges of the sorted keys (delineate_tiles.py:24-55)."""

import torch
from torch import Tensor

from .. import _native
from .common import TileGrid, record_function
from .generate_keys import PairedKeys
from .keys import to_layout


@record_function("delineate_tiles")
def delineate_tiles(sorted_paired_keys: PairedKeys, grid: TileGrid,
                    key_defines: tuple[tuple[str, int], ...], separate_shells: bool,
                    epsilon: float = 1e-8) -> Tensor:
    """int64[batch, shell or 1, tile, 2] of [start, end) key ranges; empty tiles are (0, 0).
    ``epsilon`` is accepted and unused, as in the reference."""
    keys = sorted_paired_keys.keys
    _native.require_device(keys)
    shape = [grid.num_cameras, grid.num_shells, grid.num_tiles_per_image, 2]
    if not separate_shells:
        shape[1] = 1
    tile_boundaries = torch.zeros(shape, dtype=torch.int64, device=keys.device)
    n = sorted_paired_keys.num_keys
    if n > 0:
        _native.call("mt_delineate_tiles", keys.contiguous().data_ptr(), n, to_layout(key_defines),
                     int(separate_shells), shape[0], shape[1], shape[2],
                     tile_boundaries.data_ptr(), None, None, _native.stream(keys))
    return tile_boundaries
