"""Key generation (generate_keys.py:24-70)."""

from dataclasses import dataclass

import torch
from torch import Tensor

from .. import _native
from .common import TileGrid, record_function
from .keys import to_layout


@dataclass
class PairedKeys:
    keys: Tensor  # int64[key]
    triangle_indices: Tensor  # int32[key]

    @property
    def num_keys(self) -> int:
        return self.keys.shape[0]


@record_function("generate_keys")
def generate_keys(depths: Tensor, faces: Tensor, shell_face_boundaries: Tensor, tile_minima: Tensor,
                  tile_maxima: Tensor, num_tiles_overlapped: Tensor, grid: TileGrid,
                  key_defines: tuple[tuple[str, int], ...]) -> PairedKeys:
    """One key per (camera, face, overlapped tile); keys of one (camera, face) are
    contiguous at its inclusive int64 offset, tiles in row-major order."""
    _native.require_device(depths, faces, shell_face_boundaries, tile_minima, tile_maxima,
                           num_tiles_overlapped)
    device = depths.device
    b, v = depths.shape
    num_faces, _ = faces.shape
    counts = num_tiles_overlapped.contiguous().view(-1)
    n = counts.numel()
    offsets = torch.empty((n,), dtype=torch.int64, device=device)
    stream = _native.stream(depths)
    if n > 0:
        ws = _native.workspace(_native.query("mt_scan_workspace_bytes", n), device)
        _native.call("mt_inclusive_scan_i32_to_i64", counts.data_ptr(), offsets.data_ptr(), n,
                     ws.data_ptr(), ws.numel(), stream)
    num_keys = int(offsets[-1].item()) if n > 0 else 0
    out_keys = torch.empty(num_keys, dtype=torch.int64, device=device)
    out_values = torch.empty(num_keys, dtype=torch.int32, device=device)
    if num_keys > 0:
        (s,) = shell_face_boundaries.shape
        _native.call(
            "mt_generate_keys", depths.contiguous().data_ptr(), v, faces.contiguous().data_ptr(),
            num_faces, shell_face_boundaries.contiguous().to(torch.int32).data_ptr(), s,
            tile_minima.contiguous().data_ptr(), tile_maxima.contiguous().data_ptr(),
            offsets.data_ptr(), b, grid.grid_shape[1], to_layout(key_defines), out_keys.data_ptr(),
            out_values.data_ptr(), num_keys, None, stream,
        )
    return PairedKeys(out_keys, out_values)
