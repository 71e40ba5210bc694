"""Coarse-to-fine subdivision ops of the reference (write_occupancy.py:25-46,
dilate_occupancy.py:27-48, upscale_occupancy.py:27-53, upscale_parameters.py:34-99,
index_vertices.py:6-51) on the gfx950 kernels of ``csrc/subdivision.hip``."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function
from .compute_exclusive_cumsum import compute_exclusive_cumsum
from .compute_vertex_occupancy import compute_vertex_occupancy


class UpscaledParameters(NamedTuple):
    signed_distances: Tensor  # f32[vertex]
    spherical_harmonics: Tensor  # f32[sh, vertex, 3]
    occupancy: Tensor  # int32[i, j, k_packed]


@record_function("write_occupancy")
def write_occupancy(voxel_indices: Tensor, packed_occupancy_shape: tuple[int, ...],
                    min_shell: int, max_shell: int) -> Tensor:
    """A packed grid of the given shape with the cells of ``voxel_indices`` (linear cell
    indices) set. ``min_shell`` / ``max_shell`` are accepted and unused, as in the
    reference kernel; indices outside the grid are ignored."""
    _native.require_device(voxel_indices)
    voxel_indices = voxel_indices.contiguous()
    occupancy = torch.zeros(packed_occupancy_shape, dtype=torch.int32, device=voxel_indices.device)
    _native.call("mt_write_occupancy", voxel_indices.data_ptr(), voxel_indices.numel(),
                 occupancy.data_ptr(), occupancy.numel(), int(min_shell), int(max_shell),
                 _native.stream(voxel_indices))
    return occupancy


@record_function("dilate_occupancy")
def dilate_occupancy(occupancy: Tensor, dilation: int) -> Tensor:
    """Box dilation of a packed grid by ``dilation`` cells along every axis."""
    if dilation == 0:
        return occupancy.clone()
    assert 0 < dilation < 32  # the bit shifts reach one neighbouring word at most
    _native.require_device(occupancy)
    occupancy = occupancy.contiguous()
    i, j, kp = occupancy.shape
    out = torch.empty_like(occupancy)
    _native.call("mt_dilate_occupancy", occupancy.data_ptr(), i, j, kp, int(dilation),
                 out.data_ptr(), _native.stream(occupancy))
    return out


@record_function("upscale_occupancy")
def upscale_occupancy(occupancy: Tensor, target_shape: tuple[int, int, int]) -> Tensor:
    """A finer packed grid whose cells are occupied where a coarse cell under any of their
    corners is."""
    i, j, kp = occupancy.shape
    i2, j2, k2 = target_shape
    assert k2 % 32 == 0
    kp2 = k2 // 32
    assert i2 >= i and j2 >= j and kp2 >= kp
    _native.require_device(occupancy)
    occupancy = occupancy.contiguous()
    out = torch.empty((i2, j2, kp2), dtype=torch.int32, device=occupancy.device)
    _native.call("mt_upscale_occupancy", occupancy.data_ptr(), i, j, kp, out.data_ptr(), i2, j2,
                 kp2, _native.stream(occupancy))
    return out


def _vertex_grid(occupancy: Tensor):
    vocc, offsets = compute_vertex_occupancy(occupancy)
    compute_exclusive_cumsum(offsets)
    return vocc, offsets


@record_function("upscale_parameters")
@torch.no_grad()
def upscale_parameters(old_signed_distances: Tensor, old_spherical_harmonics: Tensor,
                       old_occupancy: Tensor, occupancy_to_upscale: Tensor,
                       new_resolution: tuple[int, int, int]) -> UpscaledParameters:
    """Upscales ``occupancy_to_upscale`` (restricted to ``old_occupancy``, so that every new
    vertex has an old vertex to interpolate from) to ``new_resolution`` and resamples the
    SDF and SH onto its vertices by occupancy-weighted trilinear interpolation."""
    _native.require_device(old_signed_distances, old_spherical_harmonics, old_occupancy,
                           occupancy_to_upscale)
    device = old_signed_distances.device
    new_occupancy = upscale_occupancy(old_occupancy & occupancy_to_upscale, new_resolution)
    old_vocc, old_off = _vertex_grid(old_occupancy)
    new_vocc, new_off = _vertex_grid(new_occupancy)
    num_vertices = int(new_off[-1].item())
    sh, old_v, _ = old_spherical_harmonics.shape
    new_sdf = torch.empty((num_vertices,), dtype=torch.float32, device=device)
    new_sh = torch.empty((sh, num_vertices, 3), dtype=torch.float32, device=device)
    if num_vertices > 0:
        oi, oj, okp = old_vocc.shape
        ni, nj, nkp = new_vocc.shape
        _native.call("mt_upscale_parameters", old_vocc.data_ptr(), old_off.data_ptr(), oi, oj,
                     okp, old_signed_distances.contiguous().data_ptr(),
                     old_spherical_harmonics.contiguous().data_ptr(), old_v, sh,
                     new_vocc.data_ptr(), new_off.data_ptr(), ni, nj, nkp, new_sdf.data_ptr(),
                     new_sh.data_ptr(), num_vertices, _native.stream(old_signed_distances))
    return UpscaledParameters(new_sdf, new_sh, new_occupancy)


@record_function("index_vertices")
def index_vertices(points: Tensor, tol: float = 1e-5) -> Tensor:
    """int32[V]: for every point, the input index of the first point (in input order) at the
    same ``tol``-quantised position (index_vertices.py:6-30). A coordinate is quantised as
    ``torch.round(points / tol)`` is on the device (a float32 multiply by ``1 / tol`` rounded
    to float32, ties to even). The largest quantised magnitude is read back to size
    the packed key, and a mesh that needs more than 63 key bits raises RuntimeError, as in the
    reference. Coordinates are offset by that magnitude before packing, so negative
    coordinates never alias (the reference packs them unbiased). An empty input gives an
    empty output."""
    _native.require_device(points)
    assert points.dim() == 2 and points.shape[1] == 3 and points.dtype == torch.float32
    points = points.contiguous()
    (n, _) = points.shape
    out = torch.empty((n,), dtype=torch.int32, device=points.device)
    if n == 0:
        return out
    st = _native.stream(points)
    max_abs = torch.empty((), dtype=torch.int64, device=points.device)
    _native.call("mt_quantized_max_abs", points.data_ptr(), n, float(tol), max_abs.data_ptr(), st)
    max_abs = int(max_abs.item())
    bits = max_abs.bit_length() + 1
    if bits * 3 > 63:
        raise RuntimeError(f"need {bits * 3} bits, exceed 63 bits total")
    ws = _native.workspace(_native.query("mt_index_vertices_workspace_bytes", n), points.device)
    _native.call("mt_index_vertices", points.data_ptr(), n, float(tol), max_abs,
                 out.data_ptr(), ws.data_ptr(), ws.numel(), st)
    return out


@record_function("index_keys")
def index_keys(keys: Tensor) -> Tensor:
    """int32[N]: for every int64 key, the input index of its first occurrence in input order
    (index_vertices.py:33-51)."""
    _native.require_device(keys)
    assert keys.dim() == 1 and keys.dtype == torch.int64
    keys = keys.contiguous()
    (n,) = keys.shape
    out = torch.empty((n,), dtype=torch.int32, device=keys.device)
    if n == 0:
        return out
    ws = _native.workspace(_native.query("mt_index_vertices_workspace_bytes", n), keys.device)
    _native.call("mt_index_keys", keys.data_ptr(), n, out.data_ptr(), ws.data_ptr(), ws.numel(),
                 _native.stream(keys))
    return out
