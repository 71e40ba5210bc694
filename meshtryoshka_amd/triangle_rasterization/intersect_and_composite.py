"""Fused per-tile hit testing and NeuS compositing (intersect_and_composite.py:24-175)."""

import torch
from torch import Tensor

from .. import _native, config
from .._native import CompositeArgs
from .common import TileGrid, record_function
from .keys import to_layout


def _args(vertices, colors, signed_distances, faces, sharpness, sorted_keys,
          sorted_triangle_indices, tile_boundaries, num_shells, grid, key_defines) -> CompositeArgs:
    b, v, _ = vertices.shape
    rows, cols = grid.grid_shape
    h, w = grid.image_shape
    th, tw = grid.tile_shape
    return CompositeArgs(
        vertices.data_ptr(), colors.data_ptr(), signed_distances.data_ptr(), v, faces.data_ptr(),
        sharpness.data_ptr(), sorted_keys.data_ptr(), sorted_triangle_indices.data_ptr(),
        tile_boundaries.data_ptr(), b, num_shells, rows, cols, th, tw, h, w,
        to_layout(key_defines), config.composite_chunk,
    )


class IntersectAndComposite(torch.autograd.Function):
    """Forward: per pixel, walk the tile's depth-sorted triangles, accept a hit only if its
    shell is greater than the last accepted one, alpha = clamp((Phi(s_outer) - Phi(s_inner)
    + 1e-5) / (Phi(s_outer) + 1e-5), 0, 1), and stop after the last shell. Output
    [batch, 4, H, W] = (rgb, transmittance). Backward re-traverses (no hit storage) and
    scatters colour/SDF gradients to the triangles' corners and a sharpness gradient."""

    @staticmethod
    @record_function("intersect_and_composite_forward")
    def forward(ctx, vertices, colors, signed_distances, faces, sharpness, sorted_keys,
                sorted_triangle_indices, tile_boundaries, num_shells, grid: TileGrid, key_defines):
        tensors = [t.contiguous() for t in (vertices, colors, signed_distances, faces, sharpness,
                                            sorted_keys, sorted_triangle_indices, tile_boundaries)]
        _native.require_device(*tensors)
        (vertices, colors, signed_distances, faces, sharpness, sorted_keys,
         sorted_triangle_indices, tile_boundaries) = tensors
        b, _, _ = vertices.shape
        h, w = grid.image_shape
        out = torch.empty((b, 4, h, w), dtype=torch.float32, device=vertices.device)
        args = _args(vertices, colors, signed_distances, faces, sharpness, sorted_keys,
                     sorted_triangle_indices, tile_boundaries, num_shells, grid, key_defines)
        _native.call("mt_intersect_and_composite_forward", args, out.data_ptr(),
                     _native.stream(vertices))
        ctx.save_for_backward(vertices, colors, signed_distances, faces, sharpness, sorted_keys,
                              sorted_triangle_indices, tile_boundaries, out)
        ctx.num_shells = num_shells
        ctx.grid = grid
        ctx.key_defines = key_defines
        return out

    @staticmethod
    @record_function("intersect_and_composite_backward")
    def backward(ctx, out_image_grad):
        (vertices, colors, signed_distances, faces, sharpness, sorted_keys,
         sorted_triangle_indices, tile_boundaries, out) = ctx.saved_tensors
        out_image_grad = out_image_grad.contiguous()
        colors_grad = torch.zeros_like(colors)
        sdf_grad = torch.zeros_like(signed_distances)
        sharpness_grad = torch.zeros_like(sharpness)
        args = _args(vertices, colors, signed_distances, faces, sharpness, sorted_keys,
                     sorted_triangle_indices, tile_boundaries, ctx.num_shells, ctx.grid,
                     ctx.key_defines)
        _native.call("mt_intersect_and_composite_backward", args, out.data_ptr(),
                     out_image_grad.data_ptr(), colors_grad.data_ptr(), sdf_grad.data_ptr(),
                     sharpness_grad.data_ptr(), int(config.reference_gradient_quirks),
                     _native.stream(vertices))
        return (None, colors_grad, sdf_grad, None, sharpness_grad, None, None, None, None, None, None)


def intersect_and_composite(vertices, colors, signed_distances, faces, sharpness, sorted_keys,
                            sorted_triangle_indices, tile_boundaries, num_shells: int,
                            grid: TileGrid, key_defines) -> Tensor:
    return IntersectAndComposite.apply(vertices, colors, signed_distances, faces, sharpness,
                                       sorted_keys, sorted_triangle_indices, tile_boundaries,
                                       num_shells, grid, key_defines)
