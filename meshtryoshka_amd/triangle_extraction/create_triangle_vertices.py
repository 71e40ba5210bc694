"""Differentiable triangle-vertex creation (create_triangle_vertices.py:23-189)."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class CreateTriangleVerticesResult(NamedTuple):
    vertices: Tensor  # f32[triangle_vertex, 3]
    signed_distances: Tensor  # f32[triangle_vertex]
    spherical_harmonics: Tensor  # f32[sh, triangle_vertex, 3]
    vertex_types: Tensor  # uint8[triangle_vertex]


class CreateTriangleVertices(torch.autograd.Function):
    """Forward places a vertex on every crossing edge at t = clamp((iso - d0)/(d1 - d0), 0, 1)
    and interpolates position, SDF and SH; t is treated as a constant by the adjoint
    (create_triangle_vertices.slang:92-96), which scatters (1 - t) and t weighted
    gradients back to the two grid vertices of the edge."""

    @staticmethod
    @record_function("create_triangle_vertices_forward")
    def forward(ctx, grid_vertices, grid_signed_distances, grid_spherical_harmonics,
                voxel_lower_corners, level_sets, vertex_offsets):
        _native.require_device(grid_vertices, grid_signed_distances, grid_spherical_harmonics,
                               voxel_lower_corners, level_sets, vertex_offsets)
        device = grid_vertices.device
        (num_level_sets,) = level_sets.shape
        _, num_columns = voxel_lower_corners.shape
        num_sh, num_grid, _ = grid_spherical_harmonics.shape
        num_vertices = int(vertex_offsets[-1].item())
        grid_vertices = grid_vertices.contiguous()
        grid_signed_distances = grid_signed_distances.contiguous()
        grid_spherical_harmonics = grid_spherical_harmonics.contiguous()
        voxel_lower_corners = voxel_lower_corners.contiguous()
        level_sets = level_sets.contiguous()
        vertex_offsets = vertex_offsets.contiguous()
        f32 = dict(dtype=torch.float32, device=device)
        tv = torch.empty((num_vertices, 3), **f32)
        td = torch.empty((num_vertices,), **f32)
        tsh = torch.empty((num_sh, num_vertices, 3), **f32)
        ttype = torch.empty((num_vertices,), dtype=torch.uint8, device=device)
        _native.call(
            "mt_create_triangle_vertices_forward", grid_vertices.data_ptr(),
            grid_signed_distances.data_ptr(), grid_spherical_harmonics.data_ptr(), num_grid,
            num_sh, voxel_lower_corners.data_ptr(), num_columns, level_sets.data_ptr(),
            num_level_sets, vertex_offsets.data_ptr(), num_vertices, tv.data_ptr(), td.data_ptr(),
            tsh.data_ptr(), ttype.data_ptr(), None, _native.stream(grid_vertices),
        )
        ctx.save_for_backward(grid_vertices, grid_signed_distances, grid_spherical_harmonics,
                              voxel_lower_corners, vertex_offsets, level_sets)
        ctx.num_vertices = num_vertices
        ctx.mark_non_differentiable(ttype)
        ctx.set_materialize_grads(False)
        return tv, td, tsh, ttype

    @staticmethod
    @record_function("create_triangle_vertices_backward")
    def backward(ctx, grad_vertices, grad_signed_distances, grad_spherical_harmonics, grad_types):
        (grid_vertices, grid_sdf, grid_sh, lower, vertex_offsets, level_sets) = ctx.saved_tensors
        (num_level_sets,) = level_sets.shape
        _, num_columns = lower.shape
        num_sh, num_grid, _ = grid_sh.shape
        num_vertices = ctx.num_vertices
        need_v, need_d, need_sh = ctx.needs_input_grad[:3]
        gg_v = torch.zeros_like(grid_vertices) if need_v else None
        gg_d = torch.zeros_like(grid_sdf) if need_d else None
        gg_sh = torch.zeros_like(grid_sh) if need_sh else None
        gv = grad_vertices.contiguous() if (need_v and grad_vertices is not None) else None
        gd = grad_signed_distances.contiguous() if (need_d and grad_signed_distances is not None) else None
        gs = grad_spherical_harmonics.contiguous() if (need_sh and grad_spherical_harmonics is not None) else None
        if any(g is not None for g in (gv, gd, gs)):
            ws = _native.workspace(
                _native.query("mt_create_triangle_vertices_backward_workspace_bytes", num_grid),
                grid_sdf.device)
            _native.call(
                "mt_create_triangle_vertices_backward", grid_sdf.data_ptr(), num_grid, num_sh,
                lower.data_ptr(), num_columns, level_sets.data_ptr(), num_level_sets,
                vertex_offsets.data_ptr(), num_vertices, _native.ptr(gv), _native.ptr(gd),
                _native.ptr(gs), _native.ptr(gg_v if gv is not None else None),
                _native.ptr(gg_d if gd is not None else None),
                _native.ptr(gg_sh if gs is not None else None), ws.data_ptr(), ws.numel(),
                _native.stream(grid_sdf),
            )
        return gg_v, gg_d, gg_sh, None, None, None


def create_triangle_vertices(grid_vertices, grid_signed_distances, grid_spherical_harmonics,
                             voxel_lower_corners, level_sets, vertex_offsets) -> CreateTriangleVerticesResult:
    result = CreateTriangleVertices.apply(grid_vertices, grid_signed_distances,
                                          grid_spherical_harmonics, voxel_lower_corners,
                                          level_sets, vertex_offsets)
    return CreateTriangleVerticesResult(*result)
