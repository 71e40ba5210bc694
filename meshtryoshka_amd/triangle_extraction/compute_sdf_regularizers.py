"""SDF regularizers of the training step (compute_sdf_regularizers.py:1-139): per voxel, the
Eikonal loss at the voxel centre (split by the sign of |grad d| - 1) and a 7-point
Laplacian curvature loss, with a hand-written adjoint."""

from typing import NamedTuple

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class SDFRegularizerLosses(NamedTuple):
    eikonal_loss_pos: Tensor  # f32[voxel]
    eikonal_loss_neg: Tensor  # f32[voxel]
    curvature_loss: Tensor  # f32[voxel]


def _args(grid_signed_distances, voxel_neighbors, voxel_lower_corners, voxel_upper_corners,
          grid_size):
    i, j, k = grid_size
    _, num_columns = voxel_lower_corners.shape
    _, num_voxels = voxel_upper_corners.shape
    # voxel_x, voxel_y, voxel_z = k, j, i (compute_sdf_regularizers.py:61-66)
    return (grid_signed_distances.data_ptr(), voxel_neighbors.data_ptr(),
            voxel_lower_corners.data_ptr(), num_columns, voxel_upper_corners.data_ptr(),
            num_voxels, int(k), int(j), int(i))


class ComputeSDFRegularizers(torch.autograd.Function):
    """Forward: the three per-voxel losses. Backward: recomputes each voxel's terms and
    scatters d loss / d SDF to the voxel's corners and Laplacian neighbours (float atomics,
    as the reference's generated backward)."""

    @staticmethod
    @record_function("compute_sdf_regularizers_forward")
    def forward(ctx, grid_signed_distances, voxel_neighbors, voxel_lower_corners,
                voxel_upper_corners, grid_size):
        _native.require_device(grid_signed_distances, voxel_neighbors, voxel_lower_corners,
                               voxel_upper_corners)
        tensors = [t.contiguous() for t in (grid_signed_distances, voxel_neighbors,
                                            voxel_lower_corners, voxel_upper_corners)]
        _, num_voxels = tensors[3].shape
        f32 = dict(dtype=torch.float32, device=grid_signed_distances.device)
        pos = torch.empty((num_voxels,), **f32)
        neg = torch.empty((num_voxels,), **f32)
        curvature = torch.empty((num_voxels,), **f32)
        _native.call("mt_compute_sdf_regularizers_forward", *_args(*tensors, grid_size),
                     pos.data_ptr(), neg.data_ptr(), curvature.data_ptr(),
                     _native.stream(grid_signed_distances))
        ctx.save_for_backward(*tensors)
        ctx.grid_size = tuple(grid_size)
        return pos, neg, curvature

    @staticmethod
    @record_function("compute_sdf_regularizers_backward")
    def backward(ctx, grad_pos, grad_neg, grad_curvature):
        tensors = ctx.saved_tensors
        sdf = tensors[0]
        _, num_voxels = tensors[3].shape
        zeros = None

        def grad_or_zero(g):
            nonlocal zeros
            if g is not None:
                return g.contiguous()
            if zeros is None:
                zeros = torch.zeros((num_voxels,), dtype=torch.float32, device=sdf.device)
            return zeros

        gp, gn, gc = (grad_or_zero(g) for g in (grad_pos, grad_neg, grad_curvature))
        grad = torch.zeros_like(sdf)
        _native.call("mt_compute_sdf_regularizers_backward", *_args(*tensors, ctx.grid_size),
                     gp.data_ptr(), gn.data_ptr(), gc.data_ptr(), grad.data_ptr(),
                     _native.stream(sdf))
        return grad, None, None, None, None


def compute_sdf_regularizers(
    grid_signed_distances: Tensor,
    voxel_neighbors: Tensor,
    voxel_lower_corners: Tensor,
    voxel_upper_corners: Tensor,
    grid_size: tuple[int, int, int],
) -> SDFRegularizerLosses:
    """(eikonal_loss_pos, eikonal_loss_neg, curvature_loss), each f32[voxel]
    (compute_sdf_regularizers.py:128-139)."""
    return SDFRegularizerLosses(*ComputeSDFRegularizers.apply(
        grid_signed_distances, voxel_neighbors, voxel_lower_corners, voxel_upper_corners,
        grid_size))
