"""Explicit-mode attribute interpolation (interpolate.py:21-134)."""

import torch
from torch import Tensor

from .. import _native
from .common import record_function


class Interpolate(torch.autograd.Function):
    """(rgb, sdf) = u a + v b + (1 - u - v) c at each rendered pixel; misses give 0."""

    @staticmethod
    @record_function("interpolate_forward")
    def forward(ctx, uv, index, faces, signed_distances, colors):
        uv, index, faces, signed_distances, colors = (
            t.contiguous() for t in (uv, index, faces, signed_distances, colors))
        _native.require_device(uv, index, faces, signed_distances, colors)
        b, s, h, w, _ = uv.shape
        kw = dict(device=uv.device, dtype=torch.float32)
        out_sdf = torch.empty((b * s, h, w), **kw)
        out_col = torch.empty((b * s, h, w, 3), **kw)
        v = signed_distances.shape[0]
        _native.call("mt_interpolate_forward", uv.data_ptr(), index.data_ptr(), faces.data_ptr(),
                     signed_distances.data_ptr(), colors.data_ptr(), v, b, s, h, w,
                     out_sdf.data_ptr(), out_col.data_ptr(), _native.stream(uv))
        ctx.save_for_backward(uv, index, faces, signed_distances, colors)
        return out_sdf.view(b, s, h, w), out_col.view(b, s, h, w, 3)

    @staticmethod
    @record_function("interpolate_backward")
    def backward(ctx, out_signed_distances_grad, out_colors_grad):
        uv, index, faces, signed_distances, colors = ctx.saved_tensors
        b, s, h, w, _ = uv.shape
        if out_signed_distances_grad is None:
            out_signed_distances_grad = torch.zeros((b, s, h, w), device=uv.device)
        if out_colors_grad is None:
            out_colors_grad = torch.zeros((b, s, h, w, 3), device=uv.device)
        gs = out_signed_distances_grad.contiguous()
        gc = out_colors_grad.contiguous()
        sdf_grad = torch.zeros_like(signed_distances)
        col_grad = torch.zeros_like(colors)
        _native.call("mt_interpolate_backward", uv.data_ptr(), index.data_ptr(), faces.data_ptr(),
                     signed_distances.shape[0], b, s, h, w, gs.data_ptr(), gc.data_ptr(),
                     sdf_grad.data_ptr(), col_grad.data_ptr(), _native.stream(uv))
        return None, None, None, sdf_grad, col_grad


def interpolate(uv: Tensor, index: Tensor, faces: Tensor, signed_distances: Tensor, colors: Tensor):
    return Interpolate.apply(uv, index, faces, signed_distances, colors)
