"""Explicit-mode per-pixel compositing over shells (composite.py:22-117)."""

import torch
from torch import Tensor

from .. import _native, config
from .common import record_function


class Composite(torch.autograd.Function):
    @staticmethod
    @record_function("composite_forward")
    def forward(ctx, signed_distances, colors, mask, sharpness):
        signed_distances, colors, mask, sharpness = (
            t.contiguous() for t in (signed_distances, colors, mask, sharpness))
        _native.require_device(signed_distances, colors, mask, sharpness)
        p, s = signed_distances.shape
        kw = dict(dtype=torch.float32, device=signed_distances.device)
        out_colors = torch.empty((p, 3), **kw)
        out_trans = torch.empty((p,), **kw)
        mask_u8 = mask.view(torch.uint8) if mask.dtype == torch.bool else mask.to(torch.uint8)
        _native.call("mt_composite_forward", signed_distances.data_ptr(), colors.data_ptr(),
                     mask_u8.data_ptr(), p, s, sharpness.data_ptr(), out_colors.data_ptr(),
                     out_trans.data_ptr(), _native.stream(signed_distances))
        ctx.save_for_backward(signed_distances, colors, mask_u8, sharpness, out_colors, out_trans)
        return out_colors, out_trans

    @staticmethod
    @record_function("composite_backward")
    def backward(ctx, out_colors_grad, out_transmittances_grad):
        signed_distances, colors, mask_u8, sharpness, out_colors, out_trans = ctx.saved_tensors
        p, s = signed_distances.shape
        if out_colors_grad is None:
            out_colors_grad = torch.zeros_like(out_colors)
        if out_transmittances_grad is None:
            out_transmittances_grad = torch.zeros_like(out_trans)
        gc = out_colors_grad.contiguous()
        gt = out_transmittances_grad.contiguous()
        sdf_grad = torch.zeros_like(signed_distances)
        col_grad = torch.zeros_like(colors)
        sharp_grad = torch.zeros_like(sharpness)
        _native.call("mt_composite_backward", signed_distances.data_ptr(), colors.data_ptr(),
                     mask_u8.data_ptr(), p, s, sharpness.data_ptr(), out_colors.data_ptr(),
                     out_trans.data_ptr(), gc.data_ptr(), gt.data_ptr(), sdf_grad.data_ptr(),
                     col_grad.data_ptr(), sharp_grad.data_ptr(),
                     int(config.reference_gradient_quirks), _native.stream(signed_distances))
        return sdf_grad, col_grad, None, sharp_grad


def composite(signed_distances: Tensor, colors: Tensor, mask: Tensor, background: Tensor,
              sharpness: Tensor) -> Tensor:
    out_colors, out_transmittances = Composite.apply(signed_distances, colors, mask, sharpness)
    return out_colors + background * out_transmittances[:, None]
