"""View-dependent colour from SH (evaluate_spherical_harmonics.py:23-111)."""

import torch
from torch import Tensor

from .. import _native
from .common import record_function


def _camera_positions(extrinsics: Tensor) -> Tensor:
    b = extrinsics.shape[0]
    out = torch.empty((b, 3), dtype=torch.float32, device=extrinsics.device)
    _native.call("mt_camera_positions", extrinsics.contiguous().float().data_ptr(), b,
                 out.data_ptr(), _native.stream(extrinsics))
    return out


class EvaluateSphericalHarmonics(torch.autograd.Function):
    """colors[b, v] = max(sum_k Y_k(normalize(v - c_b)) sh[k, v] + 0.5, 0), with the
    camera centre c_b = inverse(E_b)[:3, 3]. Gradients flow to the SH coefficients only."""

    @staticmethod
    @record_function("evaluate_spherical_harmonics_forward")
    def forward(ctx, vertices: Tensor, spherical_harmonics: Tensor, extrinsics: Tensor,
                sh_degree: int) -> Tensor:
        _native.require_device(vertices, spherical_harmonics, extrinsics)
        b = extrinsics.shape[0]
        v, _ = vertices.shape
        vertices = vertices.contiguous()
        spherical_harmonics = spherical_harmonics.contiguous()
        camera_positions = _camera_positions(extrinsics)
        out = torch.empty((b, v, 3), device=vertices.device, dtype=torch.float32)
        num_sh = spherical_harmonics.shape[0]
        _native.call("mt_evaluate_sh_forward", vertices.data_ptr(), v,
                     spherical_harmonics.data_ptr(), num_sh, sh_degree,
                     camera_positions.data_ptr(), b, out.data_ptr(), _native.stream(vertices))
        ctx.save_for_backward(vertices, spherical_harmonics, camera_positions)
        ctx.sh_degree = sh_degree
        return out

    @staticmethod
    @record_function("evaluate_spherical_harmonics_backward")
    def backward(ctx, out_colors_grad: Tensor):
        vertices, spherical_harmonics, camera_positions = ctx.saved_tensors
        out_colors_grad = out_colors_grad.contiguous()
        v, _ = vertices.shape
        b = camera_positions.shape[0]
        grad = torch.zeros_like(spherical_harmonics)
        _native.call("mt_evaluate_sh_backward", vertices.data_ptr(), v,
                     spherical_harmonics.data_ptr(), spherical_harmonics.shape[0], ctx.sh_degree,
                     camera_positions.data_ptr(), b, out_colors_grad.data_ptr(), grad.data_ptr(),
                     _native.stream(vertices))
        return None, grad, None, None


def evaluate_spherical_harmonics(vertices: Tensor, spherical_harmonics: Tensor, extrinsics: Tensor,
                                 sh_degree: int) -> Tensor:
    return EvaluateSphericalHarmonics.apply(vertices, spherical_harmonics, extrinsics, sh_degree)
