"""SDF regularizers on the GPU against the oracle: the forward losses bit for bit with the
FP32 restatement, the adjoint (through the autograd entry point, with the model's mean
weighting and with random upstream gradients) within 1e-5 of the FP64 adjoint's scale."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import workloads
from oracle import regularizers as orz

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(96, 0), (64, 1)])
def scene(request):
    resolution, seed = request.param
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=seed, num_spheres=8, resolution=resolution, num_cameras=1,
                                  image_shape=(32, 32), device=dev)
    occ = spec.occupancy.to(dev)
    vox = te.extract_voxels(occ)
    _, sdf, _ = workloads.finish_fields(spec, vox.vertices)
    i, j, kp = occ.shape
    return vox, sdf.to(dev).contiguous(), (i, j, 32 * kp)


def _np(vox):
    return (vox.neighbors.cpu().numpy(), vox.lower_corners.cpu().numpy(),
            vox.upper_corners.cpu().numpy())


def test_forward_bit_exact(scene):
    vox, sdf, grid = scene
    out = te.compute_sdf_regularizers(sdf, vox.neighbors, vox.lower_corners, vox.upper_corners, grid)
    want = orz.compute_sdf_regularizers(sdf.cpu().numpy(), *_np(vox), grid)
    for got, ref in zip(out, want):
        assert got.dtype == torch.float32 and got.shape == (vox.upper_corners.shape[1],)
        np.testing.assert_array_equal(got.cpu().numpy(), ref)
    assert float(out.eikonal_loss_pos.sum() + out.eikonal_loss_neg.sum()) > 0
    assert float(out.curvature_loss.sum()) > 0


def test_backward_matches_fp64_adjoint(scene):
    vox, sdf, grid = scene
    n = vox.upper_corners.shape[1]
    gen = torch.Generator("cpu").manual_seed(7)
    ups = [torch.rand(n, generator=gen) for _ in range(3)]
    x = sdf.clone().requires_grad_(True)
    out = te.compute_sdf_regularizers(x, vox.neighbors, vox.lower_corners, vox.upper_corners, grid)
    total = sum((o * u.cuda()).sum() for o, u in zip(out, ups))
    (grad,) = torch.autograd.grad(total, x)
    ref = orz.compute_sdf_regularizers_backward(sdf.cpu().numpy(), *_np(vox), grid,
                                                *(u.numpy() for u in ups))
    err = np.abs(grad.cpu().numpy().astype(np.float64) - ref).max()
    assert err <= 1e-5 * np.abs(ref).max(), err


def test_model_weighting_and_unused_outputs(scene):
    """The reference model takes the mean of each loss (model_meshtryoshka.py:261-276); a loss
    left out of the objective contributes no gradient."""
    vox, sdf, grid = scene
    n = vox.upper_corners.shape[1]
    x = sdf.clone().requires_grad_(True)
    pos, neg, curv = te.compute_sdf_regularizers(x, vox.neighbors, vox.lower_corners,
                                                 vox.upper_corners, grid)
    (0.1 * pos.mean() + 0.1 * neg.mean()).backward()
    ones = np.full(n, 0.1 / n)
    ref = orz.compute_sdf_regularizers_backward(sdf.cpu().numpy(), *_np(vox), grid, ones, ones,
                                                np.zeros(n))
    err = np.abs(x.grad.cpu().numpy().astype(np.float64) - ref).max()
    assert err <= 1e-5 * np.abs(ref).max(), err


def test_no_voxels():
    dev = torch.device("cuda")
    sdf = torch.zeros(5, device=dev, requires_grad=True)
    nb = torch.zeros((7, 0), dtype=torch.int32, device=dev)
    lo = torch.zeros((4, 0), dtype=torch.int32, device=dev)
    up = torch.zeros((4, 0), dtype=torch.int32, device=dev)
    out = te.compute_sdf_regularizers(sdf, nb, lo, up, (1, 1, 32))
    assert all(o.shape == (0,) for o in out)
    sum(o.sum() for o in out).backward()
    assert torch.equal(sdf.grad, torch.zeros(5, device=dev))
