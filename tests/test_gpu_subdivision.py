"""Subdivision ops on the GPU against the oracle, bit for bit: write / dilate / upscale
occupancy (integers) and the SDF / SH resampling of upscale_parameters (FP32 in the
reference's operation order), at integer and non-integer upscaling ratios; index_vertices on
device tensors."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import triangle_extraction as te
from oracle import subdivision as osd
from oracle.extraction import pack_occupancy

pytestmark = pytest.mark.gpu


def _occ(seed, shape, p):
    rng = np.random.default_rng(seed)
    occ = rng.random(shape) < p
    occ[shape[0] // 4 : shape[0] // 2, 2:6, 40:80] = True
    return pack_occupancy(occ)


def test_write_occupancy():
    rng = np.random.default_rng(0)
    shape = (10, 12, 4)
    idx = rng.integers(-20, 10 * 12 * 128 + 20, 5000).astype(np.int32)
    got = te.write_occupancy(torch.from_numpy(idx).cuda(), shape, 0, 1)
    assert np.array_equal(got.cpu().numpy(), osd.write_occupancy(idx, shape))


@pytest.mark.parametrize("d", [1, 4, 31])
def test_dilate_occupancy(d):
    packed = _occ(1, (9, 11, 128), 0.01)
    got = te.dilate_occupancy(torch.from_numpy(packed).cuda(), d)
    assert np.array_equal(got.cpu().numpy(), osd.dilate_occupancy(packed, d))
    assert torch.equal(te.dilate_occupancy(torch.from_numpy(packed).cuda(), 0).cpu(),
                       torch.from_numpy(packed))


@pytest.mark.parametrize("target", [(9, 11, 128), (18, 22, 256), (13, 17, 224)])
def test_upscale_occupancy(target):
    packed = _occ(2, (9, 11, 128), 0.05)
    got = te.upscale_occupancy(torch.from_numpy(packed).cuda(), target)
    assert np.array_equal(got.cpu().numpy(), osd.upscale_occupancy(packed, target))


@pytest.mark.parametrize("target,nsh", [((16, 20, 128), 16), ((13, 17, 96), 4), ((8, 10, 64), 1)])
def test_upscale_parameters(target, nsh):
    old = _occ(3, (8, 10, 64), 0.15)
    to = _occ(4, (8, 10, 64), 0.5)
    v, _ = osd.vertex_occupancy(old)
    n = int(v.sum())
    rng = np.random.default_rng(5)
    sdf = rng.normal(0, 0.1, n).astype(np.float32)
    sh = rng.normal(0, 0.3, (nsh, n, 3)).astype(np.float32)
    got = te.upscale_parameters(torch.from_numpy(sdf).cuda(), torch.from_numpy(sh).cuda(),
                                torch.from_numpy(old).cuda(), torch.from_numpy(to).cuda(), target)
    want_sdf, want_sh, want_occ = osd.upscale_parameters(sdf, sh, old, to, target)
    assert np.array_equal(got.occupancy.cpu().numpy(), want_occ)
    assert got.signed_distances.shape[0] == want_sdf.shape[0] > 0
    np.testing.assert_array_equal(got.signed_distances.cpu().numpy(), want_sdf)
    np.testing.assert_array_equal(got.spherical_harmonics.cpu().numpy(), want_sh)


def test_index_vertices_on_device():
    rng = np.random.default_rng(6)
    base = rng.random((300, 3)).astype(np.float32)
    pts = base[rng.integers(0, 300, 2000)]
    got = te.index_vertices(torch.from_numpy(pts).cuda(), 1e-6)
    assert np.array_equal(got.cpu().numpy(), osd.index_vertices(pts, 1e-6))


def test_empty_inputs():
    dev = torch.device("cuda")
    empty_idx = torch.empty(0, dtype=torch.int32, device=dev)
    assert not bool(te.write_occupancy(empty_idx, (2, 3, 1), 0, 1).any())
    zero = torch.zeros((4, 4, 2), dtype=torch.int32, device=dev)
    assert not bool(te.dilate_occupancy(zero, 3).any())
    assert not bool(te.upscale_occupancy(zero, (8, 8, 128)).any())
    full = torch.full((4, 4, 2), -1, dtype=torch.int32, device=dev)
    v = (5 * 5 * 65)
    up = te.upscale_parameters(torch.zeros(v, device=dev), torch.zeros((1, v, 3), device=dev),
                               full, zero, (8, 8, 128))
    assert up.signed_distances.shape == (0,) and up.spherical_harmonics.shape == (1, 0, 3)
    assert not bool(up.occupancy.any())
