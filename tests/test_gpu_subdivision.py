"""Subdivision ops on the GPU against the oracle, bit for bit: write / dilate / upscale
occupancy (integers) and the SDF / SH resampling of upscale_parameters (FP32 in the
reference's operation order), at integer and non-integer upscaling ratios; vertex welding
(index_vertices, index_keys) against first-seen brute force."""

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


def _welding_points(seed, n, spread):
    """Points repeated in runs that cross the kernel's 2048-position tiles, negative
    coordinates, ties at the quantisation step, and near-duplicates (within tol)."""
    rng = np.random.default_rng(seed)
    base = (rng.random((spread, 3)) * 2 - 1).astype(np.float32)
    pts = base[rng.integers(0, spread, n)]
    jitter = rng.random(n) < 0.2
    pts[jitter] += rng.normal(0, 2e-7, (int(jitter.sum()), 3)).astype(np.float32)
    ties = rng.random(n) < 0.05
    pts[ties] = (rng.integers(-50, 50, (int(ties.sum()), 3)) + 0.5).astype(np.float32) * 1e-5
    return pts


def test_quantisation_matches_torch_on_device():
    """The oracle's quantisation is torch.round(points / tol) as torch computes it on the
    device, including the points where float32 division and multiplication by the float32
    reciprocal round differently."""
    rng = np.random.default_rng(11)
    pts = rng.random((400_000, 3)).astype(np.float32) * 2 - 1
    for tol in (1e-5, 1e-6, 1e-3):
        want = osd.quantize(pts, tol)
        got = torch.round(torch.from_numpy(pts).cuda() / tol).to(torch.int64).cpu().numpy()
        differs = np.round(pts / np.float32(tol)).astype(np.int64) != want
        assert differs.any()  # the points tell the two formulas apart
        assert np.array_equal(got, want), tol


@pytest.mark.parametrize("n,spread,tol", [(1, 1, 1e-5), (2048, 3, 1e-5), (2049, 700, 1e-6),
                                          (300_000, 40_000, 1e-5), (300_000, 2, 1e-6)])
def test_index_vertices_on_device(n, spread, tol):
    pts = _welding_points(n + spread, n, spread)
    got = te.index_vertices(torch.from_numpy(pts).cuda(), tol)
    assert got.dtype == torch.int32
    assert np.array_equal(got.cpu().numpy(), osd.index_vertices(pts, tol))


def test_index_vertices_one_point_repeated():
    """A single run over every tile: each tile finds the head in the first tile."""
    pts = np.tile(np.float32([[0.25, -0.5, 0.125]]), (100_000, 1))
    got = te.index_vertices(torch.from_numpy(pts).cuda(), 1e-5)
    assert not bool(got.any())


def test_index_vertices_limits():
    dev = torch.device("cuda")
    assert te.index_vertices(torch.empty((0, 3), device=dev), 1e-5).shape == (0,)
    with pytest.raises(RuntimeError, match="exceed 63 bits total"):
        te.index_vertices(torch.tensor([[0.0, 0.0, 1e3]], device=dev), 1e-5)  # |q| = 1e8: 3 x 28 bits
    with pytest.raises(RuntimeError, match="exceed 63 bits total"):
        te.index_vertices(torch.tensor([[0.0, float("nan"), 0.0]], device=dev), 1e-5)
    # 1e3 / 1e-2 = 1e5 < 2^17: 18 bits a field, 54 in all, accepted.
    got = te.index_vertices(torch.tensor([[1e3, 0, 0], [-1e3, 0, 0], [1e3, 0, 0]], device=dev), 1e-2)
    assert got.tolist() == [0, 1, 0]


@pytest.mark.parametrize("n,distinct", [(1, 1), (2047, 5), (2048, 2048), (4097, 1),
                                        (500_000, 1000), (500_000, 400_000)])
def test_index_keys_on_device(n, distinct):
    rng = np.random.default_rng(n + distinct)
    pool = rng.integers(-(2**63), 2**63 - 1, distinct, dtype=np.int64, endpoint=True)
    keys = pool[rng.integers(0, distinct, n)]
    got = te.index_keys(torch.from_numpy(keys).cuda())
    assert np.array_equal(got.cpu().numpy(), osd.index_keys(keys))


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
