"""The PyTorch-CPU transcription that the CPU baseline times (oracle/cpu_torch.py) against
the NumPy oracle on a small scene: extraction, tessellation, keys, sort and tile runs
bit-exact; the fused composite image within 1e-6 (exp may differ by an ulp); the adjoint
(colour, SDF and sharpness gradients of the L2 image loss) within 1e-4 of the FP64 oracle
adjoint; a subset of tiles composites exactly as in the whole image; SH and vertex adjoints
within 1e-5."""

import numpy as np
import pytest
import torch

from oracle import cpu_torch as ct
from oracle import extraction as ox
from oracle import rasterization as orr
from tests.helpers import OracleScene, small_sphere_spec


@pytest.fixture(scope="module")
def scene():
    return OracleScene(small_sphere_spec(resolution=32, image=48, sh_degree=2))


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def test_extraction_and_tessellation_bit_exact(scene):
    sc = scene
    vox = ct.extract_voxels(_t(sc.packed))
    assert torch.equal(vox.vertices, _t(sc.vox.vertices))
    for got, want in ((vox.neighbors, sc.vox.neighbors), (vox.lower, sc.vox.lower_corners),
                      (vox.upper, sc.vox.upper_corners), (vox.indices, sc.vox.indices)):
        assert np.array_equal(got.numpy(), want)
    tess = ct.tessellate(_t(sc.sdf), _t(sc.world), _t(sc.sh), vox, _t(sc.level_sets))
    t = sc.tri
    assert torch.equal(tess.vertices, _t(t.vertices)) and torch.equal(tess.sdf, _t(t.signed_distances))
    assert torch.equal(tess.sh, _t(t.spherical_harmonics))
    assert np.array_equal(tess.faces.numpy(), t.faces) and t.faces.shape[0] > 1000
    assert np.array_equal(tess.face_boundaries.numpy(), t.face_boundaries)


def test_front_end_bit_exact(scene):
    sc = scene
    xy, depth = ct.project(_t(sc.tri.vertices), _t(sc.E), _t(sc.K))
    assert torch.equal(xy, _t(sc.xy)) and torch.equal(depth, _t(sc.depth))
    sk, sv, ranges = ct.front_end(xy, depth, _t(sc.tri.faces).long(), _t(sc.tri.face_boundaries).long(),
                                  sc.tile_shape, (sc.H, sc.W), sc.spec.near_plane)
    assert np.array_equal(sk.numpy(), sc.sorted_keys)
    assert np.array_equal(sv.numpy(), sc.sorted_values)
    assert np.array_equal(ranges.numpy(), sc.ranges[:, 0])


def _loss_grad(raw, target, bg):
    pre = raw[:, :3] + raw[:, 3:4] * bg[None, :, None, None]
    diff = np.clip(pre, 0, 1) - target
    d_c = 2 * diff / target.size * ((pre >= 0) & (pre <= 1))
    grad = np.zeros_like(raw)
    grad[:, :3] = d_c
    grad[:, 3] = (d_c * bg[None, :, None, None]).sum(1)
    return float((diff.astype(np.float64) ** 2).mean()), grad


def test_fused_composite_forward_and_adjoint(scene):
    sc = scene
    t = sc.tri
    faces = _t(t.faces).long()
    sk, sv = _t(sc.sorted_keys), _t(sc.sorted_values).long()
    ranges = _t(sc.ranges[:, 0])
    rng = np.random.default_rng(3)
    target = rng.random((sc.B, 3, sc.H, sc.W)).astype(np.float32)
    bg = np.asarray(sc.spec.background, np.float32)
    res = ct.fused_composite(_t(sc.xy), faces, _t(t.signed_distances), _t(sc.colors), sk, sv, ranges,
                             sc.tile_shape, (sc.H, sc.W), sc.spec.sharpness, _t(bg), _t(target))
    assert np.abs(res.image.numpy() - sc.image).max() < 1e-6
    loss, grad = _loss_grad(sc.image, target, bg)
    assert abs(res.loss - loss) < 1e-6 * loss
    cg, sg, kg = orr.fused_backward(sc.hits, sc.colors.astype(np.float64),
                                    t.signed_distances.astype(np.float64), t.faces,
                                    sc.spec.sharpness, sc.image, grad)
    for got, want in ((res.colors_grad.numpy(), cg), (res.sdf_grad.numpy(), sg)):
        assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()
    assert abs(res.sharpness_grad - kg) <= 1e-4 * abs(kg)
    assert res.pairs >= sc.hits.pixel.size
    # A subset of tiles: those tiles composite as in the whole image, the rest stay empty.
    picked = torch.tensor([[0, 3], [0, 4], [0, 7]])
    part = ct.fused_composite(_t(sc.xy), faces, _t(t.signed_distances), _t(sc.colors), sk, sv, ranges,
                              sc.tile_shape, (sc.H, sc.W), sc.spec.sharpness, tiles=picked)
    th, tw = sc.tile_shape
    for _, tile in picked.tolist():
        y0, x0 = (tile // sc.cols) * th, (tile % sc.cols) * tw
        assert torch.equal(part.image[0, :, y0 : y0 + th, x0 : x0 + tw],
                           res.image[0, :, y0 : y0 + th, x0 : x0 + tw])
    assert float(part.image[0, 3].min()) < 1 and float(part.image[0, 3, :th, :tw].min()) == 1.0


def test_sh_and_vertex_adjoints(scene):
    sc = scene
    t = sc.tri
    deg = sc.spec.sh_degree
    col, pre = ct.sh_forward(_t(t.vertices), _t(t.spherical_harmonics), _t(sc.E), deg)
    assert np.abs(col.numpy() - sc.colors).max() < 1e-5
    rng = np.random.default_rng(4)
    g_col = rng.standard_normal(col.shape).astype(np.float32)
    g_sh = ct.sh_backward(_t(t.vertices), _t(t.spherical_harmonics), _t(sc.E), deg, pre, _t(g_col))
    want_sh = orr.evaluate_spherical_harmonics_backward(t.vertices, t.spherical_harmonics, sc.E, deg, g_col)
    assert np.abs(g_sh.numpy() - want_sh).max() <= 1e-5 * np.abs(want_sh).max()
    vox = ct.extract_voxels(_t(sc.packed))
    tess = ct.tessellate(_t(sc.sdf), _t(sc.world), _t(sc.sh), vox, _t(sc.level_sets))
    g_d = rng.standard_normal(tess.sdf.shape[0]).astype(np.float32)
    gd, gs = ct.vertex_backward(tess, sc.world.shape[0], _t(g_d), g_sh)
    _, want_d, want_s = ox.create_triangle_vertices_backward(t.record, sc.world.shape[0], sc.sh.shape[0],
                                                             None, g_d, g_sh.numpy())
    assert np.abs(gd.numpy() - want_d).max() <= 1e-5 * np.abs(want_d).max()
    assert np.abs(gs.numpy() - want_s).max() <= 1e-5 * np.abs(want_s).max()
