"""GPU parity of the public render entry points against the end-to-end oracle.

``rasterize`` (triangle_rasterization/__init__.py:22-94), ``render(mode="explicit")``
(:97-164), ``render(mode="fused")`` (:167-287) with ``ssaa`` 1 and 2 (:118-121,
:161-162, :199-203, :277-284), forward and backward, on two cameras of a non-square image
with non-square tiles. Bars: hit indices and barycentrics bit-exact; images within 1e-4
max-abs of the FP32 oracle; SDF and SH gradients within 1e-3 of the FP64 oracle adjoint
(relative to the largest entry). ``Interpolate.backward`` (``mt_interpolate_backward``)
is checked on its own against the oracle's adjoint, which is pinned by finite
differences on the CPU.
"""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import config, workloads
from meshtryoshka_amd import triangle_rasterization as tr
from oracle import rasterization as orr
from oracle import render as ore
from tests.helpers import OracleScene, small_sphere_spec, to_dev

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3
H, W = 40, 56
TILE = (8, 16)


@pytest.fixture(scope="module")
def scene():
    sc = OracleScene(small_sphere_spec(resolution=32, image=32, sh_degree=2))
    # Two cameras of a 40 x 56 image: the spec's camera and one above and to the side.
    e1 = workloads.look_at_w2c(torch.tensor([1.6, 1.2, 1.5]))
    k = workloads.pinhole((H, W), 55.0)
    sc.E2 = np.stack([sc.E[0], e1.numpy()]).astype(np.float32)
    sc.K2 = np.stack([k.numpy(), k.numpy()]).astype(np.float32)
    return sc


def _args(sc):
    t = sc.tri
    return (to_dev(sc.E2), to_dev(sc.K2), (H, W), to_dev(t.vertices), to_dev(t.faces),
            to_dev(t.face_boundaries))


def test_rasterize_matches_oracle(scene):
    sc = scene
    t = sc.tri
    E, K, shape, verts, faces, bounds = _args(sc)
    uv, index = tr.rasterize(verts, faces, bounds, E, K, shape, tile_shape=TILE)
    want_uv, want_index = ore.rasterize(t.vertices, t.faces, t.face_boundaries, sc.E2, sc.K2, (H, W),
                                        TILE)
    assert (want_index >= 0).sum() > 500
    np.testing.assert_array_equal(index.cpu().numpy(), want_index)
    assert np.array_equal(uv.cpu().numpy().view(np.uint32), want_uv.view(np.uint32))


def test_rasterize_empty_scene():
    E = torch.eye(4, device="cuda")[None].repeat(2, 1, 1)
    K = torch.tensor([[8.0, 0, 4], [0, 8, 4], [0, 0, 1]], device="cuda")[None].repeat(2, 1, 1)
    uv, index = tr.rasterize(torch.zeros((0, 3), device="cuda"),
                             torch.zeros((0, 3), dtype=torch.int32, device="cuda"),
                             torch.zeros(3, dtype=torch.int32, device="cuda"), E, K, (6, 5))
    assert uv.shape == (2, 3, 6, 5, 3) and uv.dtype == torch.float32
    assert index.shape == (2, 3, 6, 5) and index.dtype == torch.int32


def test_interpolate_backward_matches_oracle():
    rng = np.random.default_rng(11)
    b, s, h, w, v, f = 2, 3, 37, 29, 500, 800
    faces = rng.integers(0, v, (f, 3)).astype(np.int32)
    index = rng.integers(-1, f, (b, s, h, w)).astype(np.int32)
    uv = rng.uniform(0, 1, (b, s, h, w, 2)).astype(np.float32)
    sdf = rng.standard_normal(v).astype(np.float32)
    colors = rng.random((b, v, 3)).astype(np.float32)
    ts = to_dev(sdf).requires_grad_(True)
    tc = to_dev(colors).requires_grad_(True)
    out_s, out_c = tr.interpolate(to_dev(uv), to_dev(index), to_dev(faces), ts, tc)
    ws, wc = orr.interpolate(uv, index, faces, sdf, colors)
    assert np.abs(out_s.detach().cpu().numpy() - ws).max() < 1e-6
    assert np.abs(out_c.detach().cpu().numpy() - wc).max() < 1e-6
    gs = rng.standard_normal((b, s, h, w)).astype(np.float32)
    gc = rng.standard_normal((b, s, h, w, 3)).astype(np.float32)
    torch.autograd.backward([out_s, out_c], [to_dev(gs), to_dev(gc)])
    want_s, want_c = orr.interpolate_backward(uv, index, faces, v, b, gs, gc)
    for got, want in ((ts.grad, want_s), (tc.grad, want_c)):
        g = got.cpu().numpy().astype(np.float64)
        assert np.abs(g - want).max() / np.abs(want).max() < 1e-5


@pytest.mark.parametrize("mode", ["fused", "explicit"])
@pytest.mark.parametrize("ssaa", [1, 2])
@pytest.mark.parametrize("quirk", [False, True])
def test_render_matches_oracle(scene, mode, ssaa, quirk):
    sc = scene
    t = sc.tri
    E, K, shape, verts, faces, bounds = _args(sc)
    sdf = to_dev(t.signed_distances).requires_grad_(True)
    sh = to_dev(t.spherical_harmonics).requires_grad_(True)
    bg = np.array([0.25, 0.5, 1.0], np.float32)
    sharp = torch.tensor(sc.spec.sharpness, device="cuda")
    img = tr.render(E, K, shape, verts, faces, bounds, sdf, sh, sc.spec.sh_degree, sharp,
                    to_dev(bg), tile_shape=TILE, ssaa=ssaa, mode=mode)
    want, state = ore.render(sc.E2, sc.K2, (H, W), t.vertices, t.faces, t.face_boundaries,
                             t.signed_distances, t.spherical_harmonics, sc.spec.sh_degree,
                             sc.spec.sharpness, bg, tile_shape=TILE, ssaa=ssaa, mode=mode)
    assert img.shape == (2, 3, H, W)
    err = float(np.abs(img.detach().cpu().numpy() - want).max())
    assert err < IMG_TOL, f"image max-abs error {err}"
    g = np.random.default_rng(3).standard_normal(want.shape).astype(np.float32)
    prev = config.reference_gradient_quirks
    config.set_reference_gradient_quirks(quirk)
    try:
        img.backward(to_dev(g))
    finally:
        config.set_reference_gradient_quirks(prev)
    ws, wsh = ore.render_backward(state, t.vertices, t.faces, t.signed_distances,
                                  t.spherical_harmonics, sc.E2, sc.spec.sh_degree,
                                  sc.spec.sharpness, bg, g, reference_quirk=quirk)
    for got, want_g in ((sdf.grad, ws), (sh.grad, wsh)):
        gg = got.cpu().numpy().astype(np.float64)
        scale = np.abs(want_g).max()
        assert scale > 0
        assert np.abs(gg - want_g).max() / scale < GRAD_TOL


def test_render_fused_empty_scene_is_background():
    E = torch.eye(4, device="cuda")[None]
    K = torch.tensor([[8.0, 0, 4], [0, 8, 4], [0, 0, 1]], device="cuda")[None]
    bg = torch.tensor([0.1, 0.2, 0.3], device="cuda")
    img = tr.render(E, K, (6, 5), torch.zeros((0, 3), device="cuda"),
                    torch.zeros((0, 3), dtype=torch.int32, device="cuda"),
                    torch.zeros(2, dtype=torch.int32, device="cuda"), torch.zeros(0, device="cuda"),
                    torch.zeros((1, 0, 3), device="cuda"), 0, torch.tensor(50.0, device="cuda"), bg)
    assert img.shape == (1, 3, 6, 5)
    assert torch.equal(img[0, :, 2, 3], bg)
