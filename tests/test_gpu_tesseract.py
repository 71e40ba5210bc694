"""A reference trainer checkpoint of a seven-component tesseract (32^3 centre, six 32-cell
arms, sphere initializer), committed as tests/golden/tesseract_c32_a32.ckpt and loaded with
weights_only=True, on the GPU against the oracle:

* voxels: each component extracted, mapped through its corners, concatenated and welded;
  integer arrays and vertex positions bit-exact (the grids are dyadic, so the maps are
  exact), the rendered SDF (raw + sphere offset) within 1e-6 of float64;
* tessellation of the welded structure bit-exact against the oracle on the same inputs;
  vertices mapped to world space by the contraction's inverse within 2e-6 (relative) of
  float64; one camera's image within 1e-4 of the FP32 oracle render; the serving path's
  image equal to the plain render's;
* the reference's handling of missing corners (shifted and re-indexed as if they were
  vertex rows, torch wrapping the negative indices) emits extra triangle vertices that no
  face uses: the faces and the image are the same as with the markers kept.
"""

import os

import numpy as np
import pytest
import torch

from meshtryoshka_amd import serving, workloads
from meshtryoshka_amd import tesseract as ts
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import triangle_rasterization as tr
from oracle import extraction as ox
from oracle import render as ore
from oracle import tesseract as ot

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "tesseract_c32_a32.ckpt")
LAYOUT = ts.TesseractLayout({0: 32}, {0: 32}, 0.5)
SPHERE = {"name": "sphere", "radius": 0.25}
CONTRACTION = ts.Contraction(0.6, 1.5, 0.5)
CONFIG = dict(layout=LAYOUT, initializer=SPHERE, sharpness=400.0, shells=[0.9, 0.5, 0.1])
H, W = 48, 64


def _camera():
    e = workloads.look_at_w2c(torch.tensor([3.2, 2.4, 3.0], dtype=torch.float64))
    k = workloads.pinhole((H, W), 55.0)
    return e[None].float(), k[None].float()


def _np(t):
    return t.detach().cpu().numpy()


def test_components_welded_and_sdf():
    c = serving.read_scene_checkpoint(FIXTURE, **CONFIG)
    vox, bounds = ts.extract_tesseract_voxels(c["occupancy"].cuda(), LAYOUT, c["step"])
    want, want_bounds = ot.extract_tesseract(_np(c["occupancy"]), 32, 32, 0.5)
    assert bounds == want_bounds and bounds[-1] == c["signed_distances_raw"].shape[0]
    assert len(bounds) == 8  # the centre and six arms
    assert np.array_equal(_np(vox.vertices), want.vertices)
    for name in ("neighbors", "lower_corners", "upper_corners", "indices"):
        assert np.array_equal(_np(getattr(vox, name)), getattr(want, name)), name
    assert vox.grid_shape == (96, 96, 96)
    used = np.unique(_np(vox.lower_corners))
    assert used[used >= 0].size < bounds[-1]  # components share welded faces
    sdf = ts.sdf_from_raw(c["signed_distances_raw"].cuda(), vox.vertices, SPHERE, c["step"])
    ref = _np(c["signed_distances_raw"]).astype(np.float64) + ot.sphere_offset(want.vertices, 0.25)
    assert np.abs(_np(sdf) - ref).max() < 1e-6


def test_tessellate_and_render_against_the_oracle():
    scene = serving.load_scene_checkpoint(FIXTURE, "cuda", contraction=CONTRACTION, **CONFIG)
    c = serving.read_scene_checkpoint(FIXTURE, **CONFIG)
    vox, _ = ts.extract_tesseract_voxels(c["occupancy"].cuda(), LAYOUT, c["step"])
    sdf = ts.sdf_from_raw(c["signed_distances_raw"].cuda(), vox.vertices, SPHERE, c["step"])
    tri = ox.tessellate_voxels(_np(sdf), _np(vox.vertices), _np(c["spherical_harmonics"]),
                               _np(vox.neighbors), _np(vox.lower_corners), _np(vox.upper_corners),
                               _np(vox.indices), _np(c["level_sets"]))
    assert tri.faces.shape[0] > 1000
    assert np.array_equal(_np(scene.faces), tri.faces)
    assert np.array_equal(_np(scene.face_boundaries), tri.face_boundaries)
    assert np.array_equal(_np(scene.signed_distances), tri.signed_distances)
    assert np.array_equal(_np(scene.spherical_harmonics), tri.spherical_harmonics)
    world = CONTRACTION.uncontract(torch.from_numpy(tri.vertices).cuda())
    assert torch.equal(scene.vertices, world)
    want_world = ot.uncontract(tri.vertices, 0.6, 1.5, 0.5)
    assert np.allclose(_np(world), want_world, rtol=2e-6, atol=2e-6)
    E, K = _camera()
    img = tr.render(E.cuda(), K.cuda(), (H, W), scene.vertices, scene.faces, scene.face_boundaries,
                    scene.signed_distances, scene.spherical_harmonics, 0, scene.sharpness[0],
                    scene.background)
    want, _ = ore.render(_np(E), _np(K), (H, W), _np(world), tri.faces, tri.face_boundaries,
                         tri.signed_distances, tri.spherical_harmonics, 0, np.float32(400.0),
                         np.ones(3, np.float32))
    assert float((want < 0.99).mean()) > 0.05  # the scene is in view
    assert np.abs(_np(img) - want).max() < 1e-4
    served = serving.render_batch(scene, E.cuda(), K.cuda(), [(H, W)]).images[0]
    torch.cuda.synchronize()
    assert torch.equal(served, img[0])


def test_reference_missing_corner_handling_renders_the_same():
    c = serving.read_scene_checkpoint(FIXTURE, **CONFIG)
    occ = c["occupancy"].cuda()
    ours, _ = ts.extract_tesseract_voxels(occ, LAYOUT, c["step"])
    # The reference's concatenation: every lower corner shifted by the vertex rows before,
    # then indexed into the weld map, negative rows wrapping to the end.
    lowers, v0 = [[], []], 0
    for comp, grid in zip(LAYOUT.components(c["step"]), ts.split_occupancy(occ, LAYOUT, c["step"])):
        v = te.extract_voxels(grid.contiguous())
        lowers[0].append(v.lower_corners[:, : v.num_voxels] + v0)
        lowers[1].append(v.lower_corners[:, v.num_voxels :] + v0)
        v0 += v.num_vertices
    shifted = torch.cat(lowers[0] + lowers[1], dim=1).long()
    first = te.index_vertices(ours.vertices, 1e-6).long()
    wrapped = first[torch.where(shifted < 0, shifted + first.numel(), shifted)].to(torch.int32)
    assert bool((ours.lower_corners < 0).any()) and not bool((wrapped < 0).any())
    ref = ours._replace(lower_corners=wrapped.contiguous())
    sdf = ts.sdf_from_raw(c["signed_distances_raw"].cuda(), ours.vertices, SPHERE, c["step"])
    sh = c["spherical_harmonics"].cuda()
    ls = c["level_sets"].cuda()
    E, K = _camera()
    out = []
    for v in (ours, ref):
        t = te.tessellate_voxels(sdf, v.vertices, sh, v.neighbors, v.lower_corners, v.upper_corners,
                                 v.indices, ls)
        verts = CONTRACTION.uncontract(t.vertices)
        img = tr.render(E.cuda(), K.cuda(), (H, W), verts, t.faces, t.face_boundaries,
                        t.signed_distances, t.spherical_harmonics, 0,
                        torch.tensor(400.0, device="cuda"), torch.ones(3, device="cuda"))
        out.append((t, verts, img))
    (a, va, ia), (b, vb, ib) = out
    assert b.vertices.shape[0] > a.vertices.shape[0]  # the wrapped corners add crossings
    assert a.faces.shape[0] == b.faces.shape[0]
    assert torch.equal(va[a.faces.long()], vb[b.faces.long()])  # the same triangles
    assert torch.equal(ia, ib)


def test_batched_step_on_the_welded_tesseract():
    """Columns of different components share welded grid vertices, so the colour adjoint's
    grid-vertex maps are not one-to-one and its rows are added with atomics: the batched
    step's image, loss and grid gradients still match the plain path's (in the unit cube;
    cameras look at its centre)."""
    from meshtryoshka_amd import batched

    c = serving.read_scene_checkpoint(FIXTURE, **CONFIG)
    vox, _ = ts.extract_tesseract_voxels(c["occupancy"].cuda(), LAYOUT, c["step"])
    sdf = ts.sdf_from_raw(c["signed_distances_raw"].cuda(), vox.vertices, SPHERE, c["step"])
    sh = c["spherical_harmonics"].cuda().contiguous()
    ls = c["level_sets"].cuda()
    origins = vox.lower_corners[0]
    assert torch.unique(origins).numel() < origins.numel()  # shared column origins
    centre = (0.5, 0.5, 0.5)
    E = torch.stack([workloads.look_at_w2c(torch.tensor(p, dtype=torch.float64), target=centre)
                     for p in ([1.7, 1.4, 1.9], [-0.6, 1.5, 0.2], [0.4, -0.9, 1.6])]).float().cuda()
    K = workloads.pinhole((H, W), 50.0)[None].repeat(3, 1, 1).cuda()
    gen = torch.Generator("cpu").manual_seed(4)
    target = torch.rand((3, 3, H, W), generator=gen).cuda()
    sharp, bg = 400.0, (1.0, 1.0, 1.0)
    # The plain path, with autograd.
    d = sdf.clone().requires_grad_(True)
    s = sh.clone().requires_grad_(True)
    tri = te.tessellate_voxels(d, vox.vertices, s, vox.neighbors, vox.lower_corners,
                               vox.upper_corners, vox.indices, ls)
    img = tr.render(E, K, (H, W), tri.vertices.detach(), tri.faces, tri.face_boundaries,
                    tri.signed_distances, tri.spherical_harmonics, 0,
                    torch.tensor(sharp, device="cuda"), torch.tensor(bg, device="cuda"))
    loss = ((img - target) ** 2).mean()
    loss.backward()
    assert float((img < 0.99).float().mean()) > 0.05
    cap = batched.measure_capacities(vox, ls, E, K, (H, W), vox.vertices, sdf, sh).scaled(1.2)
    step = batched.BatchedTrainingStep(vox, ls, E, K, (H, W), 1, 0, sharp, bg, cap)
    step.capture(vox.vertices, sdf, sh, target)
    for i in range(2):
        step.replay()
        torch.cuda.synchronize()
        assert not step.overflowed(), i
        rgb = (step.out[:, :3] + step.out[:, 3:4] * step.background[:, None, None]).clip(0, 1)
        assert torch.equal(rgb, img.detach()), i
        assert abs(float(step.loss) - float(loss)) <= 1e-6 * float(loss), i
        for got, want in ((step.grid_sdf_grad, d.grad), (step.grid_sh_grad, s.grad)):
            assert float((got - want).abs().max() / want.abs().max()) < 1e-5, i
