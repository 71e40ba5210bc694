"""The batched, graph-captured training step against the plain per-op API path:
same image, same loss, same grid gradients (atomic-order noise only), across several
graph replays with eager work in between."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import batched
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import triangle_rasterization as tr
from meshtryoshka_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=1, num_spheres=8, resolution=128, num_cameras=3,
                                  image_shape=(96, 128), device=dev)
    vox = te.extract_voxels(spec.occupancy.to(dev))
    world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    cams = spec.cameras.to(dev)
    target = workloads.target_image(spec, (3, 3, 96, 128)).to(dev)
    return spec, vox, world.to(dev), sdf.to(dev), sh.to(dev), cams, target


def plain_loss_and_grads(spec, vox, world, sdf, sh, cams, target):
    sdf = sdf.clone().requires_grad_(True)
    sh = sh.clone().requires_grad_(True)
    ls = spec.level_sets.to(sdf.device)
    tri = te.tessellate_voxels(sdf, world, sh, vox.neighbors, vox.lower_corners, vox.upper_corners,
                               vox.indices, ls)
    img = tr.render(cams.extrinsics, cams.intrinsics, cams.image_shape, tri.vertices.detach(),
                    tri.faces, tri.face_boundaries, tri.signed_distances, tri.spherical_harmonics,
                    spec.sh_degree, torch.tensor(spec.sharpness, device=sdf.device),
                    torch.tensor(spec.background, device=sdf.device))
    loss = ((img - target) ** 2).mean()
    loss.backward()
    return loss.detach(), img.detach(), sdf.grad, sh.grad


def make_step(scene, margin=1.2, **kw):
    spec, vox, world, sdf, sh, cams, target = scene
    ls = spec.level_sets.to(sdf.device)
    cap = batched.measure_capacities(vox, ls, cams.extrinsics, cams.intrinsics, cams.image_shape,
                                     world, sdf, sh, cull_keys=kw.get("cull_keys", True)
                                     ).scaled(margin)
    return batched.BatchedTrainingStep(vox, ls, cams.extrinsics, cams.intrinsics, cams.image_shape,
                                       sh.shape[0], spec.sh_degree, spec.sharpness, spec.background,
                                       cap, **kw)


def _rel(a, b):
    return float((a - b).abs().max() / (b.abs().max() + 1e-30))


def test_batched_eager_matches_plain(scene):
    spec, vox, world, sdf, sh, cams, target = scene
    loss, img, gsdf, gsh = plain_loss_and_grads(*scene)
    step = make_step(scene)
    step.run(world, sdf, sh, target)
    assert not step.overflowed()
    rgb = (step.out[:, :3] + step.out[:, 3:4] * step.background[:, None, None]).clip(0, 1)
    assert torch.equal(rgb, img)
    assert abs(float(step.loss) - float(loss)) <= 1e-6 * float(loss)
    assert _rel(step.grid_sdf_grad, gsdf) < 1e-5
    assert _rel(step.grid_sh_grad, gsh) < 1e-5


@pytest.mark.parametrize("faces_beside_colours", [True, False])
def test_graph_replays_match_eager(scene, faces_beside_colours):
    """Both placements of the faces pass: beside the vertex and colour pass on a second
    stream (crossing-axis bytes), or after it (column words) beside the projection."""
    spec, vox, world, sdf, sh, cams, target = scene
    step = make_step(scene, faces_beside_colours=faces_beside_colours)
    step.run(world, sdf, sh, target)
    ref_out = step.out.clone()
    ref_sdf = step.grid_sdf_grad.clone()
    ref_sh = step.grid_sh_grad.clone()
    ref_loss = step.loss.clone()
    static = [t.clone() for t in (world, sdf, sh, target)]
    step.capture(*static)
    for i in range(4):
        step.replay()
        torch.cuda.synchronize()
        assert torch.equal(step.out, ref_out), i
        assert _rel(step.grid_sdf_grad, ref_sdf) < 1e-5, i
        # The vertex gradient buffers are cleared by each step's adjoint, not refilled.
        assert _rel(step.grid_sh_grad, ref_sh) < 1e-5, i
        assert abs(float(step.loss) - float(ref_loss)) <= 1e-6 * abs(float(ref_loss)), i
        # Eager work between replays.
        _ = (static[1] * 2).sum().item()
    # Replays follow the static inputs.
    static[1].add_(0.001)
    step.replay()
    other = make_step(scene, faces_beside_colours=not faces_beside_colours)
    other.run(static[0], static[1], static[2], static[3])
    assert torch.equal(step.out, other.out)


def test_overflow_is_reported_not_written(scene):
    spec, vox, world, sdf, sh, cams, target = scene
    step = make_step(scene, margin=0.5)
    step.run(world, sdf, sh, target)
    torch.cuda.synchronize()
    assert step.overflowed()


def test_culled_keys_change_nothing(scene):
    """Keys only for tiles that meet a face's candidate pixel rectangle: far fewer keys,
    the same images, the same loss, the same gradients (summation order aside)."""
    spec, vox, world, sdf, sh, cams, target = scene
    full = make_step(scene, cull_keys=False)
    full.run(world, sdf, sh, target)
    culled = make_step(scene)
    culled.run(world, sdf, sh, target)
    assert not full.overflowed() and not culled.overflowed()
    k_full, k_culled = full.live_sizes()["K"], culled.live_sizes()["K"]
    assert 0 < k_culled < k_full
    assert torch.equal(culled.out, full.out)
    assert abs(float(culled.loss) - float(full.loss)) <= 1e-6 * float(full.loss)
    assert _rel(culled.grid_sdf_grad, full.grid_sdf_grad) < 1e-5
    assert _rel(culled.grid_sh_grad, full.grid_sh_grad) < 1e-5


def test_camera_groups_match_one_group(scene):
    """Rendering the cameras in groups (one after another inside the step) gives the same
    images and loss, and grid gradients equal up to float summation order."""
    spec, vox, world, sdf, sh, cams, target = scene
    one = make_step(scene)
    one.run(world, sdf, sh, target)
    for group in (1, 2):
        step = make_step(scene, camera_group=group)
        assert len(step.groups) == -(-3 // group)
        step.run(world, sdf, sh, target)
        assert not step.overflowed()
        assert torch.equal(step.out, one.out), group
        assert abs(float(step.loss) - float(one.loss)) <= 1e-6 * float(one.loss)
        assert _rel(step.grid_sdf_grad, one.grid_sdf_grad) < 1e-5
        assert _rel(step.grid_sh_grad, one.grid_sh_grad) < 1e-5


@pytest.mark.parametrize("camera_group", [None, 2])
@pytest.mark.parametrize("tile_shape", [(16, 16), (8, 8), (32, 8)])
def test_mixed_resolutions_match_plain_per_camera(scene, tile_shape, camera_group):
    """Cameras of different resolutions (odd sizes, partial tiles) in one step: every
    camera's image equals the plain path's render of that camera alone, the loss is the mean
    over all cameras' pixels, and the grid gradients match its autograd (summation order
    only). Replays of the captured step reproduce the eager step."""
    spec, vox, world, sdf, sh, cams, _ = scene
    dev = sdf.device
    shapes = [(96, 128), (61, 83), (130, 47)]
    intr = torch.stack([workloads.pinhole(s) for s in shapes]).to(dev)
    gen = torch.Generator("cpu").manual_seed(3)
    targets = [torch.rand((3, h, w), generator=gen).to(dev) for h, w in shapes]
    flat_target = torch.cat([t.reshape(-1) for t in targets])
    ls = spec.level_sets.to(dev)

    # Plain path, camera by camera, one mean over all pixels.
    sdf_p = sdf.clone().requires_grad_(True)
    sh_p = sh.clone().requires_grad_(True)
    tri = te.tessellate_voxels(sdf_p, world, sh_p, vox.neighbors, vox.lower_corners,
                               vox.upper_corners, vox.indices, ls)
    total = sum(h * w for h, w in shapes)
    loss = 0.0
    imgs = []
    for c, (h, w) in enumerate(shapes):
        img = tr.render(cams.extrinsics[c : c + 1], intr[c : c + 1], (h, w), tri.vertices.detach(),
                        tri.faces, tri.face_boundaries, tri.signed_distances,
                        tri.spherical_harmonics, spec.sh_degree,
                        torch.tensor(spec.sharpness, device=dev),
                        torch.tensor(spec.background, device=dev), tile_shape=tile_shape)
        imgs.append(img.detach()[0])
        loss = loss + ((img[0] - targets[c]) ** 2).sum()
    loss = loss / (3 * total)
    loss.backward()

    cap = batched.measure_capacities(vox, ls, cams.extrinsics, intr, shapes, world, sdf, sh,
                                     tile_shape=tile_shape, camera_group=camera_group).scaled(1.2)
    step = batched.BatchedTrainingStep(vox, ls, cams.extrinsics, intr, shapes, sh.shape[0],
                                       spec.sh_degree, spec.sharpness, spec.background, cap,
                                       tile_shape=tile_shape, camera_group=camera_group)
    step.run(world, sdf, sh, flat_target)
    assert not step.overflowed()
    bg = step.background[:, None, None]
    for c in range(len(shapes)):
        o = step.image(c)
        rgb = (o[:3] + o[3:4] * bg).clip(0, 1)
        assert torch.equal(rgb, imgs[c]), c
    assert abs(float(step.loss) - float(loss)) <= 1e-6 * float(loss)
    assert _rel(step.grid_sdf_grad, sdf_p.grad) < 1e-5
    assert _rel(step.grid_sh_grad, sh_p.grad) < 1e-5

    ref_out = step.out.clone()
    ref_sdf = step.grid_sdf_grad.clone()
    static = [t.clone() for t in (world, sdf, sh, flat_target)]
    step.capture(*static)
    for i in range(3):
        step.replay()
        torch.cuda.synchronize()
        assert torch.equal(step.out, ref_out), i
        assert _rel(step.grid_sdf_grad, ref_sdf) < 1e-5, i
        _ = (static[1] * 2).sum().item()
