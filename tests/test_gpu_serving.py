"""Continuous-batching render server: a request's image is the same alone and in a mixed
batch, and equals the plain render API's image."""

import pytest
import torch

from meshtryoshka_amd import serving
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import triangle_rasterization as tr
from meshtryoshka_amd import workloads
from meshtryoshka_amd.tesseract import Contraction, sdf_from_raw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    dev = torch.device("cuda")
    out = {}
    for sid, n in enumerate((4, 12)):
        spec = workloads.sphere_field(seed=sid, num_spheres=n, resolution=128, num_cameras=1,
                                      image_shape=(64, 64), device=dev)
        vox = te.extract_voxels(spec.occupancy.to(dev))
        world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
        sc = serving.tessellate_scene(vox, world.to(dev), sdf.to(dev), sh.to(dev),
                                      spec.level_sets.to(dev), spec.sh_degree, spec.sharpness,
                                      spec.background)
        cams = spec.cameras.to(dev)
        sc.keys_per_pixel = serving.calibrate_key_density(sc, cams.extrinsics, cams.intrinsics, (64, 64))
        out[sid] = sc
    return out


def _poses(n, seed, shape):
    gen = torch.Generator("cpu").manual_seed(seed)
    return workloads.hemisphere_cameras(gen, n, shape)


def _request(rid, sid, pose):
    e, k, shape = pose
    return serving.Request(arrival=0.0, rid=rid, scene_id=sid, extrinsics=e, intrinsics=k,
                           image_shape=shape)


def test_same_image_alone_and_in_mix(scenes):
    target_pose = _poses(1, 7, (64, 96))[0]
    server = serving.RenderServer(scenes, hbm_budget_bytes=8 << 30, max_batch=4, keep_images=True)
    alone = _request(0, 1, target_pose)
    server.submit(alone)
    server.drain()
    torch.cuda.synchronize()
    # Scene 1 requests at two other resolutions and the same one: batched together with
    # the target (per-camera resolution table); scene 0 requests in their own batches.
    mix = [_request(1, 1, p) for p in _poses(2, 8, (64, 96))]
    mix += [_request(10, 0, p) for p in _poses(3, 9, (48, 48))]
    mix += [_request(20, 1, p) for p in _poses(2, 10, (40, 72))]
    mix += [_request(30, 1, p) for p in _poses(1, 12, (97, 61))]
    mixed = _request(99, 1, target_pose)
    mix.insert(4, mixed)
    server2 = serving.RenderServer(scenes, hbm_budget_bytes=8 << 30, max_batch=8, keep_images=True)
    for r in mix:
        server2.submit(r)
    server2.drain()
    torch.cuda.synchronize()
    # Scene 1's six cameras in one mixed-resolution batch (a batch whose keys overflowed the
    # calibrated capacity is rendered again with the exact count, not split).
    assert server2.batches - server2.rerenders == 2
    assert torch.equal(alone.image, mixed.image)
    # Every request of the mix equals its own render alone.
    for r in mix:
        solo = serving.RenderServer(scenes, hbm_budget_bytes=8 << 30, max_batch=1, keep_images=True)
        again = _request(r.rid, r.scene_id, (r.extrinsics, r.intrinsics, r.image_shape))
        solo.submit(again)
        solo.drain()
        assert torch.equal(again.image, r.image), (r.rid, r.image_shape)
    # And both equal the plain API render of that one camera.
    sc = scenes[1]
    e, k, shape = target_pose
    img = tr.render(e[None].cuda(), k[None].cuda(), shape, sc.vertices, sc.faces, sc.face_boundaries,
                    sc.signed_distances, sc.spherical_harmonics, sc.sh_degree, sc.sharpness[0],
                    sc.background)
    assert torch.equal(img[0], alone.image)


def test_budget_limits_inflight_work(scenes):
    server = serving.RenderServer(scenes, hbm_budget_bytes=1, max_batch=2, keep_images=False)
    for i, p in enumerate(_poses(5, 11, (64, 64))):
        server.submit(_request(i, 0, p))
    server.schedule()
    # With no memory to spare, only the head request is admitted.
    assert len(server.inflight) == 1 and server.queued == 4
    server.drain()
    assert len(server.completed) == 5


def test_scene_checkpoint_renders_the_same(tmp_path):
    """A scene written with save_scene_checkpoint and loaded back (voxels extracted and shells
    tessellated on the GPU, unit-cube vertices mapped to world space by a contraction whose
    inverse is the generator's 2 v - 1) renders the same image as the scene built in memory.
    With the sphere initializer the file stores SDF - offset, and a reader gets the SDF back
    to float32 rounding."""
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=3, num_spheres=5, resolution=64, num_cameras=1,
                                  image_shape=(48, 64), device=dev)
    occ = spec.occupancy.to(dev)
    vox = te.extract_voxels(occ)
    _, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    to_world = lambda v: 2 * v - 1  # noqa: E731 (the generator's world frame)
    mem = serving.tessellate_scene(vox, vox.vertices, sdf.to(dev), sh.to(dev),
                                   spec.level_sets.to(dev), spec.sh_degree, spec.sharpness,
                                   spec.background)
    mem.vertices = to_world(mem.vertices).contiguous()
    path = tmp_path / "scene.ckpt"
    serving.save_scene_checkpoint(path, occ, sdf, sh, spec.level_sets, spec.sh_degree,
                                  spec.sharpness, spec.background, initializer={"name": "none"},
                                  contraction=Contraction(1.0, 0.0, 1.0))
    disk = serving.load_scene_checkpoint(path, dev)
    assert torch.equal(disk.faces, mem.faces) and torch.equal(disk.vertices, mem.vertices)
    e, k, shape = _poses(1, 21, (48, 64))[0]
    imgs = [serving.render_batch(s, e[None].cuda(), k[None].cuda(), [shape]).images[0]
            for s in (mem, disk)]
    torch.cuda.synchronize()
    assert torch.equal(imgs[0], imgs[1])
    assert float((imgs[0] < 0.999).float().mean()) > 0.01  # the scene is in view
    sphere = tmp_path / "sphere.ckpt"
    serving.save_scene_checkpoint(sphere, occ, sdf, sh, spec.level_sets, spec.sh_degree,
                                  spec.sharpness, spec.background,
                                  initializer={"name": "sphere", "radius": 0.25})
    c = serving.read_scene_checkpoint(sphere)
    stored = c["signed_distances_raw"].to(dev)
    assert (stored - sdf.to(dev)).abs().max() > 0.1  # the offset was taken out
    back = sdf_from_raw(stored, vox.vertices, c["initializer"], c["step"])
    assert (back - sdf.to(dev)).abs().max() < 1e-6


def test_grid_colours_render_the_same(scenes):
    """A resident scene that keeps the grid SH as rows and colours each batch from them
    renders the same bits as one that keeps every triangle vertex's SH, in less memory."""
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=1, num_spheres=12, resolution=128, num_cameras=1,
                                  image_shape=(64, 64), device=dev)
    vox = te.extract_voxels(spec.occupancy.to(dev))
    world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    args = (vox, world.to(dev), sdf.to(dev), sh.to(dev), spec.level_sets.to(dev), spec.sh_degree,
            spec.sharpness, spec.background)
    per_vertex = serving.tessellate_scene(*args)
    grid = serving.tessellate_scene(*args, colours="grid")
    assert grid.spherical_harmonics is None and grid.grid is not None
    assert torch.equal(grid.vertices, per_vertex.vertices) and torch.equal(grid.faces, per_vertex.faces)
    assert grid.resident_bytes() < per_vertex.resident_bytes()
    poses = _poses(3, 31, (64, 96)) + _poses(2, 32, (40, 72))
    E = torch.stack([p[0] for p in poses]).cuda()
    K = torch.stack([p[1] for p in poses]).cuda()
    shapes = [p[2] for p in poses]
    a = serving.render_batch(per_vertex, E, K, shapes, key_capacity=1 << 20)
    b = serving.render_batch(grid, E, K, shapes, key_capacity=1 << 20)
    torch.cuda.synchronize()
    assert int(a.overflow) == 0 and int(b.overflow) == 0
    for x, y in zip(a.images, b.images):
        assert torch.equal(x, y)


def test_arrivals_while_batches_are_in_flight(scenes):
    """A second wave of requests submitted after the first wave's batches were launched
    (and before they were retired) is admitted onto the free streams at once; polling
    never blocks; every image equals the same request rendered alone."""
    server = serving.RenderServer(scenes, hbm_budget_bytes=8 << 30, max_batch=4, num_streams=4,
                                  keep_images=True)
    wave1 = [_request(i, i % 2, p) for i, p in enumerate(_poses(6, 40, (64, 96)))]
    for r in wave1:
        server.submit(r)
    launched1 = server.schedule()
    assert launched1 == 2 and len(server.inflight) == 2  # one batch per scene
    busy = sum(not job.event.query() for job in server.inflight)
    wave2 = [_request(100 + i, i % 2, p) for i, p in enumerate(_poses(4, 41, (48, 80)))]
    for r in wave2:
        server.submit(r)
    launched2 = server.schedule()
    # Wave 2 took the two free streams while wave 1 was still counted in flight.
    assert launched2 == 2 and len(server.inflight) == 4
    print(f"wave 1 batches still running on the device when wave 2 was launched: {busy} of 2")
    while server.queued or server.inflight:
        server.poll()
        server.schedule()
    assert len(server.completed) == 10
    for r in wave1 + wave2:
        solo = serving.RenderServer(scenes, hbm_budget_bytes=8 << 30, max_batch=1, keep_images=True)
        again = _request(r.rid, r.scene_id, (r.extrinsics, r.intrinsics, r.image_shape))
        solo.submit(again)
        solo.drain()
        assert torch.equal(again.image, r.image), r.rid
