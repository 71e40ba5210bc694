"""The batched and serving configurations at their stated sizes.

* C3: the sphere-field generator at 1024^3 with 256 spheres (40M occupied cells), 7 shells,
  32 cameras whose resolutions are drawn from {320x240, 640x480, 1237x822, 1920x1080,
  2473x1643}, rendered in camera groups of 4 inside one captured step. The mesh has more
  than 2^31 face corners, so the step runs on int64 face offsets. One camera of every
  resolution drawn: the batched step's image equals the plain per-op render's.
* C4: resident scenes of the same generator at 1024^3 (8 and 256 spheres, SH degree 3, 5
  shells), colours computed from the grid per batch, requests at {256^2, 512^2, 1024x768,
  1920x1080}: every image is the same alone and in the mix, and equals the plain render.
"""

import gc

import pytest
import torch

from meshtryoshka_amd import batched, serving, workloads
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import triangle_rasterization as tr

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

C3_RESOLUTIONS = [(240, 320), (480, 640), (822, 1237), (1080, 1920), (1643, 2473)]


def _free():
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_c3_batched_step_matches_plain_render_per_camera():
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=0, num_spheres=256, resolution=1024, num_shells=7,
                                  num_cameras=32, device=dev, resolutions=C3_RESOLUTIONS)
    vox = te.extract_voxels(spec.occupancy.to(dev))
    world, sdf, sh = (t.to(dev) for t in workloads.finish_fields(spec, vox.vertices))
    cams = spec.cameras.to(dev)
    shapes = batched.camera_shapes(cams.image_shape, 32)
    num_pixels = sum(h * w for h, w in shapes)
    target = workloads.target_image(spec, (3 * num_pixels,)).to(dev)
    ls = spec.level_sets.to(dev)
    cap = batched.measure_capacities(vox, ls, cams.extrinsics, cams.intrinsics, shapes, world, sdf,
                                     sh, tile_shape=(8, 8), camera_group=4).scaled(1.1)
    step = batched.BatchedTrainingStep(vox, ls, cams.extrinsics, cams.intrinsics, shapes, 16, 3,
                                       spec.sharpness, spec.background, cap, tile_shape=(8, 8),
                                       chunk=128, camera_group=4)
    step.capture(world, sdf, sh, target)
    step.replay()
    torch.cuda.synchronize()
    assert not step.overflowed()
    live = step.live_sizes()
    print(f"C3: {vox.num_voxels} cells, {live}, peak "
          f"{torch.cuda.max_memory_allocated() / 2**30:.1f} GiB")
    assert 3 * live["F"] >= 2**31  # the int64 face offsets are exercised
    # One camera of each resolution drawn.
    picks = {}
    for b, hw in enumerate(shapes):
        picks.setdefault(hw, b)
    bg = step.background[:, None, None]
    want = {b: (step.image(b)[:3] + step.image(b)[3:4] * bg).clip(0, 1).clone()
            for b in picks.values()}
    del step
    _free()
    with torch.no_grad():
        tri = te.tessellate_voxels(sdf, world, sh, vox.neighbors, vox.lower_corners,
                                   vox.upper_corners, vox.indices, ls)
    assert tri.faces.shape[0] == live["F"] and tri.vertices.shape[0] == live["V_t"]
    for hw, b in sorted(picks.items()):
        with torch.no_grad():
            img = tr.render(cams.extrinsics[b : b + 1], cams.intrinsics[b : b + 1], hw,
                            tri.vertices, tri.faces, tri.face_boundaries, tri.signed_distances,
                            tri.spherical_harmonics, 3, torch.tensor(spec.sharpness, device=dev),
                            torch.tensor(spec.background, device=dev), tile_shape=(8, 8))
        assert torch.equal(img[0], want[b]), (b, hw)
        print(f"C3 camera {b} ({hw[1]}x{hw[0]}): batched image identical to the plain render")
        del img
        _free()
    del tri
    _free()


def test_c4_scale_serving_alone_vs_mix_and_plain():
    dev = torch.device("cuda")
    scenes, specs = {}, {}
    for sid, n in enumerate((8, 256)):
        spec = workloads.sphere_field(seed=100 + sid, num_spheres=n, resolution=1024, num_shells=5,
                                      num_cameras=1, image_shape=(512, 512), device=dev)
        vox = te.extract_voxels(spec.occupancy.to(dev))
        world, sdf, sh = (t.to(dev) for t in workloads.finish_fields(spec, vox.vertices))
        sc = serving.tessellate_scene(vox, world, sdf, sh, spec.level_sets.to(dev), 3,
                                      spec.sharpness, spec.background, colours="grid")
        gen = torch.Generator("cpu").manual_seed(sid)
        cal = {}
        for shape in ((256, 256), (1080, 1920)):
            poses = workloads.hemisphere_cameras(gen, 2, shape)
            cal[shape] = (torch.stack([p[0] for p in poses]).to(dev),
                          torch.stack([p[1] for p in poses]).to(dev))
        serving.calibrate_keys(sc, cal)
        scenes[sid] = sc
        specs[sid] = (spec, vox, world, sdf, sh)
        print(f"C4 scene {sid} ({n} spheres): {vox.num_voxels} cells, "
              f"{sc.faces.shape[0]} faces, resident {sc.resident_bytes() / 2**30:.1f} GiB")
        del vox
    gen = torch.Generator("cpu").manual_seed(5)
    res = [(256, 256), (512, 512), (768, 1024), (1080, 1920)]

    def request(rid, sid, pose):
        return serving.Request(arrival=0.0, rid=rid, scene_id=sid, extrinsics=pose[0],
                               intrinsics=pose[1], image_shape=pose[2])

    mix = []
    for i in range(8):
        pose = workloads.hemisphere_cameras(gen, 1, res[i % 4])[0]
        mix.append(request(i, i % 2, pose))
    server = serving.RenderServer(scenes, hbm_budget_bytes=96 << 30, max_batch=16,
                                  keep_images=True)
    for r in mix:
        server.submit(r)
    server.drain()
    torch.cuda.synchronize()
    assert len(server.completed) == 8
    for r in mix:
        solo = serving.RenderServer(scenes, hbm_budget_bytes=96 << 30, max_batch=1,
                                    keep_images=True)
        again = request(r.rid, r.scene_id, (r.extrinsics, r.intrinsics, r.image_shape))
        solo.submit(again)
        solo.drain()
        assert torch.equal(again.image, r.image), (r.rid, r.image_shape)
        del solo
    print("C4 eight requests: every image the same alone and in the mix")
    # The plain render (per-vertex SH) of the largest request of each scene.
    for sid in (0, 1):
        spec, vox_, world, sdf, sh = specs[sid]
        vox = te.extract_voxels(spec.occupancy.to(dev))
        with torch.no_grad():
            tri = te.tessellate_voxels(sdf, world, sh, vox.neighbors, vox.lower_corners,
                                       vox.upper_corners, vox.indices, spec.level_sets.to(dev))
        r = [q for q in mix if q.scene_id == sid][-1]
        with torch.no_grad():
            img = tr.render(r.extrinsics[None].to(dev), r.intrinsics[None].to(dev), r.image_shape,
                            tri.vertices, tri.faces, tri.face_boundaries, tri.signed_distances,
                            tri.spherical_harmonics, 3, torch.tensor(spec.sharpness, device=dev),
                            torch.tensor(spec.background, device=dev))
        assert torch.equal(img[0], r.image), sid
        del tri, img, vox
        _free()
    del server, scenes
    _free()
