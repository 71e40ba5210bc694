"""Writes tests/golden/dense_sphere_small.npz: inputs and oracle outputs of a small
seeded scene (16 x 16 x 32 dense grid, 3 level sets, one 40 x 40 camera, SH degree 1).

The reference implementation cannot run in this environment (its Slang/CUDA toolchain
is absent), so these vectors come from the CPU oracle; they pin the oracle against
silent drift and give the GPU tests a fixed, committed target.
Run: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from meshtryoshka_amd import workloads  # noqa: E402
from oracle import extraction as ox  # noqa: E402
from oracle import rasterization as orr  # noqa: E402

PATH = os.path.join(HERE, "dense_sphere_small.npz")


def scene():
    spec = workloads.dense_sphere(seed=3, resolution=32, image_size=40)
    occ = np.zeros((16, 16, 32), dtype=bool)
    occ[:, :, :] = True
    packed = ox.pack_occupancy(occ)
    vox = ox.extract_voxels(packed)
    gen = torch.Generator("cpu").manual_seed(3)
    unit = torch.from_numpy(vox.vertices)
    p = (2 * unit - 1).numpy().astype(np.float32)
    sdf = (np.linalg.norm(p, axis=1) - 0.6 + 0.03 * torch.randn(p.shape[0], generator=gen).numpy()).astype(np.float32)
    sh = (torch.rand((4, p.shape[0], 3), generator=gen) * 0.5).numpy().astype(np.float32)
    level_sets = np.array([0.08, 0.0, -0.08], np.float32)
    E = spec.cameras.extrinsics.numpy()
    E[0, 2, 3] = 2.2
    K = spec.cameras.intrinsics.numpy()
    return packed, vox, p, sdf, sh, level_sets, E, K, (40, 40)


def compute():
    packed, vox, p, sdf, sh, ls, E, K, shape = scene()
    tri = ox.tessellate_voxels(sdf, p, sh, vox.neighbors, vox.lower_corners, vox.upper_corners,
                               vox.indices, ls)
    H, W = shape
    rows, cols = -(-H // 16), -(-W // 16)
    xy, depth = orr.project_vertices(tri.vertices, E, K)
    tmin, tmax, cnt = orr.bound_triangles(xy, depth, tri.faces, (16, 16), (rows, cols), 0.2)
    lay = orr.fused_layout()
    keys, vals, _ = orr.generate_keys(depth, tri.faces, tri.face_boundaries, tmin, tmax, cnt, cols, lay)
    sk, sv = orr.sort_keys(keys, vals)
    ranges = orr.delineate_tiles(sk, lay, 1, 3, rows * cols, False)
    col, _ = orr.evaluate_spherical_harmonics(tri.vertices, tri.spherical_harmonics, E, 1)
    col = col.astype(np.float32)
    hits = orr.fused_hits(xy, tri.faces, sk, sv, ranges[:, 0], lay, 3, (16, 16), shape)
    img = orr.fused_forward_image(hits, col, tri.signed_distances, tri.faces, 40.0, 1, shape)
    return dict(
        occupancy=packed, sdf=sdf, sh=sh, level_sets=ls, extrinsics=E, intrinsics=K,
        grid_vertices=p, faces=tri.faces, triangle_vertices=tri.vertices,
        face_boundaries=tri.face_boundaries, vertex_boundaries=tri.vertex_boundaries,
        sorted_keys=sk, sorted_values=sv, tile_ranges=ranges, image=img,
        sharpness=np.float32(40.0),
    )


if __name__ == "__main__":
    out = compute()
    np.savez_compressed(PATH, **out)
    print(PATH, os.path.getsize(PATH), "bytes")
