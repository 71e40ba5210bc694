"""Shared builders for the parity tests: one seeded scene, run through the oracle and
(on a GPU) through the HIP path."""

from __future__ import annotations

import numpy as np
import torch

from meshtryoshka_amd import workloads
from oracle import extraction as ox
from oracle import rasterization as orr


def small_sphere_spec(resolution=16, image=32, seed=0, sh_degree=0, level_sets=(0.06, 0.0, -0.06)):
    """A scaled-down dense sphere: resolution^3 cells, image x image pixels."""
    spec = workloads.dense_sphere(seed=seed, resolution=resolution, image_size=image)
    spec.level_sets = torch.tensor(level_sets, dtype=torch.float32)
    if sh_degree:
        nsh = (sh_degree + 1) ** 2
        base = spec.field_fn

        def field_fn(unit, gen):
            p, sdf, sh0 = base(unit, gen)
            extra = torch.randn((nsh - 1, p.shape[0], 3), generator=gen) * 0.1
            return p, sdf, torch.cat([sh0, extra], 0).contiguous()

        spec.field_fn = field_fn
        spec.sh_degree = sh_degree
    return spec


class OracleScene:
    """Everything the oracle computes for a scene, FP32-exact."""

    def __init__(self, spec, tile_shape=(16, 16), explicit=False):
        self.spec = spec
        self.packed = spec.occupancy.numpy()
        self.vox = ox.extract_voxels(self.packed)
        w, s, h = workloads.finish_fields(spec, torch.from_numpy(self.vox.vertices))
        self.world, self.sdf, self.sh = w.numpy(), s.numpy(), h.numpy()
        self.level_sets = spec.level_sets.numpy()
        v = self.vox
        self.tri = ox.tessellate_voxels(self.sdf, self.world, self.sh, v.neighbors, v.lower_corners,
                                        v.upper_corners, v.indices, self.level_sets)
        cams = spec.cameras
        self.E = cams.extrinsics.numpy()
        self.K = cams.intrinsics.numpy()
        self.H, self.W = cams.image_shape
        self.B = self.E.shape[0]
        self.S = self.level_sets.shape[0]
        self.tile_shape = tile_shape
        th, tw = tile_shape
        self.rows, self.cols = -(-self.H // th), -(-self.W // tw)
        self.T = self.rows * self.cols
        t = self.tri
        self.xy, self.depth = orr.project_vertices(t.vertices, self.E, self.K)
        self.tmin, self.tmax, self.counts = orr.bound_triangles(
            self.xy, self.depth, t.faces, tile_shape, (self.rows, self.cols), spec.near_plane)
        self.layout = orr.explicit_layout() if explicit else orr.fused_layout()
        self.keys, self.values, self.offsets = orr.generate_keys(
            self.depth, t.faces, t.face_boundaries, self.tmin, self.tmax, self.counts, self.cols,
            self.layout)
        self.sorted_keys, self.sorted_values = orr.sort_keys(self.keys, self.values)
        self.ranges = orr.delineate_tiles(self.sorted_keys, self.layout, self.B, self.S, self.T,
                                          explicit)
        col, _ = orr.evaluate_spherical_harmonics(t.vertices, t.spherical_harmonics, self.E,
                                                  spec.sh_degree)
        self.colors = col.astype(np.float32)
        if not explicit:
            self.hits = orr.fused_hits(self.xy, t.faces, self.sorted_keys, self.sorted_values,
                                       self.ranges[:, 0], self.layout, self.S, tile_shape,
                                       (self.H, self.W))
            self.image = orr.fused_forward_image(self.hits, self.colors, t.signed_distances, t.faces,
                                                 spec.sharpness, self.B, (self.H, self.W))


def to_dev(x, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)
