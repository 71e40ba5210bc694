"""Drop-in replacement for the reference ``triangle_rasterization`` package
(triangle_rasterization/__init__.py:1-329), running on hand-written gfx950 HIP kernels.

``render(..., mode="fused" | "explicit")`` keeps the reference signature and semantics:
OpenCV world-to-camera extrinsics (+Z look, -Y up, +X right), unnormalized pixel
intrinsics, keys packed camera|tile|depth|shell (fused) or camera|shell|tile|depth
(explicit), background composite, SSAA, and the fused path's final clip to [0, 1].
"""

from typing import Literal

import torch
from einops import rearrange, reduce, repeat
from torch import Tensor

from .bound_triangles import TriangleBounds, bound_triangles
from .common import TileGrid, record_function
from .composite import Composite, composite
from .delineate_tiles import delineate_tiles
from .evaluate_spherical_harmonics import EvaluateSphericalHarmonics, evaluate_spherical_harmonics
from .generate_keys import PairedKeys, generate_keys
from .interpolate import Interpolate, interpolate
from .intersect import RenderedImages, intersect
from .intersect_and_composite import IntersectAndComposite, intersect_and_composite
from .keys import create_key_defines
from .project_vertices import ProjectedVertices, project_vertices
from .sort_keys import sort_keys


@record_function("rasterize")
def rasterize(vertices: Tensor, faces: Tensor, shell_face_boundaries: Tensor, extrinsics: Tensor,
              intrinsics: Tensor, image_shape: tuple[int, int], tile_shape: tuple[int, int] = (16, 16),
              near_plane: float = 0.2, num_camera_bits: int = 6, num_shell_bits: int = 6,
              num_tile_bits: int = 20) -> RenderedImages:
    """Per-shell first-hit rasterization (triangle_rasterization/__init__.py:22-94)."""
    h, w = image_shape
    device = vertices.device
    c, _, _ = extrinsics.shape
    (s,) = shell_face_boundaries.shape
    if vertices.numel() == 0 or faces.numel() == 0:
        # The reference's empty path (which misspells float32, __init__.py:48) intends this.
        uvw = torch.zeros((c, s, h, w, 3), device=device, dtype=torch.float32)
        index = torch.zeros((c, s, h, w), device=device, dtype=torch.int32)
        return RenderedImages(uvw, index)
    grid = TileGrid(tile_shape, image_shape, c, s)
    key_defines = create_key_defines(("camera", num_camera_bits), ("shell", num_shell_bits),
                                     ("tile", num_tile_bits), ("depth", 32))
    projected_positions, depths = project_vertices(vertices, extrinsics, intrinsics)
    bounds = bound_triangles(projected_positions, depths, faces, grid, near_plane)
    paired_keys = generate_keys(depths, faces, shell_face_boundaries, bounds.tile_minima,
                                bounds.tile_maxima, bounds.num_tiles_overlapped, grid, key_defines)
    del bounds, depths
    paired_keys = sort_keys(paired_keys, grid)
    tile_boundaries = delineate_tiles(paired_keys, grid, key_defines, True)
    return intersect(projected_positions, faces, paired_keys.triangle_indices, tile_boundaries,
                     grid, key_defines)


@record_function("render_explicit")
def render_explicit(extrinsics: Tensor, intrinsics: Tensor, image_shape: tuple[int, int],
                    vertices: Tensor, faces: Tensor, shell_face_boundaries: Tensor,
                    signed_distances: Tensor, spherical_harmonics: Tensor, sh_degree: int,
                    sharpness: Tensor, background: Tensor, tile_shape: tuple[int, int] = (16, 16),
                    near_plane: float = 0.2, num_camera_bits: int = 6, num_shell_bits: int = 6,
                    num_tile_bits: int = 20, ssaa: int = 1) -> Tensor:
    """triangle_rasterization/__init__.py:97-164 (no final clip, as in the reference)."""
    if ssaa > 1:
        h, w = image_shape
        image_shape = (h * ssaa, w * ssaa)
        multiplier = torch.tensor((ssaa, ssaa, 1), device=intrinsics.device)
        intrinsics = intrinsics * multiplier[:, None]
    uv, index = rasterize(vertices, faces, shell_face_boundaries, extrinsics, intrinsics,
                          image_shape, tile_shape=tile_shape, near_plane=near_plane,
                          num_camera_bits=num_camera_bits, num_shell_bits=num_shell_bits,
                          num_tile_bits=num_tile_bits)
    colors = evaluate_spherical_harmonics(vertices, spherical_harmonics, extrinsics, sh_degree)
    signed_distances, colors = interpolate(uv, index, faces, signed_distances, colors)
    mask = index != -1
    signed_distances = rearrange(signed_distances, "b l h w -> (b h w) l").contiguous()
    colors = rearrange(colors, "b l h w c -> (b h w) l c").contiguous()
    mask = rearrange(mask, "b l h w -> (b h w) l").contiguous()
    image = composite(signed_distances, colors, mask, background, sharpness)
    h, w = image_shape
    b, _, _ = extrinsics.shape
    image = rearrange(image, "(b h w) c -> b c h w", b=b, h=h, w=w)
    if ssaa > 1:
        image = reduce(image, "b c (h mh) (w mw) -> b c h w", "mean", mh=ssaa, mw=ssaa)
    return image


@record_function("render_fused")
def render_fused(extrinsics: Tensor, intrinsics: Tensor, image_shape: tuple[int, int],
                 vertices: Tensor, faces: Tensor, shell_face_boundaries: Tensor,
                 signed_distances: Tensor, spherical_harmonics: Tensor, sh_degree: int,
                 sharpness: Tensor, background: Tensor, tile_shape: tuple[int, int] = (16, 16),
                 near_plane: float = 0.2, ssaa: int = 1, num_camera_bits: int = 6,
                 num_shell_bits: int = 6, num_tile_bits: int = 20) -> Tensor:
    """triangle_rasterization/__init__.py:167-287."""
    b, _, _ = extrinsics.shape
    h, w = image_shape
    if vertices.numel() == 0 or faces.numel() == 0:
        return repeat(background, "c -> b c h w", b=b, h=h, w=w)
    if ssaa > 1:
        image_shape = (h * ssaa, w * ssaa)
        multiplier = torch.tensor((ssaa, ssaa, 1), device=intrinsics.device)
        intrinsics = intrinsics * multiplier[:, None]
    (s,) = shell_face_boundaries.shape
    grid = TileGrid(tile_shape, image_shape, b, s)
    key_defines = create_key_defines(("camera", num_camera_bits), ("tile", num_tile_bits),
                                     ("depth", 32), ("shell", num_shell_bits))
    projected_positions, depths = project_vertices(vertices, extrinsics, intrinsics)
    bounds = bound_triangles(projected_positions, depths, faces, grid, near_plane)
    paired_keys = generate_keys(depths, faces, shell_face_boundaries, bounds.tile_minima,
                                bounds.tile_maxima, bounds.num_tiles_overlapped, grid, key_defines)
    del bounds, depths
    paired_keys = sort_keys(paired_keys, grid)
    tile_boundaries = delineate_tiles(paired_keys, grid, key_defines, False)
    colors = evaluate_spherical_harmonics(vertices, spherical_harmonics, extrinsics, sh_degree)
    _, _, t, _ = tile_boundaries.shape
    images = intersect_and_composite(projected_positions, colors, signed_distances, faces,
                                     sharpness, paired_keys.keys, paired_keys.triangle_indices,
                                     tile_boundaries.view(b, t, 2), s, grid, key_defines)
    images = images[:, :3] + images[:, 3:4] * background[:, None, None]
    if ssaa > 1:
        images = reduce(images, "b c (h mh) (w mw) -> b c h w", "mean", mh=ssaa, mw=ssaa)
    return images.clip(min=0, max=1)


def render(extrinsics: Tensor, intrinsics: Tensor, image_shape: tuple[int, int], vertices: Tensor,
           faces: Tensor, shell_face_boundaries: Tensor, signed_distances: Tensor,
           spherical_harmonics: Tensor, sh_degree: int, sharpness: Tensor, background: Tensor,
           tile_shape: tuple[int, int] = (16, 16), near_plane: float = 0.2,
           num_camera_bits: int = 6, num_shell_bits: int = 6, num_tile_bits: int = 20,
           ssaa: int = 1, mode: Literal["explicit", "fused"] = "fused") -> Tensor:
    """triangle_rasterization/__init__.py:290-329."""
    _render = render_fused if mode == "fused" else render_explicit
    return _render(extrinsics, intrinsics, image_shape, vertices, faces, shell_face_boundaries,
                   signed_distances, spherical_harmonics, sh_degree, sharpness, background,
                   tile_shape=tile_shape, near_plane=near_plane, num_camera_bits=num_camera_bits,
                   num_shell_bits=num_shell_bits, num_tile_bits=num_tile_bits, ssaa=ssaa)


__all__ = [
    "Composite", "EvaluateSphericalHarmonics", "Interpolate", "IntersectAndComposite",
    "PairedKeys", "ProjectedVertices", "RenderedImages", "TileGrid", "TriangleBounds",
    "bound_triangles", "composite", "create_key_defines", "delineate_tiles",
    "evaluate_spherical_harmonics", "generate_keys", "interpolate", "intersect",
    "intersect_and_composite", "project_vertices", "rasterize", "render", "render_explicit",
    "render_fused", "sort_keys",
]
