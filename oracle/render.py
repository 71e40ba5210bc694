"""End-to-end oracle renders: ``render_fused`` / ``render_explicit`` and their adjoints.

Test infrastructure only (see oracle/__init__.py). These compose the per-op oracle
functions of :mod:`oracle.rasterization` in the order of the reference's
orchestration, triangle_rasterization/__init__.py:

* ``rasterize`` (:22-94): project, bound, explicit keys camera|shell|tile|depth, sort,
  delineate per shell, first hit per (camera, shell, pixel);
* ``render_explicit`` (:97-164): SSAA scales rows 0 and 1 of the intrinsics
  (:118-121), SH colours, interpolate, composite with the background, SSAA mean
  (:161-162), **no clip**;
* ``render_fused`` (:167-287): empty scene -> background (:196-197), fused keys
  camera|tile|depth|shell, per-pixel traversal, ``rgb + T * background`` (:274), SSAA
  mean (:277-284), clip to [0, 1] (:287).

Forward images are FP32 (the per-op FP32 oracle); adjoints are float64 on the frozen
discrete decisions (hit lists, masks), as in :mod:`oracle.rasterization`.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import rasterization as orr

F32 = np.float32


def ssaa_intrinsics(intrinsics, ssaa):
    """intrinsics * (ssaa, ssaa, 1)[:, None]: rows 0 and 1 scaled (__init__.py:118-121)."""
    if ssaa <= 1:
        return intrinsics.astype(F32)
    mult = np.array([ssaa, ssaa, 1], dtype=F32)[:, None]
    return (intrinsics.astype(F32) * mult).astype(F32)


def ssaa_mean(image, ssaa):
    """Mean over ssaa x ssaa pixel blocks: [B, C, H*m, W*m] -> [B, C, H, W]."""
    if ssaa <= 1:
        return image
    b, c, h, w = image.shape
    blocks = image.reshape(b, c, h // ssaa, ssaa, w // ssaa, ssaa).astype(np.float64)
    return blocks.mean(axis=(3, 5)).astype(image.dtype)


def ssaa_spread(grad, ssaa):
    """Adjoint of :func:`ssaa_mean`: every fine pixel gets grad / ssaa^2."""
    if ssaa <= 1:
        return grad
    g = np.repeat(np.repeat(grad, ssaa, axis=2), ssaa, axis=3)
    return g / float(ssaa * ssaa)


@dataclass
class FrontEnd:
    xy: np.ndarray
    depth: np.ndarray
    sorted_keys: np.ndarray
    sorted_values: np.ndarray
    ranges: np.ndarray  # int64[B, S', T, 2]
    layout: dict
    grid_shape: tuple


def front_end(vertices, faces, boundaries, extrinsics, intrinsics, image_shape, tile_shape,
              near_plane, explicit):
    """project -> bound -> keys -> sort -> delineate (__init__.py:61-85 / :221-250)."""
    h, w = image_shape
    th, tw = tile_shape
    rows, cols = -(-h // th), -(-w // tw)
    b = extrinsics.shape[0]
    s = boundaries.shape[0]
    xy, depth = orr.project_vertices(vertices, extrinsics, intrinsics)
    tmin, tmax, cnt = orr.bound_triangles(xy, depth, faces, tile_shape, (rows, cols), near_plane)
    layout = orr.explicit_layout() if explicit else orr.fused_layout()
    keys, values, _ = orr.generate_keys(depth, faces, boundaries, tmin, tmax, cnt, cols, layout)
    sk, sv = orr.sort_keys(keys, values)
    ranges = orr.delineate_tiles(sk, layout, b, s, rows * cols, explicit)
    return FrontEnd(xy, depth, sk, sv, ranges, layout, (rows, cols))


def rasterize(vertices, faces, boundaries, extrinsics, intrinsics, image_shape,
              tile_shape=(16, 16), near_plane=0.2):
    """(uv f32[B, S, H, W, 2], index i32[B, S, H, W]) (__init__.py:22-94)."""
    fe = front_end(vertices, faces, boundaries, extrinsics, intrinsics, image_shape, tile_shape,
                   near_plane, explicit=True)
    return orr.explicit_intersect(fe.xy, faces, fe.sorted_values, fe.ranges, tile_shape,
                                  image_shape)


@dataclass
class RenderState:
    """What a render's adjoint needs from its forward."""

    mode: str
    ssaa: int
    image_shape: tuple  # the rendered (super-sampled) shape
    pre_clip: np.ndarray  # image after the SSAA mean, before the clip (fused)
    colors: np.ndarray  # f32[B, V, 3]
    raw: object  # fused: [B, 4, H, W]; explicit: (color [P, 3], trans [P])
    hits: object = None  # fused HitList
    uv: np.ndarray = None
    index: np.ndarray = None
    planes: tuple = None  # explicit: (sdf_pl [P, S], rgb_pl [P, S, 3], mask_pl [P, S])


def render(extrinsics, intrinsics, image_shape, vertices, faces, boundaries, sdf, sh, sh_degree,
           sharpness, background, tile_shape=(16, 16), near_plane=0.2, ssaa=1, mode="fused"):
    """FP32 image [B, 3, H, W] and the state its adjoint needs."""
    h, w = image_shape
    b = extrinsics.shape[0]
    bg = np.asarray(background, dtype=F32)
    big = (h * ssaa, w * ssaa) if ssaa > 1 else (h, w)
    K = ssaa_intrinsics(intrinsics, ssaa)
    s = boundaries.shape[0]
    colors = orr.evaluate_spherical_harmonics(vertices, sh, extrinsics, sh_degree)[0].astype(F32)
    if mode == "fused":
        if faces.shape[0] == 0 or vertices.shape[0] == 0:
            img = np.broadcast_to(bg[None, :, None, None], (b, 3, h, w)).astype(F32)
            return img, RenderState(mode, ssaa, big, img.copy(), colors, None)
        fe = front_end(vertices, faces, boundaries, extrinsics, K, big, tile_shape, near_plane,
                       explicit=False)
        hits = orr.fused_hits(fe.xy, faces, fe.sorted_keys, fe.sorted_values, fe.ranges[:, 0],
                              fe.layout, s, tile_shape, big)
        raw = orr.fused_forward_image(hits, colors, sdf, faces, sharpness, b, big)
        img = (raw[:, :3] + raw[:, 3:4] * bg[None, :, None, None]).astype(F32)
        img = ssaa_mean(img, ssaa)
        return np.clip(img, 0, 1), RenderState(mode, ssaa, big, img, colors, raw, hits=hits)
    uv, index = rasterize(vertices, faces, boundaries, extrinsics, K, big, tile_shape, near_plane)
    s_img, c_img = orr.interpolate(uv, index, faces, sdf, colors)
    hh, ww = big
    sdf_pl = s_img.transpose(0, 2, 3, 1).reshape(-1, s)
    rgb_pl = c_img.transpose(0, 2, 3, 1, 4).reshape(-1, s, 3)
    mask_pl = (index != -1).transpose(0, 2, 3, 1).reshape(-1, s)
    col, trans = orr.composite(sdf_pl, rgb_pl, mask_pl, sharpness)
    flat = (col + bg[None] * trans[:, None]).astype(F32)
    img = flat.reshape(b, hh, ww, 3).transpose(0, 3, 1, 2)
    img = ssaa_mean(np.ascontiguousarray(img), ssaa)
    return img, RenderState(mode, ssaa, big, img, colors, (col, trans), uv=uv, index=index,
                            planes=(sdf_pl, rgb_pl, mask_pl))


def render_backward(state: RenderState, vertices, faces, sdf, sh, extrinsics, sh_degree,
                    sharpness, background, grad_image, reference_quirk=False):
    """Adjoint of :func:`render` w.r.t. the triangle vertices' SDF and SH, float64.

    Returns (sdf_grad [V], sh_grad [NSH, V, 3])."""
    bg = np.asarray(background, dtype=np.float64)
    g = grad_image.astype(np.float64)
    if state.mode == "fused":
        # clamp's adjoint passes where min <= x <= max (torch.clamp).
        x = state.pre_clip
        g = g * ((x >= 0) & (x <= 1))
    g = ssaa_spread(g, state.ssaa)
    b = g.shape[0]
    hh, ww = state.image_shape
    v = vertices.shape[0]
    sdf64 = sdf.astype(np.float64)
    if state.raw is None:
        return np.zeros(v), np.zeros(sh.shape)
    if state.mode == "fused":
        grad4 = np.empty((b, 4, hh, ww))
        grad4[:, :3] = g
        grad4[:, 3] = np.einsum("bchw,c->bhw", g, bg)
        cg, sg, _ = orr.fused_backward(state.hits, state.colors.astype(np.float64), sdf64, faces,
                                       sharpness, state.raw, grad4, reference_quirk=reference_quirk)
    else:
        s = state.index.shape[1]
        gpix = g.transpose(0, 2, 3, 1).reshape(-1, 3)
        gt = gpix @ bg
        col, trans = state.raw
        sdf_pl, rgb_pl, mask_pl = state.planes
        gs_pl, gc_pl, _ = orr.composite_backward(sdf_pl, rgb_pl, mask_pl, sharpness, col, trans,
                                                 gpix, gt, reference_quirk=reference_quirk)
        gs_img = gs_pl.reshape(b, hh, ww, s).transpose(0, 3, 1, 2)
        gc_img = gc_pl.reshape(b, hh, ww, s, 3).transpose(0, 3, 1, 2, 4)
        sg, cg = orr.interpolate_backward(state.uv, state.index, faces, v, b, gs_img, gc_img)
    shg = orr.evaluate_spherical_harmonics_backward(vertices, sh, extrinsics, sh_degree, cg)
    return sg, shg
