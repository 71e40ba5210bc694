"""Render serving with device-resident scenes and continuous batching.

* ``ResidentScene``: a scene's extracted shells (vertices, faces, shell boundaries, SDF,
  and either the triangle vertices' SH or the grid SH as rows, ``GridColours``) kept on
  the device; tessellation happens once, at load.
* ``render_batch``: forward-only fused render of B cameras of one scene, each at its own
  resolution (a per-camera table of image and tile-grid sizes drives the key,
  delineation and compositing kernels), launched on the current stream with no host
  synchronisation: the key buffers are sized from the scene's calibrated keys-per-pixel
  density (with margin) and a device overflow flag reports the rare batch that needs a
  re-render with more room.
* ``RenderServer``: requests (scene, pose, resolution) wait in one FIFO per scene. Each of
  ``num_streams`` streams owns a device arena of budget / num_streams bytes, allocated
  once; every scheduling round, while a stream is free, the server takes the oldest
  request and gathers the next queued requests of the same scene (any resolution) into a
  batch of up to ``max_batch`` cameras whose buffers fit the stream's arena (the exact
  byte plan of ``render_batch``), and launches the batch on that stream, so batches of
  different scenes overlap on the device and no batch allocates device memory. Completed
  batches (event query, no blocking) free their stream and arena; requests that arrive
  meanwhile join the next batches. Requests that waited longer than ``deadline_s``
  without being admitted are shed (reported, never rendered).

Each camera's pixels depend only on that camera's keys (the sort is stable and keys carry
the camera in their top bits), so a request's image is the same alone and in any batch.
"""

from __future__ import annotations

import heapq
import itertools
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Mapping

import torch

from . import _native
from .tesseract import (Contraction, TesseractLayout, _initializer_offset, extract_tesseract_voxels,
                        level_sets_from_shells, raw_from_sdf, sdf_from_raw, split_occupancy)
from .triangle_rasterization.keys import create_key_defines, to_layout

_LAYOUT = to_layout(create_key_defines(("camera", 6), ("tile", 20), ("depth", 32), ("shell", 6)))


@dataclass
class GridColours:
    """What a resident scene keeps to colour its triangle vertices from the grid for each
    batch (``mt_create_triangle_vertices_colors_forward``, colours only) instead of
    keeping their SH coefficients (192 bytes per vertex at degree 3). The colours are the
    same bits as evaluating the per-vertex SH."""

    grid_vertices: torch.Tensor  # f32[V_grid, 3]
    grid_sdf: torch.Tensor  # f32[V_grid]
    sh_rows: torch.Tensor  # f32[V_grid, mt_sh_row_floats] (mt_pack_sh_rows)
    num_sh: int
    lower_corners: torch.Tensor  # int32[4, columns]
    level_sets: torch.Tensor  # f32[S]
    vertex_offsets: torch.Tensor  # int32[S * columns + 1]


@dataclass
class ResidentScene:
    vertices: torch.Tensor  # f32[V, 3]
    faces: torch.Tensor  # int32[F, 3]
    face_boundaries: torch.Tensor  # int32[S]
    signed_distances: torch.Tensor  # f32[V]
    spherical_harmonics: torch.Tensor | None  # f32[NSH, V, 3], or None with ``grid``
    sh_degree: int
    sharpness: torch.Tensor  # f32[1]
    background: torch.Tensor  # f32[3]
    keys_per_pixel: float = 8.0  # measured density of (face, tile) keys per image pixel
    # (pixels, keys) of calibrated resolutions: the most keys one camera needed at that size
    key_table: list = field(default_factory=list)
    grid: GridColours | None = None

    def resident_bytes(self) -> int:
        ts = [self.vertices, self.faces, self.face_boundaries, self.signed_distances,
              self.spherical_harmonics]
        if self.grid is not None:
            g = self.grid
            ts += [g.grid_vertices, g.grid_sdf, g.sh_rows, g.lower_corners, g.level_sets,
                   g.vertex_offsets]
        return sum(t.numel() * t.element_size() for t in ts if t is not None)

    @property
    def num_shells(self) -> int:
        return int(self.face_boundaries.shape[0])

    def keys_for(self, h: int, w: int, margin: float) -> int:
        """Key capacity for one camera. Keys grow with resolution but not in proportion
        (at low resolution nearly every visible face still needs one key), so a calibrated
        resolution at least as large bounds smaller ones; above the largest, keys are
        scaled by pixels. Without calibration: the keys-per-pixel density."""
        px = h * w
        est = None
        for cal_px, cal_keys in sorted(self.key_table):
            if cal_px >= px:
                est = cal_keys
                break
        if est is None:
            if self.key_table:
                cal_px, cal_keys = max(self.key_table)
                est = cal_keys * px / cal_px
            else:
                est = self.keys_per_pixel * px
        return int(est * margin) + 4096


def tessellate_scene(voxels, grid_vertices, grid_sdf, grid_sh, level_sets, sh_degree, sharpness,
                     background, colours: str = "vertex") -> ResidentScene:
    """Tessellate once and keep the shells resident. ``colours="vertex"`` keeps every
    triangle vertex's SH (the plain render's inputs); ``colours="grid"`` keeps the grid SH
    as rows and recomputes the batch cameras' colours from them per batch (far less
    memory for large scenes; the same images)."""
    from . import triangle_extraction as te

    if colours not in ("vertex", "grid"):
        raise ValueError(f"colours must be 'vertex' or 'grid', not {colours!r}")
    dev = grid_sdf.device
    with torch.no_grad():
        sh_in = grid_sh if colours == "vertex" else grid_sh[:1]  # positions need no SH
        tri = te.tessellate_voxels(grid_sdf, grid_vertices, sh_in, voxels.neighbors,
                                   voxels.lower_corners, voxels.upper_corners, voxels.indices,
                                   level_sets)
        grid = None
        if colours == "grid":
            offsets = te.count_triangle_vertices(grid_sdf, voxels.lower_corners, level_sets)
            te.compute_exclusive_cumsum(offsets)
            n_grid = grid_vertices.shape[0]
            rows = torch.empty((n_grid, _native.query("mt_sh_row_floats", sh_degree)),
                               dtype=torch.float32, device=dev)
            num_sh = grid_sh.shape[0]
            _native.call("mt_pack_sh_rows", grid_sh.contiguous().data_ptr(), n_grid, num_sh,
                         sh_degree, rows.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
            grid = GridColours(grid_vertices.contiguous(), grid_sdf.detach().contiguous(), rows,
                               num_sh, voxels.lower_corners.contiguous(),
                               level_sets.contiguous(), offsets)
    return ResidentScene(
        tri.vertices.detach(), tri.faces, tri.face_boundaries, tri.signed_distances.detach(),
        tri.spherical_harmonics.detach() if colours == "vertex" else None, sh_degree,
        torch.tensor([float(sharpness)], device=dev, dtype=torch.float32),
        torch.as_tensor(background, dtype=torch.float32, device=dev).contiguous(),
        grid=grid,
    )


# ------------------------------------------------------------------------------------
# Scenes on disk, in the reference trainer's checkpoint layout (trainer.py:111-137): a dict
# whose "model" entry is the model's state dict. Its tesseract entries hold the flat packed
# occupancy of all components, the raw per-vertex SDF parameters (one row per component
# vertex) and the SH split into the DC band and the rest (tesseract.py:161-214); "step" and
# "tesseract.step" hold the training step. The reference keeps the grid layout, initializer,
# shells, sharpness and contraction in its config: for its files they are passed in. A file
# written here carries them in an extra "scene" entry as well.
# ------------------------------------------------------------------------------------


def _layout_dict(layout: TesseractLayout) -> dict:
    def plain(x):
        return {int(k): int(v) for k, v in x.items()} if isinstance(x, Mapping) else int(x)

    return {"center_resolution": plain(layout.center_resolution),
            "arm_resolution": plain(layout.arm_resolution),
            "center_extent": float(layout.center_extent),
            "disable_tesseract": bool(layout.disable_tesseract)}


def save_scene_checkpoint(path, occupancy, signed_distances, spherical_harmonics, level_sets,
                          sh_degree, sharpness, background, *, initializer, step: int = 0,
                          layout: TesseractLayout | None = None, contraction: Contraction | None = None
                          ) -> None:
    """Writes a scene in the layout above. ``occupancy`` is the flat packed occupancy of the
    layout's components (or one cubic grid int32[n, n, n / 32], ``layout`` then defaulting
    to that single grid); ``signed_distances`` f32[V] is the SDF to render at the
    components' vertices and ``spherical_harmonics`` f32[NSH, V, 3]. ``initializer`` (as in
    the reference's config: {"name": "sphere", "radius": r}, {"name": "point_cloud", ...} or
    {"name": "none"}) says what a reader adds to the stored parameters, so they are stored
    as the SDF minus that offset at ``step`` (to float32 rounding); the vertices it needs
    are extracted on the occupancy's device."""
    if layout is None:
        n = occupancy.shape[0]
        if occupancy.dim() != 3 or tuple(occupancy.shape) != (n, n, n // 32):
            raise ValueError("without a layout the occupancy must be one cubic grid int32[n, n, n/32]")
        layout = TesseractLayout.single_grid(n)
    sdf = signed_distances.detach().to(torch.float32)
    # Refuses unknown initializers and offsets that need data outside the file.
    if _initializer_offset(initializer, torch.zeros((1, 3)), step) is not None:
        vox, _ = extract_tesseract_voxels(occupancy, layout, step)
        sdf = raw_from_sdf(sdf.to(vox.vertices.device), vox.vertices, initializer, step)
    sh = spherical_harmonics.detach().cpu().to(torch.float32)
    model = {
        "step": torch.tensor([step], dtype=torch.int32),
        "background": torch.as_tensor(background, dtype=torch.float32).cpu(),
        "tesseract.step": torch.tensor(step, dtype=torch.int32),
        "tesseract.occupancy": occupancy.detach().cpu().reshape(-1).to(torch.int32),
        "tesseract.signed_distances_raw": sdf.cpu(),
        "tesseract.sh_rgb": sh[:1].clone(),
        "tesseract.sh_rest": sh[1:].clone() if sh.shape[0] > 1 else torch.empty([], dtype=torch.float32),
        "tessellator.step": torch.tensor([step], dtype=torch.int32),
    }
    scene = {"layout": _layout_dict(layout), "initializer": dict(initializer),
             "level_sets": [float(x) for x in torch.as_tensor(level_sets).cpu().tolist()],
             "sh_degree": int(sh_degree), "sharpness": float(sharpness)}
    if contraction is not None:
        scene["contraction"] = {"center_radius": contraction.center_radius,
                                "arm_ratio": contraction.arm_ratio,
                                "center_extent": contraction.center_extent,
                                "disable_exponential": contraction.disable_exponential}
    torch.save({"model": model, "step": step, "scene": scene}, path)


def read_scene_checkpoint(path, layout: TesseractLayout | None = None, initializer=None,
                          level_sets=None, shells=None, sh_degree=None, sharpness=None,
                          contraction: Contraction | None = None) -> dict:
    """The scene of a checkpoint on the CPU (loaded with ``weights_only=True``: nothing in
    the file is executed). Arguments override or supply what the file lacks; a reference
    checkpoint needs the layout, the initializer, the sharpness and the level sets (or the
    shells as CDF values, converted with the sharpness). The SDF stays the stored raw
    parameters until the vertices are known (``load_scene_checkpoint``)."""
    content = torch.load(path, map_location="cpu", weights_only=True)
    model = content["model"]
    extra = content.get("scene", {})
    if layout is None and "layout" in extra:
        layout = TesseractLayout(**extra["layout"])
    if layout is None:
        raise ValueError("the tesseract layout (centre and arm resolutions, centre extent) is "
                         "needed: reference checkpoints keep it in their config")
    if initializer is None:
        initializer = extra.get("initializer")
    if initializer is None:
        raise ValueError("the initializer is needed (the rendered SDF is the stored parameters "
                         "plus its offset): reference checkpoints keep it in their config")
    step_t = model.get("tesseract.step", model.get("step", content.get("step", 0)))
    step = int(torch.as_tensor(step_t).reshape(-1)[0])
    occupancy = model["tesseract.occupancy"].to(torch.int32).reshape(-1)
    split_occupancy(occupancy, layout, step)  # checks the word count
    raw = model["tesseract.signed_distances_raw"].to(torch.float32)
    sh = model["tesseract.sh_rgb"].to(torch.float32)
    rest = model.get("tesseract.sh_rest")
    if rest is not None and rest.dim() == 3 and rest.shape[0] > 0:
        sh = torch.cat([sh, rest.to(torch.float32)], dim=0)
    if sh_degree is None:
        sh_degree = extra.get("sh_degree", int(round(sh.shape[0] ** 0.5)) - 1)
    if sharpness is None:
        sharpness = extra.get("sharpness")
        if sharpness is None:
            raise ValueError("the sharpness is needed: reference checkpoints keep it in their config")
    if level_sets is None and shells is not None:
        level_sets = level_sets_from_shells(shells, float(sharpness))
    if level_sets is None:
        level_sets = extra.get("level_sets")
        if level_sets is None:
            raise ValueError("the level sets (or shells) are needed: reference checkpoints keep "
                             "them in their config")
    if contraction is None and "contraction" in extra:
        contraction = Contraction(**extra["contraction"])
    _initializer_offset(initializer, torch.zeros((1, 3)), step)  # refuses what cannot be applied
    return {"occupancy": occupancy, "layout": layout, "initializer": dict(initializer),
            "step": step, "signed_distances_raw": raw, "spherical_harmonics": sh,
            "level_sets": torch.as_tensor(level_sets, dtype=torch.float32),
            "sh_degree": int(sh_degree), "sharpness": float(sharpness),
            "background": model.get("background", torch.zeros(3)).to(torch.float32),
            "contraction": contraction}


def load_scene_checkpoint(path, device, colours: str = "vertex", **overrides) -> ResidentScene:
    """A resident scene from a checkpoint: the tesseract's voxels extracted, mapped and
    welded on ``device``, the SDF the reference renders (stored parameters plus the
    initializer's offset at the checkpoint's step), shells tessellated at the level sets,
    and, with a contraction, the triangle vertices mapped to world space as the reference's
    tessellator does (tessellator.py:57)."""
    c = read_scene_checkpoint(path, **overrides)
    vox, _ = extract_tesseract_voxels(c["occupancy"].to(device), c["layout"], c["step"])
    if vox.vertices.shape[0] != c["signed_distances_raw"].shape[0]:
        raise ValueError(f"{c['signed_distances_raw'].shape[0]} SDF values for "
                         f"{vox.vertices.shape[0]} component vertices")
    sdf = sdf_from_raw(c["signed_distances_raw"].to(device), vox.vertices, c["initializer"],
                       c["step"]).contiguous()
    if c["contraction"] is not None and colours == "grid":
        raise ValueError("a contraction needs colours='vertex' (grid colours are computed at "
                         "the uncontracted grid vertices)")
    scene = tessellate_scene(vox, vox.vertices, sdf, c["spherical_harmonics"].to(device),
                             c["level_sets"].to(device), c["sh_degree"], c["sharpness"],
                             c["background"].to(device), colours=colours)
    if c["contraction"] is not None:
        scene.vertices = c["contraction"].uncontract(scene.vertices).contiguous()
    return scene


@dataclass
class BatchResult:
    images: list  # per camera f32[3, H_b, W_b] (background composited, clipped)
    overflow: torch.Tensor  # int32[1]
    keys: torch.Tensor  # int64[1]: keys the batch needed
    nbytes: int


class ArenaFull(RuntimeError):
    pass


class _Buffers:
    """Hands out a batch's buffers: from torch's allocator, or carved (256-byte aligned)
    out of a preallocated byte arena, or (``arena`` = "count") only counts the bytes."""

    def __init__(self, device, arena=None):
        self.device = device
        self.arena = arena
        self.offset = 0

    def take(self, shape, dtype):
        shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        numel = 1
        for x in shape:
            numel *= int(x)
        nbytes = numel * torch.empty((), dtype=dtype).element_size()
        start = (self.offset + 255) // 256 * 256
        self.offset = start + nbytes
        if self.arena is None:
            return torch.empty(shape, dtype=dtype, device=self.device)
        if isinstance(self.arena, str):
            return None
        if self.offset > self.arena.numel():
            raise ArenaFull(f"batch needs {self.offset} bytes, arena holds {self.arena.numel()}")
        return self.arena[start : start + nbytes].view(dtype).view(shape)


def _batch_plan(scene: ResidentScene, shapes, tile_shape, k_cap):
    th, tw = tile_shape
    dims = [(h, w, -(-h // th), -(-w // tw)) for h, w in shapes]
    if max(max(d[2], d[3]) for d in dims) > 65535:
        raise ValueError("tile grids are limited to 65535 rows and columns")
    tile_base, pixel_base = [0], [0]
    for h, w, r, c in dims:
        tile_base.append(tile_base[-1] + r * c)
        pixel_base.append(pixel_base[-1] + h * w)
    b = len(shapes)
    v = scene.vertices.shape[0]
    f32, i32, i64, u8 = torch.float32, torch.int32, torch.int64, torch.uint8
    plan = [
        ("xy", (b, v, 2), f32), ("depth", (b, v), f32), ("colors", (b, v, 3), f32),
        ("n_keys", 1, i64),
        ("keys", k_cap, i64), ("vals", k_cap, i32), ("skeys", k_cap, i64), ("svals", k_cap, i32),
        ("ranges", (tile_base[-1], 2), i64), ("out", 4 * pixel_base[-1], f32),
        ("campos", (b, 3), f32), ("overflow", 1, i32), ("live_x3", 1, i64),
        ("keys_ws", _native.query("mt_generate_keys_single_pass_workspace_bytes",
                                  scene.faces.shape[0], b), u8),
        ("sort_ws", _native.query("mt_sort_workspace_bytes", k_cap), u8),
        ("range_ws", _native.query("mt_sort_key_ranges_workspace_bytes", tile_base[-1]), u8),
    ]
    return dims, tile_base, pixel_base, plan


def batch_bytes(scene: ResidentScene, image_shapes, tile_shape=(16, 16), key_margin=1.25,
                key_capacity=None) -> int:
    """Device bytes ``render_batch`` takes for these cameras (what an arena must hold)."""
    shapes = [tuple(x) for x in image_shapes]
    k_cap = key_capacity or sum(scene.keys_for(h, w, key_margin) for h, w in shapes)
    _, _, _, plan = _batch_plan(scene, shapes, tile_shape, k_cap)
    bufs = _Buffers(None, "count")
    for _, shape, dtype in plan:
        bufs.take(shape, dtype)
    return bufs.offset


def render_batch(scene: ResidentScene, extrinsics: torch.Tensor, intrinsics: torch.Tensor,
                 image_shapes, tile_shape=(16, 16), chunk=0, near_plane=0.2,
                 key_margin=1.25, key_capacity=None, arena=None) -> BatchResult:
    """Forward-only fused render of B cameras of one scene, no host sync.

    ``image_shapes`` is one (H, W) for every camera or a per-camera list. With ``arena``
    (a uint8 device tensor), every working buffer is carved out of it (the images
    are fresh tensors); ``ArenaFull`` if it is too small."""
    from .batched import camera_shapes

    call = _native.call
    dev = scene.vertices.device
    st = torch.cuda.current_stream(dev).cuda_stream
    b = extrinsics.shape[0]
    if not 0 < b <= 64:
        raise ValueError(f"render_batch: {b} cameras (1..64 fit the key's camera field)")
    shapes = camera_shapes(image_shapes, b)
    table = len(set(shapes)) > 1
    th, tw = tile_shape
    s = scene.num_shells
    v = scene.vertices.shape[0]
    f = scene.faces.shape[0]
    k_cap = key_capacity or sum(scene.keys_for(h, w, key_margin) for h, w in shapes)
    dims, tile_base, pixel_base, plan = _batch_plan(scene, shapes, tile_shape, k_cap)
    rows, cols = dims[0][2], dims[0][3]
    bufs = _Buffers(dev, arena)
    t = {name: bufs.take(shape, dtype) for name, shape, dtype in plan}
    E = extrinsics.contiguous().float()
    K = intrinsics.contiguous().float()
    # Small per-batch tables travel with the launch (host -> device copies on the stream).
    cam_dims = torch.tensor([list(d) for d in dims], dtype=torch.int32).to(dev, non_blocking=True)
    tbase = torch.tensor(tile_base, dtype=torch.int64).to(dev, non_blocking=True)
    pbase = torch.tensor(pixel_base, dtype=torch.int64).to(dev, non_blocking=True)
    t["live_x3"].fill_(3 * f)
    t["overflow"].zero_()
    t["ranges"].zero_()
    xy, depth, colors, out = t["xy"], t["depth"], t["colors"], t["out"]
    keys, vals, skeys, svals = t["keys"], t["vals"], t["skeys"], t["svals"]
    keys_ws, sort_ws, range_ws, n_dev = t["keys_ws"], t["sort_ws"], t["range_ws"], t["n_keys"]
    live_x3, overflow, ranges, campos = t["live_x3"], t["overflow"], t["ranges"], t["campos"]
    dims_p = cam_dims.data_ptr() if table else None
    call("mt_project_vertices", scene.vertices.data_ptr(), v, E.data_ptr(), K.data_ptr(), b,
         xy.data_ptr(), depth.data_ptr(), st)
    call("mt_generate_keys_single_pass", xy.data_ptr(), depth.data_ptr(), v, scene.faces.data_ptr(),
         f, live_x3.data_ptr(), scene.face_boundaries.data_ptr(), s, b, th, tw, 0, 0,
         max(d[2] for d in dims), max(d[3] for d in dims),
         float(near_plane), dims_p, 1, cols, _LAYOUT, keys.data_ptr(), vals.data_ptr(), k_cap,
         n_dev.data_ptr(), overflow.data_ptr(), skeys.data_ptr(), svals.data_ptr(),
         keys_ws.data_ptr(), keys_ws.numel(), st)
    # Tile-bucketed sort (as in the training step): (camera, tile) bits, tile ranges, then
    # each tile's keys by (depth, shell), with the unsorted key arrays as temporary space.
    low = _LAYOUT.tile_offset
    call("mt_sort_keys_high", keys.data_ptr(), vals.data_ptr(), k_cap, skeys.data_ptr(),
         svals.data_ptr(), sort_ws.data_ptr(), sort_ws.numel(), n_dev.data_ptr(), low, st)
    call("mt_delineate_tiles", skeys.data_ptr(), k_cap, _LAYOUT, 0, b, 1, rows * cols,
         ranges.data_ptr(), n_dev.data_ptr(), tbase.data_ptr() if table else None, st)
    call("mt_sort_key_ranges", skeys.data_ptr(), svals.data_ptr(), k_cap, ranges.data_ptr(),
         tile_base[-1], low, keys.data_ptr(), vals.data_ptr(), range_ws.data_ptr(),
         range_ws.numel(), st)
    call("mt_camera_positions", E.data_ptr(), b, campos.data_ptr(), st)
    if scene.grid is None:
        call("mt_evaluate_sh_forward", scene.vertices.data_ptr(), v,
             scene.spherical_harmonics.data_ptr(), scene.spherical_harmonics.shape[0],
             scene.sh_degree, campos.data_ptr(), b, colors.data_ptr(), st)
    else:  # colours only (parts 2), from the grid's SH rows (+4)
        g = scene.grid
        call("mt_create_triangle_vertices_colors_forward", g.grid_vertices.data_ptr(),
             g.grid_sdf.data_ptr(), g.sh_rows.data_ptr(), g.grid_vertices.shape[0], g.num_sh,
             scene.sh_degree, g.lower_corners.data_ptr(), g.lower_corners.shape[1],
             g.level_sets.data_ptr(), s, g.vertex_offsets.data_ptr(), v, campos.data_ptr(), b,
             None, None, None, colors.data_ptr(), overflow.data_ptr(), 2 | 4, 0, None, st)
    h0, w0 = shapes[0]
    args = _native.CompositeArgs(
        xy.data_ptr(), colors.data_ptr(), scene.signed_distances.data_ptr(), v,
        scene.faces.data_ptr(), scene.sharpness.data_ptr(), skeys.data_ptr(), svals.data_ptr(),
        ranges.data_ptr(), b, s, rows, cols, th, tw, h0, w0, _LAYOUT, chunk, None,
        dims_p, tbase.data_ptr() if table else None, pbase.data_ptr() if table else None,
        tile_base[-1] if table else 0,
    )
    call("mt_intersect_and_composite_forward", args, out.data_ptr(), st)
    # Background and clip (the plain render's expression); the images are the batch's only
    # allocations outside the arena.
    bg = scene.background[:, None, None]
    images = []
    for (h, w), p in zip(shapes, pixel_base):
        o = out[4 * p : 4 * (p + h * w)].view(4, h, w)
        images.append((o[:3] + o[3:4] * bg).clip(0, 1))
    return BatchResult(images, overflow, n_dev, bufs.offset)


def _max_keys(scene, extrinsics, intrinsics, image_shape, tile_shape, near_plane) -> int:
    from .batched import count_keys

    best = 0
    with torch.no_grad():
        for c in range(extrinsics.shape[0]):
            best = max(best, count_keys(scene.vertices, scene.faces, extrinsics[c : c + 1],
                                        intrinsics[c : c + 1], image_shape, tile_shape, near_plane))
    return best


def calibrate_key_density(scene: ResidentScene, extrinsics, intrinsics, image_shape,
                          tile_shape=(16, 16), near_plane=0.2) -> float:
    """The largest keys-per-pixel density over a sample of poses (reads sizes back; done
    at scene load)."""
    h, w = image_shape
    return _max_keys(scene, extrinsics, intrinsics, image_shape, tile_shape, near_plane) / (h * w)


def calibrate_keys(scene: ResidentScene, poses_by_shape: dict, tile_shape=(16, 16),
                   near_plane=0.2) -> list:
    """(pixels, most keys of one camera) per resolution over sample poses
    ``{(H, W): (extrinsics[N, 4, 4], intrinsics[N, 3, 3])}``; sets ``scene.key_table``."""
    table = []
    for (h, w), (E, K) in poses_by_shape.items():
        table.append((h * w, _max_keys(scene, E, K, (h, w), tile_shape, near_plane)))
    scene.key_table = sorted(table)
    return scene.key_table


# ------------------------------------------------------------------------------------


@dataclass(order=True)
class Request:
    arrival: float
    rid: int
    scene_id: int = field(compare=False)
    extrinsics: torch.Tensor = field(compare=False)  # f32[4, 4]
    intrinsics: torch.Tensor = field(compare=False)  # f32[3, 3]
    image_shape: tuple = field(compare=False)
    done: float | None = field(default=None, compare=False)
    image: torch.Tensor | None = field(default=None, compare=False)


@dataclass
class _InFlight:
    requests: list
    result: BatchResult
    event: torch.cuda.Event
    stream: torch.cuda.Stream


class RenderServer:
    def __init__(self, scenes: dict, hbm_budget_bytes: int, max_batch: int = 8, num_streams: int = 4,
                 tile_shape=(16, 16), chunk=0, key_margin=1.25, keep_images=False,
                 deadline_s: float | None = None):
        self.scenes = scenes
        self.budget = int(hbm_budget_bytes)
        self.max_batch = min(int(max_batch), 64)
        self.tile_shape = tile_shape
        self.chunk = chunk
        self.key_margin = key_margin
        self.keep_images = keep_images
        self.deadline_s = deadline_s
        dev = next(iter(scenes.values())).vertices.device
        self.device = dev
        # Each stream owns an arena of budget / num_streams bytes, allocated once: a batch's
        # buffers are carved out of its stream's arena, so serving never allocates device
        # memory per batch (only the small output images).
        self.arena_bytes = self.budget // num_streams // 256 * 256
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(num_streams)]
        self.arenas = {st: torch.empty(self.arena_bytes, dtype=torch.uint8, device=dev)
                       for st in self.streams}
        self.free_streams = list(self.streams)
        # One FIFO per scene; submission order across scenes by sequence number.
        self.queues: dict = {sid: deque() for sid in scenes}
        self._seq = itertools.count()
        self.inflight: list[_InFlight] = []
        self.completed: list[Request] = []
        self.shed: list[Request] = []
        self.rerenders = 0
        self.batches = 0
        self.cameras = 0
        self.oversize = 0

    def reset_stats(self) -> None:
        """Clears completed / shed lists and counters (the queues must be empty)."""
        if self.queued or self.inflight:
            raise RuntimeError("reset_stats: requests still queued or in flight")
        self.completed, self.shed = [], []
        self.rerenders = self.batches = self.cameras = self.oversize = 0

    @property
    def queued(self) -> int:
        return sum(len(q) for q in self.queues.values())

    def submit(self, req: Request) -> None:
        self.queues[req.scene_id].append((next(self._seq), req))

    def _bytes(self, sid, shapes, key_capacity=None) -> int:
        return batch_bytes(self.scenes[sid], shapes, self.tile_shape, self.key_margin, key_capacity)

    def _shed_late(self) -> None:
        if self.deadline_s is None:
            return
        cutoff = time.perf_counter() - self.deadline_s
        for q in self.queues.values():
            while q and q[0][1].arrival < cutoff:
                self.shed.append(q.popleft()[1])

    def schedule(self) -> int:
        """Launch batches while a stream is free; returns batches launched. A batch = the
        oldest queued request plus the next queued requests of the same scene (any
        resolution, in order), up to max_batch cameras and what the stream's arena holds.
        A single request larger than an arena runs alone, from the general allocator."""
        self._shed_late()
        launched = 0
        while self.free_streams:
            heads = [(q[0][0], sid) for sid, q in self.queues.items() if q]
            if not heads:
                break
            _, sid = min(heads)
            q = self.queues[sid]
            shapes = [tuple(q[0][1].image_shape)]
            if self._bytes(sid, shapes) > self.arena_bytes:
                if self.inflight:
                    break  # runs alone once the device is idle (never overtaken)
                self._launch(sid, [q.popleft()[1]], use_arena=False)
                launched += 1
                continue
            batch = [q.popleft()[1]]
            while q and len(batch) < self.max_batch:
                more = shapes + [tuple(q[0][1].image_shape)]
                if self._bytes(sid, more) > self.arena_bytes:
                    break
                shapes = more
                batch.append(q.popleft()[1])
            self._launch(sid, batch)
            launched += 1
        return launched

    def _launch(self, sid, reqs, key_capacity=None, use_arena=True) -> None:
        scene = self.scenes[sid]
        stream = self.free_streams.pop()
        arena = self.arenas[stream] if use_arena else None
        if arena is not None and key_capacity is not None:
            if self._bytes(sid, [tuple(r.image_shape) for r in reqs], key_capacity) > self.arena_bytes:
                arena = None
        if arena is None:
            self.oversize += 1
        with torch.cuda.stream(stream):
            E = torch.stack([r.extrinsics for r in reqs]).to(self.device, non_blocking=True)
            K = torch.stack([r.intrinsics for r in reqs]).to(self.device, non_blocking=True)
            res = render_batch(scene, E, K, [tuple(r.image_shape) for r in reqs], self.tile_shape,
                               self.chunk, key_margin=self.key_margin, key_capacity=key_capacity,
                               arena=arena)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.batches += 1
        self.cameras += len(reqs)
        self.inflight.append(_InFlight(reqs, res, ev, stream))

    def poll(self) -> int:
        """Retire finished batches without blocking; returns requests completed."""
        now = time.perf_counter()
        still, finished = [], []
        for job in self.inflight:
            (finished if job.event.query() else still).append(job)
        self.inflight = still
        done = 0
        for job in finished:
            self.free_streams.append(job.stream)
            if int(job.result.overflow.item()):  # the batch has completed: these reads do not wait
                self.rerenders += 1
                need = int(job.result.keys.item()) + 4096
                self._launch(job.requests[0].scene_id, job.requests, key_capacity=need)
                continue
            for i, r in enumerate(job.requests):
                r.done = now
                if self.keep_images:
                    r.image = job.result.images[i]
                self.completed.append(r)
                done += 1
        return done

    def drain(self) -> None:
        while self.queued or self.inflight:
            self.schedule()
            self.poll()


def latency_stats(reqs) -> dict:
    lat = sorted((r.done - r.arrival) * 1000 for r in reqs if r.done is not None)
    if not lat:
        return {"n": 0}

    def pct(p):
        i = min(len(lat) - 1, max(0, int(round(p / 100 * (len(lat) - 1)))))
        return lat[i]

    return {"n": len(lat), "p50_ms": round(pct(50), 3), "p99_ms": round(pct(99), 3),
            "mean_ms": round(sum(lat) / len(lat), 3)}


def poisson_load(server: RenderServer, make_request, rate_per_s: float, duration_s: float,
                 seed: int = 0) -> dict:
    """Submit a Poisson stream of requests for ``duration_s`` seconds (wall clock) while
    scheduling and polling continuously, then drain; returns completed throughput, the
    latency percentiles of completed requests and the number shed past the deadline."""
    import random

    rng = random.Random(seed)
    t0 = time.perf_counter()
    arrivals = []
    t = 0.0
    while True:
        t += rng.expovariate(rate_per_s)
        if t >= duration_s:
            break
        arrivals.append(t)
    heapq.heapify(arrivals)
    submitted = []
    rid = 0
    while arrivals or server.queued or server.inflight:
        now = time.perf_counter() - t0
        while arrivals and arrivals[0] <= now:
            a = heapq.heappop(arrivals)
            req = make_request(rid, t0 + a)
            rid += 1
            server.submit(req)
            submitted.append(req)
        server.schedule()
        server.poll()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = latency_stats(submitted)
    stats.update({"offered_req_per_s": rate_per_s, "submitted": len(submitted),
                  "completed": len(server.completed), "shed": len(server.shed),
                  "req_per_s": round(len(server.completed) / wall, 2), "wall_s": round(wall, 3),
                  "batches": server.batches,
                  "mean_batch": round(server.cameras / max(server.batches, 1), 2),
                  "rerenders": server.rerenders, "oversize": server.oversize})
    return stats
