"""Batched, graph-capturable training step: tessellate all shells, render all cameras,
image loss, and the backward chain to the grid SDF and SH, as one launch sequence with
no host synchronisation.

The plain API (``triangle_extraction`` / ``triangle_rasterization``) mirrors the
reference and therefore reads three sizes back to the host every step (vertex, face and
key counts) to allocate exactly-sized tensors. Here every buffer is preallocated at a
capacity, kernels read the live counts from device memory (``num_faces_x3`` /
``n_device`` arguments of the C ABI) and write nothing past a capacity (they raise a
device overflow flag instead), so the whole step can be captured once into a hipGraph
(``torch.cuda.CUDAGraph``) and replayed.

Unlike the plain path, the per-triangle-vertex SH array ([NSH, V, 3], 192 bytes per vertex
at degree 3) is never stored: vertices and their colours come out of one kernel that
interpolates the grid SH in registers, and the SH adjoint is recomputed per triangle
vertex and distributed to the grid. Keys are generated straight from the projected
geometry (no per-face tile bounds, counts or offsets are stored), and only for tiles that
meet the face's candidate pixel rectangle (the bound the composites apply per tile): a
face that covers no pixel centre of a tile cannot change that tile, and most sub-pixel
faces cover none, so the sort and the composites see a fraction of the reference's keys.
Images are bit-identical to the plain path's; gradients equal it up to float summation
order.

Cameras may have different resolutions (a per-camera table of image and tile-grid sizes
drives the key, delineation, compositing and loss kernels) and may be rendered in groups
of ``camera_group`` cameras one after another inside the same captured step, which bounds
the per-camera buffers (projected vertices, colours and their gradients, keys) to one
group: the tessellation runs once, each group adds its share of the loss and of the grid
gradients.

Capacities come from one eager step (``measure_capacities``) times a margin; after a
replay, ``overflowed()`` (one host read) tells whether any capacity was exceeded, in
which case the caller rebuilds the step with larger capacities and runs it again.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _native
from .triangle_extraction import Voxels
from .triangle_rasterization.keys import create_key_defines, to_layout


@dataclass(frozen=True)
class Capacities:
    vertices: int
    faces: int
    keys: int  # per camera group

    def scaled(self, margin: float) -> "Capacities":
        return Capacities(int(self.vertices * margin) + 1024, int(self.faces * margin) + 1024,
                          int(self.keys * margin) + 4096)


def camera_shapes(image_shape, b: int) -> list[tuple[int, int]]:
    """(H, W) of every camera from one shared (H, W) or a per-camera list of them."""
    if len(image_shape) == 2 and all(isinstance(x, int) for x in image_shape):
        return [(int(image_shape[0]), int(image_shape[1]))] * b
    shapes = [(int(s[0]), int(s[1])) for s in image_shape]
    if len(shapes) != b:
        raise ValueError(f"{len(shapes)} image shapes for {b} cameras")
    return shapes


def camera_groups(b: int, camera_group: int | None) -> list[tuple[int, int]]:
    g = b if not camera_group else max(1, min(int(camera_group), b))
    return [(c, min(b, c + g)) for c in range(0, b, g)]


class BatchedTrainingStep:
    """One fused forward + backward step over B cameras (any resolutions) and S shells."""

    def __init__(self, voxels: Voxels, level_sets: torch.Tensor, extrinsics: torch.Tensor,
                 intrinsics: torch.Tensor, image_shape, num_sh: int, sh_degree: int,
                 sharpness: float, background, capacities: Capacities, tile_shape=(16, 16),
                 chunk: int = 0, near_plane: float = 0.2, reference_quirk: bool = False,
                 camera_group: int | None = None, cull_keys: bool = True,
                 faces_beside_colours: bool = True):
        dev = voxels.lower_corners.device
        self.device = dev
        self.voxels = voxels
        self.level_sets = level_sets.contiguous().to(dev)
        self.extrinsics = extrinsics.contiguous().float().to(dev)
        self.intrinsics = intrinsics.contiguous().float().to(dev)
        self.B = self.extrinsics.shape[0]
        self.shapes = camera_shapes(image_shape, self.B)
        self.uniform = len(set(self.shapes)) == 1
        self.H, self.W = self.shapes[0]
        self.S = self.level_sets.shape[0]
        self.num_sh = num_sh
        self.sh_degree = sh_degree
        self.tile_shape = tuple(tile_shape)
        self.chunk = chunk
        self.near_plane = float(near_plane)
        self.quirk = int(reference_quirk)
        self.cull = int(cull_keys)
        self.cap = capacities
        self.layout = to_layout(create_key_defines(("camera", 6), ("tile", 20), ("depth", 32),
                                                   ("shell", 6)))
        self.n_col = voxels.lower_corners.shape[1]
        self.n_vox = voxels.upper_corners.shape[1]
        self.n_grid = voxels.vertices.shape[0]
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        V, F, K, S = self.cap.vertices, self.cap.faces, self.cap.keys, self.S
        th, tw = self.tile_shape

        # Camera table: image and tile-grid sizes; first pixel of every camera.
        dims = [(h, w, -(-h // th), -(-w // tw)) for h, w in self.shapes]
        if max(max(d[2], d[3]) for d in dims) > 65535:
            raise ValueError("tile grids are limited to 65535 rows and columns")
        first = [0]
        for h, w in self.shapes:
            first.append(first[-1] + h * w)
        self.pixel_first = first
        self.num_pixels = first[-1]
        self.rows, self.cols = dims[0][2], dims[0][3]
        # Bounds of every camera's tile grid (the key kernel sizes its count check by them).
        self.max_rows, self.max_cols = max(d[2] for d in dims), max(d[3] for d in dims)
        self.T = self.rows * self.cols
        # Camera groups, each with its device-side table (camera indices local to the group).
        self.groups = []
        for c0, c1 in camera_groups(self.B, camera_group):
            tb, pb = [0], [0]
            for c in range(c0, c1):
                tb.append(tb[-1] + dims[c][2] * dims[c][3])
                pb.append(pb[-1] + dims[c][0] * dims[c][1])
            self.groups.append({
                "c0": c0, "c1": c1, "n": c1 - c0, "tiles": tb[-1], "pixels": pb[-1],
                "cam_dims": torch.tensor([list(d) for d in dims[c0:c1]], **i32),
                "tile_base": torch.tensor(tb, **i64),
                "pixel_base": torch.tensor(pb, **i64),
            })
        self.G = max(g["n"] for g in self.groups)
        self.total_tiles = sum(g["tiles"] for g in self.groups)
        G = self.G

        self.vertex_offsets = torch.zeros(S * self.n_col + 1, **i32)
        # Face-corner counts per (level set, cell) and their exclusive offsets in int64: a
        # mesh may hold 2^31 face corners or more (716M faces) before face ids (int32) run out.
        self.face_counts = torch.zeros(S * self.n_vox + 1, **i32)
        self.face_offsets = torch.zeros(S * self.n_vox + 1, **i64)
        self.codes = torch.empty((S, self.n_vox), dtype=torch.uint8, device=dev)
        self.tv = torch.zeros((V, 3), **f32)
        self.td = torch.zeros(V, **f32)
        self.ttype = torch.zeros(V, dtype=torch.uint8, device=dev)
        self.faces = torch.zeros((F, 3), **i32)
        # The step needs no voxel index per face. Faces find their corners' vertices from one word per (level set, column), (first
        # vertex << 3) | crossing axes, written by the vertex pass; past 2^29 vertices from the
        # offsets and the count pass's crossing-axis bytes.
        # faces_beside_colours: the faces pass (crossing-axis bytes) runs on a second stream
        # beside the first group's vertex and colour pass instead of after it.
        self.faces_early = bool(faces_beside_colours)
        self.use_words = V <= (1 << 29) and not self.faces_early
        if self.use_words:
            self.column_words = torch.zeros(S * self.n_col, dtype=torch.int32, device=dev)
        else:
            self.column_axes = torch.zeros(S * self.n_col, dtype=torch.uint8, device=dev)
        self.xy = torch.zeros((G, V, 2), **f32)
        self.depth = torch.zeros((G, V), **f32)
        self.n_keys = torch.zeros(1, **i64)  # key count of the current camera group
        self.keys = torch.zeros(K, **i64)
        self.values = torch.zeros(K, **i32)
        self.sorted_keys = torch.zeros(K, **i64)
        self.sorted_values = torch.zeros(K, **i32)
        self.ranges = torch.zeros((max(g["tiles"] for g in self.groups), 2), **i64)
        self.campos = torch.zeros((self.B, 3), **f32)
        # Vertex-major colours [V, G, 3]: a vertex's cameras are one contiguous record, which
        # the colour adjoint reads (and whose gradients it reads and clears) at once.
        self.colors = torch.zeros((V, G, 3), **f32)
        # Each grid vertex's SH coefficients as one contiguous row, repacked every step.
        row = _native.query("mt_sh_row_floats", self.sh_degree)
        self.sh_rows = torch.empty((self.n_grid, row), **f32)
        self.maps_built = False
        self.out_flat = torch.zeros(4 * self.num_pixels, **f32)
        self.grad_flat = torch.zeros_like(self.out_flat)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
        self.colors_grad = torch.zeros_like(self.colors)
        self.td_grad = torch.zeros_like(self.td)
        # Vertices that received a gradient in the current group (set by the composite
        # backward, read by the colour adjoint to skip the rest).
        self.touched = torch.zeros(V, dtype=torch.uint8, device=dev)
        self.sharp_grad = torch.zeros(1, **f32)
        self.grid_sdf_grad = torch.zeros(self.n_grid, **f32)
        self.grid_sh_grad = torch.zeros((num_sh, self.n_grid, 3), **f32)
        self.overflow = torch.zeros(1, **i32)
        self.face_boundaries = torch.zeros(S, **i32)  # last face (exclusive) of each shell
        self.face_bound64 = torch.zeros(S, **i64)
        self.side = torch.cuda.Stream(device=dev)  # default priority, like every stream here
        self.group_keys = torch.zeros(len(self.groups), **i64)  # key count of each group
        self.sharpness = torch.tensor([float(sharpness)], **f32)
        self.background = torch.as_tensor(background, dtype=torch.float32, device=dev).contiguous()
        scan_n = max(S * self.n_col + 1, S * self.n_vox)
        self.scan_ws = _native.workspace(_native.query("mt_scan_workspace_bytes", scan_n), dev)
        self.keys_ws = _native.workspace(
            _native.query("mt_generate_keys_single_pass_workspace_bytes", F, G), dev)
        self.sort_ws = _native.workspace(_native.query("mt_sort_workspace_bytes", K), dev)
        self.range_ws = _native.workspace(
            _native.query("mt_sort_key_ranges_workspace_bytes", self.ranges.shape[0]), dev)
        self.vbwd_ws = _native.workspace(
            _native.query("mt_create_triangle_vertices_colors_backward_workspace_bytes",
                          self.n_grid, self.n_col, sh_degree), dev)
        # Live counts, read by kernels from device memory (int64 face corners).
        self.num_faces_x3 = self.face_offsets[S * self.n_vox :]
        self.graph = None
        self.static_inputs = None

    # ------------------------------------------------------------------------------
    @property
    def out(self) -> torch.Tensor:
        """The rendered (rgb, transmittance) planes: [B, 4, H, W] when every camera has the
        same resolution, else the flat per-camera concatenation (camera b's [4, H_b, W_b]
        at 4 * pixel_first[b])."""
        if self.uniform:
            return self.out_flat.view(self.B, 4, self.H, self.W)
        return self.out_flat

    def image(self, b: int) -> torch.Tensor:
        """Camera b's (rgb, transmittance) planes [4, H_b, W_b]."""
        h, w = self.shapes[b]
        p = self.pixel_first[b]
        return self.out_flat[4 * p : 4 * (p + h * w)].view(4, h, w)

    def _composite_args(self, g=None):
        g = self.groups[0] if g is None else g
        th, tw = self.tile_shape
        table = not self.uniform
        h, w = self.shapes[g["c0"]]
        return _native.CompositeArgs(
            self.xy.data_ptr(), self.colors.data_ptr(), self.td.data_ptr(), self.cap.vertices,
            self.faces.data_ptr(), self.sharpness.data_ptr(), self.sorted_keys.data_ptr(),
            self.sorted_values.data_ptr(), self.ranges.data_ptr(), g["n"], self.S, self.rows,
            self.cols, th, tw, h, w, self.layout, self.chunk, None,
            g["cam_dims"].data_ptr() if table else None,
            g["tile_base"].data_ptr() if table else None,
            g["pixel_base"].data_ptr() if table else None, g["tiles"] if table else 0,
            None, self.G,
        )

    def run(self, grid_vertices: torch.Tensor, grid_sdf: torch.Tensor, grid_sh: torch.Tensor,
            target: torch.Tensor, hook=None) -> None:
        """Enqueue one forward + backward on the current stream (no host sync).

        ``target`` is [B, 3, H, W] (one resolution) or the flat per-camera concatenation of
        the [3, H_b, W_b] planes. ``hook(name)``, if given, is called after each library
        launch (the benchmark records a HIP event there to time every kernel stage)."""
        if hook is None:
            call = _native.call
        else:
            def call(name, *args):
                _native.call(name, *args)
                hook(name)
        st = torch.cuda.current_stream(self.device).cuda_stream
        v = self.voxels
        S = self.S
        V, F, K = self.cap.vertices, self.cap.faces, self.cap.keys
        th, tw = self.tile_shape
        target = target.contiguous().view(-1)
        table = not self.uniform
        # Each run reports only its own capacity overflows (a fill kernel, so it is part of a
        # captured replay too).
        self.overflow.zero_()
        self.loss.zero_()
        self.sharp_grad.zero_()
        # The SH rows need only grid_sh: without per-launch timing they are packed on a second
        # stream beside the tessellation counts and scans.
        main = torch.cuda.current_stream(self.device)
        if hook is None:
            self.side.wait_stream(main)
        with torch.cuda.stream(self.side if hook is None else main):
            call("mt_pack_sh_rows", grid_sh.data_ptr(), self.n_grid, self.num_sh, self.sh_degree,
                 self.sh_rows.data_ptr(), torch.cuda.current_stream(self.device).cuda_stream)
        # --- tessellation counts (once for all camera groups)
        if self.use_words:
            call("mt_count_triangle_vertices", grid_sdf.data_ptr(), v.lower_corners.data_ptr(),
                 self.n_col, self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(), st)
        else:
            call("mt_count_triangle_vertices_axes", grid_sdf.data_ptr(), v.lower_corners.data_ptr(),
                 self.n_col, self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(),
                 self.column_axes.data_ptr(), st)
        call("mt_exclusive_scan_i32", self.vertex_offsets.data_ptr(), self.vertex_offsets.numel(),
             self.scan_ws.data_ptr(), self.scan_ws.numel(), st)
        call("mt_count_triangle_faces", grid_sdf.data_ptr(), v.lower_corners.data_ptr(), self.n_col,
             v.upper_corners.data_ptr(), self.n_vox, self.level_sets.data_ptr(), S,
             self.face_counts.data_ptr(), self.codes.data_ptr(), st)
        # face_offsets[0] stays 0; entries 1.. are the inclusive sums of the counts.
        call("mt_inclusive_scan_i32_to_i64", self.face_counts.data_ptr(),
             self.face_offsets[1:].data_ptr(), S * self.n_vox, self.scan_ws.data_ptr(),
             self.scan_ws.numel(), st)
        call("mt_camera_positions", self.extrinsics.data_ptr(), self.B, self.campos.data_ptr(), st)
        if hook is None:
            main.wait_stream(self.side)
        # The first camera group's adjoint writes every grid gradient (zero where no vertex
        # contributes), so they need no clearing; later groups add to them over the grid
        # vertices that receive a gradient.
        for gi, g in enumerate(self.groups):
            c0, n = g["c0"], g["n"]
            cam_dims = g["cam_dims"].data_ptr() if table else None
            # colors_grad and td_grad hold zeros here: they start as zeros and the adjoint
            # clears every entry the composite backward wrote (the flagged vertices).
            self.touched.zero_()
            self.ranges[: g["tiles"]].zero_()
            if hook is not None:
                hook("zero_fill")
            main = torch.cuda.current_stream(self.device)
            overlap = gi == 0 and hook is None
            early = overlap and self.faces_early

            def faces_pass():
                with torch.cuda.stream(self.side if overlap else main):
                    fst = torch.cuda.current_stream(self.device).cuda_stream
                    if self.use_words:
                        call("mt_create_triangle_faces_words_i64", v.neighbors.data_ptr(),
                             self.n_vox, self.n_col, self.column_words.data_ptr(),
                             self.codes.data_ptr(), S, self.face_offsets.data_ptr(),
                             self.faces.data_ptr(), F, V, self.overflow.data_ptr(), fst)
                    else:
                        call("mt_create_triangle_faces_axes_i64", v.neighbors.data_ptr(),
                             self.n_vox, self.n_col, self.vertex_offsets.data_ptr(),
                             self.column_axes.data_ptr(), self.codes.data_ptr(), S,
                             self.face_offsets.data_ptr(), self.faces.data_ptr(), F, V,
                             self.overflow.data_ptr(), fst)
                    torch.floor_divide(self.face_offsets[self.n_vox :: self.n_vox][:S], 3,
                                       out=self.face_bound64)
                    self.face_boundaries.copy_(self.face_bound64)

            if early:  # the faces need only the counts: beside the vertex and colour pass
                self.side.wait_stream(main)
                faces_pass()
            # Vertices and the first group's colours in one pass, later groups' colours alone
            # (the positions are the same); the per-vertex SH array is never stored.
            call("mt_create_triangle_vertices_colors_forward", grid_vertices.data_ptr(),
                 grid_sdf.data_ptr(), self.sh_rows.data_ptr(), self.n_grid, self.num_sh,
                 self.sh_degree, v.lower_corners.data_ptr(), self.n_col,
                 self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(), V,
                 self.campos[c0:].data_ptr(), n, self.tv.data_ptr(), self.td.data_ptr(),
                 self.ttype.data_ptr(), self.colors.data_ptr(), self.overflow.data_ptr(),
                 (3 if gi == 0 else 2) | 4, self.G,
                 self.column_words.data_ptr() if gi == 0 and self.use_words else None, st)
            # Otherwise the faces (first group only) and the group's projection, which both
            # need only the vertices, run side by side on two streams (without per-launch
            # timing).
            if gi == 0 and not early:
                if overlap:
                    self.side.wait_stream(main)
                faces_pass()
            if gi == 0 and hook is not None:
                hook("face_boundaries")
            face_boundaries = self.face_boundaries
            # --- rasterization of the group's cameras
            # Live vertices only (the arrays' capacity has a margin above the vertex count).
            call("mt_project_vertices_live", self.tv.data_ptr(), V,
                 self.vertex_offsets[S * self.n_col :].data_ptr(), self.extrinsics[c0:].data_ptr(),
                 self.intrinsics[c0:].data_ptr(), n, self.xy.data_ptr(), self.depth.data_ptr(), st)
            if overlap:
                main.wait_stream(self.side)
            # Keys of the group's (camera, face, tile) triples in one pass (counts, their scan
            # by look-back, and the keys); the total lands in n_keys on the device.
            n_keys = self.n_keys
            call("mt_generate_keys_single_pass", self.xy.data_ptr(), self.depth.data_ptr(), V,
                 self.faces.data_ptr(), F, self.num_faces_x3.data_ptr(), face_boundaries.data_ptr(),
                 S, n, th, tw, 0, 0, self.max_rows, self.max_cols, self.near_plane, cam_dims, self.cull,
                 self.cols, self.layout, self.keys.data_ptr(), self.values.data_ptr(), K,
                 n_keys.data_ptr(), self.overflow.data_ptr(), self.sorted_keys.data_ptr(),
                 self.sorted_values.data_ptr(), self.keys_ws.data_ptr(), self.keys_ws.numel(), st)
            self.group_keys[gi : gi + 1].copy_(n_keys)
            # Tile-bucketed sort: radix passes over the (camera, tile) bits only, the tile
            # ranges, then each tile's keys ordered by (depth, shell) inside one workgroup
            # (the unsorted key arrays serve as its temporary space): the same order as the
            # full sort of the keys.
            low = self.layout.tile_offset
            call("mt_sort_keys_high", self.keys.data_ptr(), self.values.data_ptr(), K,
                 self.sorted_keys.data_ptr(), self.sorted_values.data_ptr(), self.sort_ws.data_ptr(),
                 self.sort_ws.numel(), n_keys.data_ptr(), low, st)
            call("mt_delineate_tiles", self.sorted_keys.data_ptr(), K, self.layout, 0, n, 1, self.T,
                 self.ranges.data_ptr(), n_keys.data_ptr(),
                 g["tile_base"].data_ptr() if table else None, st)
            call("mt_sort_key_ranges", self.sorted_keys.data_ptr(), self.sorted_values.data_ptr(), K,
                 self.ranges.data_ptr(), g["tiles"], low, self.keys.data_ptr(),
                 self.values.data_ptr(), self.range_ws.data_ptr(), self.range_ws.numel(), st)
            args = self._composite_args(g)
            p0 = self.pixel_first[c0]
            out = self.out_flat[4 * p0 :]
            grad = self.grad_flat[4 * p0 :]
            call("mt_intersect_and_composite_forward", args, out.data_ptr(), st)
            # --- loss (one mean over every camera's pixels) and backward chain
            h, w = self.shapes[c0]
            call("mt_image_l2_loss", out.data_ptr(), self.background.data_ptr(),
                 target[3 * p0 :].data_ptr(), n, h, w,
                 g["pixel_base"].data_ptr() if table else None, g["pixels"], self.num_pixels,
                 self.loss.data_ptr(), grad.data_ptr(), st)
            args.vertex_touched = self.touched.data_ptr()
            call("mt_intersect_and_composite_backward", args, out.data_ptr(), grad.data_ptr(),
                 self.colors_grad.data_ptr(), self.td_grad.data_ptr(), self.sharp_grad.data_ptr(),
                 self.quirk, st)
            # SH adjoint fused into the vertex adjoint: the first group writes the grid
            # gradients, later groups add to them.
            call("mt_create_triangle_vertices_colors_backward", grid_sdf.data_ptr(), self.n_grid,
                 self.num_sh, self.sh_degree, v.lower_corners.data_ptr(), self.n_col,
                 self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(), V,
                 self.tv.data_ptr(), self.colors.data_ptr(), self.campos[c0:].data_ptr(), n,
                 self.td_grad.data_ptr(), self.colors_grad.data_ptr(),
                 self.grid_sdf_grad.data_ptr(), self.grid_sh_grad.data_ptr(),
                 self.vbwd_ws.data_ptr(), self.vbwd_ws.numel(), int(gi > 0),
                 self.touched.data_ptr(), 1,
                 int(self.maps_built),
                 self.overflow.data_ptr(), self.G, st)
            # The adjoint's grid-vertex maps stay in its workspace (they depend only on the
            # columns): every later call reuses them, captured replays included.
            self.maps_built = True

    # ------------------------------------------------------------------------------
    def capture(self, grid_vertices, grid_sdf, grid_sh, target) -> None:
        """Record the step into a hipGraph. The inputs become static: replays read whatever
        their storage holds at replay time."""
        self.static_inputs = (grid_vertices, grid_sdf, grid_sh, target)
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.run(*self.static_inputs)  # warm-up outside the capture
        torch.cuda.current_stream(self.device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run(*self.static_inputs)

    def replay(self) -> None:
        self.graph.replay()

    def overflowed(self) -> bool:
        return bool(self.overflow.item())

    def live_sizes(self) -> dict:
        """Vertex and face counts, the key count summed over camera groups, and the
        largest and the last group's key counts."""
        keys = self.group_keys.tolist()
        return {
            "V_t": int(self.vertex_offsets[-1]),
            "F": int(self.face_offsets[-1]) // 3,
            "K": int(sum(keys)),
            "K_max_group": int(max(keys)),
            "K_last_group": int(keys[-1]),
        }


def count_keys(vertices: torch.Tensor, faces: torch.Tensor, extrinsics: torch.Tensor,
               intrinsics: torch.Tensor, image_shape, tile_shape=(16, 16), near_plane=0.2,
               cull: bool = True) -> int:
    """Keys one camera needs (reads the count back): ``mt_key_block_counts`` on its
    projected vertices, with the batched step's culling."""
    from . import triangle_rasterization as tr

    th, tw = tile_shape
    h, w = image_shape
    f = faces.shape[0]
    if f == 0:
        return 0
    xy, depth = tr.project_vertices(vertices, extrinsics, intrinsics)
    n_blk = (f + 255) // 256
    counts = torch.empty(n_blk, dtype=torch.int32, device=vertices.device)
    st = torch.cuda.current_stream(vertices.device).cuda_stream
    _native.call("mt_key_block_counts", xy.contiguous().data_ptr(), depth.contiguous().data_ptr(),
                 vertices.shape[0], faces.contiguous().data_ptr(), f, None, 1, th, tw, 0, 0,
                 -(-h // th), -(-w // tw), float(near_plane), None, int(cull), counts.data_ptr(), st)
    return int(counts.to(torch.int64).sum())


def measure_capacities(voxels: Voxels, level_sets, extrinsics, intrinsics, image_shape, grid_vertices,
                       grid_sdf, grid_sh, tile_shape=(16, 16), near_plane=0.2,
                       camera_group: int | None = None, cull_keys: bool = True) -> Capacities:
    """Exact vertex and face counts of one eager forward and the largest key count of a
    camera group (reads the sizes back)."""
    from . import triangle_extraction as te

    b = extrinsics.shape[0]
    shapes = camera_shapes(image_shape, b)
    with torch.no_grad():
        # Positions and faces do not depend on the SH: tessellate with the DC band only.
        tri = te.tessellate_voxels(grid_sdf, grid_vertices, grid_sh[:1], voxels.neighbors,
                                   voxels.lower_corners, voxels.upper_corners, voxels.indices,
                                   level_sets)
        per_camera = [count_keys(tri.vertices, tri.faces, extrinsics[c : c + 1],
                                 intrinsics[c : c + 1], shapes[c], tile_shape, near_plane,
                                 cull_keys) for c in range(b)]
    k = max(sum(per_camera[c0:c1]) for c0, c1 in camera_groups(b, camera_group))
    return Capacities(int(tri.vertices.shape[0]), int(tri.faces.shape[0]), k)
