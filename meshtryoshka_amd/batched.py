"""Batched, graph-capturable training step: tessellate all shells, render all cameras,
image loss, and the backward chain to the grid SDF and SH, as one launch sequence with
no host synchronisation.

The plain API (``triangle_extraction`` / ``triangle_rasterization``) mirrors the
reference and therefore reads three sizes back to the host every step (vertex, face and
key counts) to allocate exactly-sized tensors. Here every buffer is preallocated at a
capacity, kernels read the live counts from device memory (``num_faces_x3`` /
``n_device`` arguments of the C ABI) and write nothing past a capacity (they raise a
device overflow flag instead), so the whole step can be captured once into a hipGraph
(``torch.cuda.CUDAGraph``) and replayed. Results are bit-identical to the plain path:
the same arithmetic runs on the same data, only the sizes come from device memory.

Unlike the plain path, the per-triangle-vertex SH array ([NSH, V, 3], 192 bytes per vertex
at degree 3) is never stored: vertices and their colours for all cameras come out of one
kernel that interpolates the grid SH in registers, and the SH adjoint is recomputed per
triangle vertex inside the gather of the vertex adjoint. Colours stay bit-identical to the
plain path's; gradients equal it up to float summation order.

Capacities come from one eager step (``measure_capacities``) times a margin; after a
replay, ``overflowed()`` (one host read) tells whether any capacity was exceeded, in
which case the caller rebuilds the step with larger capacities and runs it again.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _native
from .triangle_extraction import Voxels
from .triangle_rasterization.common import TileGrid
from .triangle_rasterization.keys import create_key_defines, to_layout

_ptr = _native.ptr


@dataclass(frozen=True)
class Capacities:
    vertices: int
    faces: int
    keys: int

    def scaled(self, margin: float) -> "Capacities":
        return Capacities(int(self.vertices * margin) + 1024, int(self.faces * margin) + 1024,
                          int(self.keys * margin) + 4096)


class BatchedTrainingStep:
    """One fused forward + backward step over B same-resolution cameras and S shells."""

    def __init__(self, voxels: Voxels, level_sets: torch.Tensor, extrinsics: torch.Tensor,
                 intrinsics: torch.Tensor, image_shape: tuple[int, int], num_sh: int, sh_degree: int,
                 sharpness: float, background, capacities: Capacities, tile_shape=(16, 16),
                 chunk: int = 0, near_plane: float = 0.2, reference_quirk: bool = False):
        dev = voxels.lower_corners.device
        self.device = dev
        self.voxels = voxels
        self.level_sets = level_sets.contiguous().to(dev)
        self.extrinsics = extrinsics.contiguous().float().to(dev)
        self.intrinsics = intrinsics.contiguous().float().to(dev)
        self.B = self.extrinsics.shape[0]
        self.H, self.W = image_shape
        self.S = self.level_sets.shape[0]
        self.num_sh = num_sh
        self.sh_degree = sh_degree
        self.tile_shape = tuple(tile_shape)
        self.chunk = chunk
        self.near_plane = float(near_plane)
        self.quirk = int(reference_quirk)
        self.cap = capacities
        self.grid = TileGrid(self.tile_shape, (self.H, self.W), self.B, self.S)
        self.rows, self.cols = self.grid.grid_shape
        self.T = self.rows * self.cols
        self.layout = to_layout(create_key_defines(("camera", 6), ("tile", 20), ("depth", 32),
                                                   ("shell", 6)))
        self.n_col = voxels.lower_corners.shape[1]
        self.n_vox = voxels.upper_corners.shape[1]
        self.n_grid = voxels.vertices.shape[0]
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        V, F, K, B, S = self.cap.vertices, self.cap.faces, self.cap.keys, self.B, self.S
        self.vertex_offsets = torch.zeros(S * self.n_col + 1, **i32)
        self.face_offsets = torch.zeros(S * self.n_vox + 1, **i32)
        self.codes = torch.empty((S, self.n_vox), dtype=torch.uint8, device=dev)
        self.tv = torch.zeros((V, 3), **f32)
        self.td = torch.zeros(V, **f32)
        self.ttype = torch.zeros(V, dtype=torch.uint8, device=dev)
        self.faces = torch.zeros((F, 3), **i32)
        self.voxel_indices = torch.zeros(F, **i32)
        self.xy = torch.zeros((B, V, 2), **f32)
        self.depth = torch.zeros((B, V), **f32)
        # Keys per (camera, 256-face block) and their inclusive scan (the key kernels need
        # no per-face tile bounds, counts or offsets).
        self.n_blk = (F + 255) // 256
        self.block_counts = torch.zeros(B * self.n_blk, **i32)
        self.block_offsets = torch.zeros(B * self.n_blk, **i64)
        self.keys = torch.zeros(K, **i64)
        self.values = torch.zeros(K, **i32)
        self.sorted_keys = torch.zeros(K, **i64)
        self.sorted_values = torch.zeros(K, **i32)
        self.ranges = torch.zeros((B, 1, self.T, 2), **i64)
        self.campos = torch.zeros((B, 3), **f32)
        self.colors = torch.zeros((B, V, 3), **f32)
        self.out = torch.zeros((B, 4, self.H, self.W), **f32)
        self.grad_out = torch.zeros_like(self.out)
        self.loss = torch.zeros(1, **f32)
        self.colors_grad = torch.zeros_like(self.colors)
        self.td_grad = torch.zeros_like(self.td)
        self.sharp_grad = torch.zeros(1, **f32)
        self.grid_sdf_grad = torch.zeros(self.n_grid, **f32)
        self.grid_sh_grad = torch.zeros((num_sh, self.n_grid, 3), **f32)
        self.overflow = torch.zeros(1, **i32)
        self.sharpness = torch.tensor([float(sharpness)], **f32)
        self.background = torch.as_tensor(background, dtype=torch.float32, device=dev).contiguous()
        scan_n = max(S * self.n_col + 1, S * self.n_vox + 1, B * self.n_blk)
        self.scan_ws = _native.workspace(_native.query("mt_scan_workspace_bytes", scan_n), dev)
        self.sort_ws = _native.workspace(_native.query("mt_sort_workspace_bytes", K), dev)
        self.vbwd_ws = _native.workspace(
            _native.query("mt_create_triangle_vertices_colors_backward_workspace_bytes",
                          self.n_grid, self.n_col, sh_degree), dev)
        # Live counts, read by kernels from device memory.
        self.num_faces_x3 = self.face_offsets[S * self.n_vox :]
        self.num_keys = self.block_offsets[B * self.n_blk - 1 :]
        self.graph = None
        self.static_inputs = None

    # ------------------------------------------------------------------------------
    def _composite_args(self):
        th, tw = self.tile_shape
        return _native.CompositeArgs(
            self.xy.data_ptr(), self.colors.data_ptr(), self.td.data_ptr(), self.cap.vertices,
            self.faces.data_ptr(), self.sharpness.data_ptr(), self.sorted_keys.data_ptr(),
            self.sorted_values.data_ptr(), self.ranges.data_ptr(), self.B, self.S, self.rows,
            self.cols, th, tw, self.H, self.W, self.layout, self.chunk, None,
        )

    def run(self, grid_vertices: torch.Tensor, grid_sdf: torch.Tensor, grid_sh: torch.Tensor,
            target: torch.Tensor, hook=None) -> None:
        """Enqueue one forward + backward on the current stream (no host sync).

        ``hook(name)``, if given, is called after each library launch (the benchmark
        records a HIP event there to time every kernel stage)."""
        if hook is None:
            call = _native.call
        else:
            def call(name, *args):
                _native.call(name, *args)
                hook(name)
        st = torch.cuda.current_stream(self.device).cuda_stream
        v = self.voxels
        S, B = self.S, self.B
        V, F, K = self.cap.vertices, self.cap.faces, self.cap.keys
        self.loss.zero_()
        self.ranges.zero_()
        self.colors_grad.zero_()
        self.td_grad.zero_()
        self.sharp_grad.zero_()
        if hook is not None:
            hook("zero_fill")
        # --- tessellation
        call("mt_count_triangle_vertices", grid_sdf.data_ptr(), v.lower_corners.data_ptr(),
             self.n_col, self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(), st)
        call("mt_exclusive_scan_i32", self.vertex_offsets.data_ptr(), self.vertex_offsets.numel(),
             self.scan_ws.data_ptr(), self.scan_ws.numel(), st)
        call("mt_camera_positions", self.extrinsics.data_ptr(), B, self.campos.data_ptr(), st)
        # Vertices and their colours in one pass; the per-vertex SH array is never stored.
        call("mt_create_triangle_vertices_colors_forward", grid_vertices.data_ptr(),
             grid_sdf.data_ptr(), grid_sh.data_ptr(), self.n_grid, self.num_sh, self.sh_degree,
             v.lower_corners.data_ptr(), self.n_col, self.level_sets.data_ptr(), S,
             self.vertex_offsets.data_ptr(), V, self.campos.data_ptr(), B, self.tv.data_ptr(),
             self.td.data_ptr(), self.ttype.data_ptr(), self.colors.data_ptr(),
             self.overflow.data_ptr(), st)
        call("mt_count_triangle_faces", grid_sdf.data_ptr(), v.lower_corners.data_ptr(), self.n_col,
             v.upper_corners.data_ptr(), self.n_vox, self.level_sets.data_ptr(), S,
             self.face_offsets.data_ptr(), self.codes.data_ptr(), st)
        call("mt_exclusive_scan_i32", self.face_offsets.data_ptr(), self.face_offsets.numel(),
             self.scan_ws.data_ptr(), self.scan_ws.numel(), st)
        call("mt_create_triangle_faces", v.neighbors.data_ptr(), v.indices.data_ptr(), self.n_vox,
             self.n_col, self.vertex_offsets.data_ptr(), self.ttype.data_ptr(),
             self.codes.data_ptr(), S, self.face_offsets.data_ptr(), self.faces.data_ptr(),
             self.voxel_indices.data_ptr(), F, V, self.overflow.data_ptr(), st)
        face_boundaries = (self.face_offsets[self.n_vox :: self.n_vox][:S] // 3).contiguous()
        if hook is not None:
            hook("face_boundaries")
        # --- rasterization
        call("mt_project_vertices", self.tv.data_ptr(), V, self.extrinsics.data_ptr(),
             self.intrinsics.data_ptr(), B, self.xy.data_ptr(), self.depth.data_ptr(), st)
        th, tw = self.tile_shape
        call("mt_key_block_counts", self.xy.data_ptr(), self.depth.data_ptr(), V,
             self.faces.data_ptr(), F, self.num_faces_x3.data_ptr(), B, th, tw, 0, 0, self.rows,
             self.cols, self.near_plane, self.block_counts.data_ptr(), st)
        call("mt_inclusive_scan_i32_to_i64", self.block_counts.data_ptr(),
             self.block_offsets.data_ptr(), B * self.n_blk, self.scan_ws.data_ptr(),
             self.scan_ws.numel(), st)
        call("mt_generate_keys_from_geometry", self.xy.data_ptr(), self.depth.data_ptr(), V,
             self.faces.data_ptr(), F, self.num_faces_x3.data_ptr(), face_boundaries.data_ptr(), S,
             B, th, tw, 0, 0, self.rows, self.cols, self.near_plane, self.block_offsets.data_ptr(),
             self.cols, self.layout, self.keys.data_ptr(), self.values.data_ptr(), K,
             self.overflow.data_ptr(), st)
        call("mt_sort_keys", self.keys.data_ptr(), self.values.data_ptr(), K,
             self.sorted_keys.data_ptr(), self.sorted_values.data_ptr(), self.sort_ws.data_ptr(),
             self.sort_ws.numel(), self.num_keys.data_ptr(), st)
        call("mt_delineate_tiles", self.sorted_keys.data_ptr(), K, self.layout, 0, B, 1, self.T,
             self.ranges.data_ptr(), self.num_keys.data_ptr(), st)
        args = self._composite_args()
        call("mt_intersect_and_composite_forward", args, self.out.data_ptr(), st)
        # --- loss and backward chain
        call("mt_image_l2_loss", self.out.data_ptr(), self.background.data_ptr(), target.data_ptr(),
             B, self.H, self.W, self.loss.data_ptr(), self.grad_out.data_ptr(), st)
        call("mt_intersect_and_composite_backward", args, self.out.data_ptr(),
             self.grad_out.data_ptr(), self.colors_grad.data_ptr(), self.td_grad.data_ptr(),
             self.sharp_grad.data_ptr(), self.quirk, st)
        # SH adjoint fused into the gather of the vertex adjoint (overwrites the grid grads).
        call("mt_create_triangle_vertices_colors_backward", grid_sdf.data_ptr(), self.n_grid,
             self.num_sh, self.sh_degree, v.lower_corners.data_ptr(), self.n_col,
             self.level_sets.data_ptr(), S, self.vertex_offsets.data_ptr(), V, self.tv.data_ptr(),
             self.colors.data_ptr(), self.campos.data_ptr(), B, self.td_grad.data_ptr(),
             self.colors_grad.data_ptr(), self.grid_sdf_grad.data_ptr(),
             self.grid_sh_grad.data_ptr(), self.vbwd_ws.data_ptr(), self.vbwd_ws.numel(),
             self.overflow.data_ptr(), st)

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
        return {
            "V_t": int(self.vertex_offsets[-1]),
            "F": int(self.face_offsets[-1]) // 3,
            "K": int(self.num_keys[0]),
        }


def measure_capacities(voxels: Voxels, level_sets, extrinsics, intrinsics, image_shape, grid_vertices,
                       grid_sdf, grid_sh, tile_shape=(16, 16), near_plane=0.2) -> Capacities:
    """Exact vertex / face / key counts of one eager forward (reads three sizes back)."""
    from . import triangle_extraction as te
    from . import triangle_rasterization as tr

    with torch.no_grad():
        tri = te.tessellate_voxels(grid_sdf, grid_vertices, grid_sh, voxels.neighbors,
                                   voxels.lower_corners, voxels.upper_corners, voxels.indices,
                                   level_sets)
        b = extrinsics.shape[0]
        grid = TileGrid(tuple(tile_shape), tuple(image_shape), b, level_sets.shape[0])
        xy, depth = tr.project_vertices(tri.vertices, extrinsics, intrinsics)
        bounds = tr.bound_triangles(xy, depth, tri.faces, grid, near_plane)
        k = int(bounds.num_tiles_overlapped.to(torch.int64).sum())
    return Capacities(int(tri.vertices.shape[0]), int(tri.faces.shape[0]), k)
