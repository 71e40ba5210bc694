"""Render serving with device-resident scenes and continuous batching.

* ``ResidentScene``: a scene's extracted shells (vertices, faces, shell boundaries, SDF,
  SH) kept on the device; tessellation happens once, at load.
* ``render_batch``: forward-only fused render of B cameras of one scene and one
  resolution, launched on the current stream with no host synchronisation: the key
  buffers are sized from the scene's measured keys-per-pixel density (with margin) and a
  device overflow flag reports the rare batch that needs a re-render with more room.
* ``RenderServer``: requests (scene, pose, resolution) wait in a FIFO queue; every
  scheduling round admits queued requests while the estimated HBM footprint of the
  in-flight work stays within the budget, groups the admitted ones by (scene,
  resolution) into batches of up to ``max_batch`` cameras, and launches each batch on
  its own stream, so batches of different scenes and resolutions overlap on the device.
  Completed batches (event query, no blocking) release their memory, which admits more
  requests: requests that arrive while others run join at the next round.

Each camera's pixels depend only on that camera's keys (the sort is stable and keys carry
the camera in their top bits), so a request's image is the same alone and in any batch.
"""

from __future__ import annotations

import heapq
import itertools
import time
from collections import deque
from dataclasses import dataclass, field

import torch

from . import _native
from .triangle_rasterization.common import TileGrid
from .triangle_rasterization.keys import create_key_defines, to_layout

_LAYOUT = to_layout(create_key_defines(("camera", 6), ("tile", 20), ("depth", 32), ("shell", 6)))


@dataclass
class ResidentScene:
    vertices: torch.Tensor  # f32[V, 3]
    faces: torch.Tensor  # int32[F, 3]
    face_boundaries: torch.Tensor  # int32[S]
    signed_distances: torch.Tensor  # f32[V]
    spherical_harmonics: torch.Tensor  # f32[NSH, V, 3]
    sh_degree: int
    sharpness: torch.Tensor  # f32[1]
    background: torch.Tensor  # f32[3]
    keys_per_pixel: float = 8.0  # measured density of (face, tile) keys per image pixel

    @property
    def num_shells(self) -> int:
        return int(self.face_boundaries.shape[0])

    def bytes_per_camera(self, h: int, w: int, margin: float) -> int:
        v = self.vertices.shape[0]
        f = self.faces.shape[0]
        keys = int(self.keys_per_pixel * h * w * margin) + 4096
        # xy + depth + colours per vertex, bounds + counts + offsets per face,
        # keys / values twice + sort workspace, output image.
        return v * 24 + f * 28 + keys * 37 + h * w * 16


def tessellate_scene(voxels, grid_vertices, grid_sdf, grid_sh, level_sets, sh_degree, sharpness,
                     background) -> ResidentScene:
    from . import triangle_extraction as te

    with torch.no_grad():
        tri = te.tessellate_voxels(grid_sdf, grid_vertices, grid_sh, voxels.neighbors,
                                   voxels.lower_corners, voxels.upper_corners, voxels.indices,
                                   level_sets)
    dev = grid_sdf.device
    return ResidentScene(
        tri.vertices.detach(), tri.faces, tri.face_boundaries, tri.signed_distances.detach(),
        tri.spherical_harmonics.detach(), sh_degree,
        torch.tensor([float(sharpness)], device=dev, dtype=torch.float32),
        torch.as_tensor(background, dtype=torch.float32, device=dev).contiguous(),
    )


@dataclass
class BatchResult:
    image: torch.Tensor  # f32[B, 3, H, W] (background composited, clipped)
    overflow: torch.Tensor  # int32[1]
    nbytes: int


def render_batch(scene: ResidentScene, extrinsics: torch.Tensor, intrinsics: torch.Tensor,
                 image_shape, tile_shape=(16, 16), chunk=0, near_plane=0.2,
                 key_margin=1.5, key_capacity=None) -> BatchResult:
    """Forward-only fused render of B cameras of one scene, no host sync."""
    call = _native.call
    dev = scene.vertices.device
    st = torch.cuda.current_stream(dev).cuda_stream
    b = extrinsics.shape[0]
    h, w = image_shape
    s = scene.num_shells
    grid = TileGrid(tuple(tile_shape), (h, w), b, s)
    rows, cols = grid.grid_shape
    t = rows * cols
    v = scene.vertices.shape[0]
    f = scene.faces.shape[0]
    k_cap = key_capacity or int(scene.keys_per_pixel * h * w * b * key_margin) + 4096
    E = extrinsics.contiguous().float()
    K = intrinsics.contiguous().float()
    f32 = dict(dtype=torch.float32, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    xy = torch.empty((b, v, 2), **f32)
    depth = torch.empty((b, v), **f32)
    tmin = torch.empty((b, f, 2), **i32)
    tmax = torch.empty((b, f, 2), **i32)
    counts = torch.empty((b, f), **i32)
    offsets = torch.empty(b * f, dtype=torch.int64, device=dev)
    keys = torch.empty(k_cap, dtype=torch.int64, device=dev)
    vals = torch.empty(k_cap, **i32)
    skeys = torch.empty_like(keys)
    svals = torch.empty_like(vals)
    ranges = torch.zeros((b, 1, t, 2), dtype=torch.int64, device=dev)
    campos = torch.empty((b, 3), **f32)
    colors = torch.empty((b, v, 3), **f32)
    out = torch.empty((b, 4, h, w), **f32)
    overflow = torch.zeros(1, **i32)
    scan_ws = _native.workspace(_native.query("mt_scan_workspace_bytes", b * f), dev)
    sort_ws = _native.workspace(_native.query("mt_sort_workspace_bytes", k_cap), dev)
    th, tw = tile_shape
    call("mt_project_vertices", scene.vertices.data_ptr(), v, E.data_ptr(), K.data_ptr(), b,
         xy.data_ptr(), depth.data_ptr(), st)
    call("mt_bound_triangles", xy.data_ptr(), depth.data_ptr(), v, scene.faces.data_ptr(), f, b,
         th, tw, 0, 0, rows, cols, float(near_plane), tmin.data_ptr(), tmax.data_ptr(),
         counts.data_ptr(), None, st)
    call("mt_inclusive_scan_i32_to_i64", counts.data_ptr(), offsets.data_ptr(), b * f,
         scan_ws.data_ptr(), scan_ws.numel(), st)
    n_dev = offsets[b * f - 1 :]
    call("mt_generate_keys", depth.data_ptr(), v, scene.faces.data_ptr(), f,
         scene.face_boundaries.data_ptr(), s, tmin.data_ptr(), tmax.data_ptr(), offsets.data_ptr(),
         b, cols, _LAYOUT, keys.data_ptr(), vals.data_ptr(), k_cap, overflow.data_ptr(), st)
    call("mt_sort_keys", keys.data_ptr(), vals.data_ptr(), k_cap, skeys.data_ptr(), svals.data_ptr(),
         sort_ws.data_ptr(), sort_ws.numel(), n_dev.data_ptr(), st)
    call("mt_delineate_tiles", skeys.data_ptr(), k_cap, _LAYOUT, 0, b, 1, t, ranges.data_ptr(),
         n_dev.data_ptr(), None, st)
    call("mt_camera_positions", E.data_ptr(), b, campos.data_ptr(), st)
    call("mt_evaluate_sh_forward", scene.vertices.data_ptr(), v, scene.spherical_harmonics.data_ptr(),
         scene.spherical_harmonics.shape[0], scene.sh_degree, campos.data_ptr(), b,
         colors.data_ptr(), st)
    args = _native.CompositeArgs(
        xy.data_ptr(), colors.data_ptr(), scene.signed_distances.data_ptr(), v,
        scene.faces.data_ptr(), scene.sharpness.data_ptr(), skeys.data_ptr(), svals.data_ptr(),
        ranges.data_ptr(), b, s, rows, cols, th, tw, h, w, _LAYOUT, chunk, None,
    )
    call("mt_intersect_and_composite_forward", args, out.data_ptr(), st)
    image = (out[:, :3] + out[:, 3:4] * scene.background[:, None, None]).clip(0, 1)
    nbytes = b * scene.bytes_per_camera(h, w, key_margin)
    return BatchResult(image, overflow, nbytes)


def calibrate_key_density(scene: ResidentScene, extrinsics, intrinsics, image_shape,
                          tile_shape=(16, 16), near_plane=0.2) -> float:
    """Keys per pixel of a sample of poses (reads one size back; done at scene load)."""
    from . import triangle_rasterization as tr

    b = extrinsics.shape[0]
    grid = TileGrid(tuple(tile_shape), tuple(image_shape), b, scene.num_shells)
    with torch.no_grad():
        xy, depth = tr.project_vertices(scene.vertices, extrinsics, intrinsics)
        bounds = tr.bound_triangles(xy, depth, scene.faces, grid, near_plane)
        k = int(bounds.num_tiles_overlapped.to(torch.int64).sum())
    h, w = image_shape
    return k / (b * h * w)


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
                 tile_shape=(16, 16), chunk=0, key_margin=1.5, keep_images=False):
        self.scenes = scenes
        self.budget = int(hbm_budget_bytes)
        self.max_batch = max_batch
        self.tile_shape = tile_shape
        self.chunk = chunk
        self.key_margin = key_margin
        self.keep_images = keep_images
        dev = next(iter(scenes.values())).vertices.device
        self.device = dev
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(num_streams)]
        self._stream_cycle = itertools.cycle(self.streams)
        self.queue: deque[Request] = deque()
        self.inflight: list[_InFlight] = []
        self.in_use = 0
        self.completed: list[Request] = []
        self.rerenders = 0

    def submit(self, req: Request) -> None:
        self.queue.append(req)

    def _footprint(self, req: Request) -> int:
        h, w = req.image_shape
        return self.scenes[req.scene_id].bytes_per_camera(h, w, self.key_margin)

    def schedule(self) -> int:
        """Admit queued requests that fit the budget and launch them; returns batches launched."""
        admitted = []
        budget_left = self.budget - self.in_use
        while self.queue:
            need = self._footprint(self.queue[0])
            if need > budget_left and (self.inflight or admitted):
                break  # wait for memory to free up (FIFO: no overtaking)
            req = self.queue.popleft()
            admitted.append(req)
            budget_left -= need
        groups: dict = {}
        for r in admitted:
            groups.setdefault((r.scene_id, tuple(r.image_shape)), []).append(r)
        launched = 0
        for (sid, shape), reqs in groups.items():
            for i in range(0, len(reqs), self.max_batch):
                self._launch(sid, shape, reqs[i : i + self.max_batch])
                launched += 1
        return launched

    def _launch(self, sid, shape, reqs, key_capacity=None) -> None:
        scene = self.scenes[sid]
        stream = next(self._stream_cycle)
        with torch.cuda.stream(stream):
            E = torch.stack([r.extrinsics for r in reqs]).to(self.device, non_blocking=True)
            K = torch.stack([r.intrinsics for r in reqs]).to(self.device, non_blocking=True)
            res = render_batch(scene, E, K, shape, self.tile_shape, self.chunk,
                               key_margin=self.key_margin, key_capacity=key_capacity)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.in_use += res.nbytes
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
            self.in_use -= job.result.nbytes
            if int(job.result.overflow.item()):  # the batch has completed: this read does not wait
                self.rerenders += 1
                b = len(job.requests)
                h, w = job.requests[0].image_shape
                sid = job.requests[0].scene_id
                cap = int(self.scenes[sid].keys_per_pixel * h * w * b * self.key_margin * 4) + 4096
                self._launch(sid, (h, w), job.requests, key_capacity=cap)
                continue
            for i, r in enumerate(job.requests):
                r.done = now
                if self.keep_images:
                    r.image = job.result.image[i]
                self.completed.append(r)
                done += 1
        return done

    def drain(self) -> None:
        while self.queue or self.inflight:
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
    scheduling and polling continuously; returns throughput and latency percentiles."""
    import random

    rng = random.Random(seed)
    t0 = time.perf_counter()
    arrivals = []
    t = 0.0
    while t < duration_s:
        t += rng.expovariate(rate_per_s)
        arrivals.append(t)
    heapq.heapify(arrivals)
    submitted = []
    rid = 0
    while arrivals or server.queue or server.inflight:
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
    stats.update({"offered_req_per_s": rate_per_s, "completed": len(server.completed),
                  "req_per_s": round(len(submitted) / wall, 2), "wall_s": round(wall, 3),
                  "rerenders": server.rerenders})
    return stats
