"""Autotuner for the batched training step: tile shape x LDS key-chunk size.

The tile shape changes the number of keys (a triangle is keyed once per overlapped
tile) and the workgroup size of the compositing kernels; the chunk size changes how many
triangles a workgroup stages per LDS round. Images do not depend on either (the
traversal order within a tile is the global (depth, shell, face) order), so the tuner
only measures time: each candidate is built with its own capacities, warmed up, and
timed with HIP events over a few eager steps on the current stream.
"""

from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass

import torch

from . import batched

# (height, width) tile shapes and staged-triangle chunk sizes searched.
TILE_SHAPES = ((8, 8), (16, 16), (32, 8), (16, 32))
CHUNKS = (128, 256, 512)


@dataclass
class Choice:
    tile_shape: tuple[int, int]
    chunk: int
    ms: float


def _time_step(step, inputs, warmup: int, reps: int) -> float:
    for _ in range(warmup):
        step.run(*inputs)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        step.run(*inputs)
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def autotune(voxels, level_sets, extrinsics, intrinsics, image_shape, num_sh, sh_degree, sharpness,
             background, inputs, tile_shapes=TILE_SHAPES, chunks=CHUNKS, margin=1.15, warmup=1,
             reps=3, near_plane=0.2, log=None,
             camera_group: int | None = None) -> tuple[Choice, list[Choice]]:
    """Returns (best, all measured candidates). ``inputs`` = (grid_vertices, grid_sdf,
    grid_sh, target) as passed to ``BatchedTrainingStep.run``."""
    grid_vertices, grid_sdf, grid_sh, _ = inputs
    results = []
    for ts in tile_shapes:
        cap = batched.measure_capacities(voxels, level_sets, extrinsics, intrinsics, image_shape,
                                         grid_vertices, grid_sdf, grid_sh, tile_shape=ts,
                                         near_plane=near_plane,
                                         camera_group=camera_group).scaled(margin)
        for ch in chunks:
            step = batched.BatchedTrainingStep(voxels, level_sets, extrinsics, intrinsics,
                                               image_shape, num_sh, sh_degree, sharpness,
                                               background, cap, tile_shape=ts, chunk=ch,
                                               near_plane=near_plane, camera_group=camera_group)
            ms = _time_step(step, inputs, warmup, reps)
            if step.overflowed():
                raise RuntimeError(f"capacity overflow while tuning tile {ts} chunk {ch}")
            results.append(Choice(tuple(ts), ch, ms))
            if log:
                log(f"[autotune] tile {ts[0]}x{ts[1]} chunk {ch}: {ms:.3f} ms/step")
            del step
            torch.cuda.empty_cache()
    best = min(results, key=lambda c: c.ms)
    return best, results


def save(path: str, key: str, best: Choice, results: list[Choice]) -> None:
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = {"best": asdict(best), "all": [asdict(r) for r in results]}
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
