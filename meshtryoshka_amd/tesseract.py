"""Multi-component sparse grids ("tesseracts") as the reference model stores them, so that a
trained scene's checkpoint can be loaded and rendered (meshtryoshka/model/tesseract.py).

A tesseract is a cubic centre grid plus, when the arm resolution is positive, six arm
grids, one per face of the centre. Each component is a regular grid of cells over the unit
cube, bent into its place by a trilinear map of its eight corners (tesseract.py:85-141): the
centre's corners span [0.5 - e / 2, 0.5 + e / 2]^3 (e = ``center_extent``); an arm's corners
are the centre face it grows from and the matching face of the outer unit cube, so the six
arms fill the shell between the centre and the unit cube as truncated pyramids. The
components' packed occupancies are stored one after another in one flat int32 tensor
(tesseract.py:261-264, 328-344).

``extract_tesseract_voxels`` builds the scene's voxel structure the way the reference does
(tesseract.py:346-372): voxels of each component (``extract_voxels``, on the GPU), vertices
mapped through the component's corners, the components concatenated with every index
shifted (voxel columns of all components first, then their subvoxel columns), and vertices
that coincide (within 1e-6) welded by pointing all corner lists at the first copy
(``index_vertices``, on the GPU; tesseract.py:555-560). Vertex rows are not removed: the
per-vertex parameters of a checkpoint have one row per component vertex.

``sdf_from_raw`` turns the stored SDF parameters into the field the reference renders
(Tesseract.signed_distances, tesseract.py:229-236, through the initializer's
get_sdf_modify); ``Contraction.uncontract`` maps tessellated vertices from the unit cube to
world space (contraction.py:74-108, applied at tessellator.py:57).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Mapping, NamedTuple

import torch
from torch import Tensor

from . import triangle_extraction as te
from .triangle_extraction.create_voxels import Voxels

WELD_TOLERANCE = 1e-6  # tesseract.py:425, 522


def value_at_step(schedule, step: int):
    """The value of a step schedule ({first step: value}) at ``step``: the entry with the
    largest first step not after ``step`` (utils.py:35-48). A plain value is its own
    schedule."""
    if not isinstance(schedule, Mapping):
        return schedule
    live = [int(s) for s in schedule if int(s) <= step]
    if not live:
        raise ValueError(f"the schedule {dict(schedule)} has no entry at or before step {step}")
    best = max(live)
    for s, v in schedule.items():
        if int(s) == best:
            return v
    raise AssertionError


def unit_cube_corners() -> Tensor:
    """f32[2, 2, 2, 3]: corner (i, j, k) of the unit cube is the point (x, y, z) = (k, j, i);
    the grid's first index runs along z, its last along x, as in the voxel vertices."""
    c = torch.empty((2, 2, 2, 3), dtype=torch.float32)
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                c[i, j, k] = torch.tensor([k, j, i], dtype=torch.float32)
    return c


class Component(NamedTuple):
    shape: tuple[int, int, int]  # cells along (i, j, k)
    corners: Tensor  # f32[2, 2, 2, 3]: where each unit-cube corner goes

    def map(self, xyz: Tensor) -> Tensor:
        """Trilinear map of unit-cube points f32[..., 3] through the corners: the sum over
        the eight corners, in (i, j, k) order, of the corner's point times the product of
        its per-axis weights (x where the corner's coordinate is 1, 1 - x where it is 0)."""
        unit = unit_cube_corners().to(xyz.device)
        corners = self.corners.to(xyz.device)
        out = torch.zeros_like(xyz)
        for i in (0, 1):
            for j in (0, 1):
                for k in (0, 1):
                    c = unit[i, j, k]
                    per_axis = c * xyz + (1 - xyz) * (1 - c)
                    w = per_axis[..., 0] * per_axis[..., 1] * per_axis[..., 2]
                    out = out + w[..., None] * corners[i, j, k]
        return out


@dataclass(frozen=True)
class TesseractLayout:
    """The grid-shape part of the reference's TesseractCfg (tesseract.py:45-72): resolution
    schedules ({first step: cells}, or a plain int) for the centre and the arms, the
    centre's share of the unit cube, and the no-tesseract ablation (one centre grid of
    centre + 2 arm cells, spanning the whole cube)."""

    center_resolution: Mapping[int, int] | int
    arm_resolution: Mapping[int, int] | int
    center_extent: float
    disable_tesseract: bool = False

    @classmethod
    def single_grid(cls, resolution: int) -> "TesseractLayout":
        """One cubic grid over the whole unit cube (arm resolution 0, extent 1)."""
        return cls(int(resolution), 0, 1.0)

    def has_arms(self) -> bool:
        return value_at_step(self.arm_resolution, 0) > 0 and not self.disable_tesseract

    def extent(self, step: int) -> int:
        """Cells along each axis of the cube that holds the whole tesseract."""
        return value_at_step(self.center_resolution, step) + 2 * value_at_step(self.arm_resolution, step)

    def center_corners(self) -> Tensor:
        e = torch.tensor(1.0 if self.disable_tesseract else self.center_extent, dtype=torch.float32)
        lo, hi = 0.5 - 0.5 * e, 0.5 + 0.5 * e
        return lo + (hi - lo) * unit_cube_corners()

    def components(self, step: int) -> list[Component]:
        """The centre, then (with arms) the arms along grid axis 0, 1, 2, each towards the
        low side first (tesseract.py:183-194)."""
        c = value_at_step(self.center_resolution, step)
        a = value_at_step(self.arm_resolution, step)
        centre = self.center_corners()
        if self.disable_tesseract:
            return [Component((c + 2 * a,) * 3, centre)]
        out = [Component((c, c, c), centre)]
        if not self.has_arms():
            return out
        unit = unit_cube_corners()
        for axis in (0, 1, 2):
            for side in (0, 1):
                # The centre's face on `side` becomes the arm's inner face (the flip puts it
                # opposite the outer face); the corners on `side` move to the unit cube.
                corners = centre.flip(axis).clone()
                sel = [slice(None)] * 3
                sel[axis] = side
                corners[tuple(sel)] = unit[tuple(sel)]
                shape = [c, c, c]
                shape[axis] = a
                out.append(Component(tuple(shape), corners))
        return out

    def occupancy_words(self, step: int) -> list[tuple[int, int]]:
        """[start, end) of each component's words in the flat packed occupancy."""
        bounds, start = [], 0
        for comp in self.components(step):
            cells = comp.shape[0] * comp.shape[1] * comp.shape[2]
            if comp.shape[2] % 32:
                raise ValueError(f"component shape {comp.shape}: the last axis must be a multiple of 32")
            bounds.append((start // 32, (start + cells) // 32))
            start += cells
        return bounds


def split_occupancy(occupancy: Tensor, layout: TesseractLayout, step: int) -> list[Tensor]:
    """The components' packed grids int32[i, j, k / 32] (views of the flat tensor)."""
    flat = occupancy.reshape(-1)
    words = layout.occupancy_words(step)
    if flat.numel() != words[-1][1]:
        raise ValueError(f"the occupancy holds {flat.numel()} words; the layout at step {step} "
                         f"has {words[-1][1]} ({[c.shape for c in layout.components(step)]})")
    return [flat[w0:w1].view(c.shape[0], c.shape[1], c.shape[2] // 32)
            for c, (w0, w1) in zip(layout.components(step), words)]


def concatenate_voxels(parts: list[Voxels], cell_counts: list[int]) -> Voxels:
    """One voxel structure over several components: vertices stacked; corner lists shifted
    by the vertex rows before them; voxel columns of every component first, then every
    component's subvoxel columns, and neighbour indices shifted to match; cell indices
    shifted by the cells of the components before (tesseract.py:266-325). A negative lower
    corner (a subvoxel column's missing +x / +y / +z vertex) stays negative; the reference
    shifts it like an index."""
    n_vox = [p.num_voxels for p in parts]
    n_sub = [p.num_subvoxels for p in parts]
    total_vox = sum(n_vox)
    vert0 = vox0 = sub0 = cell0 = 0
    neighbors, vox_cols, sub_cols, uppers, indices = [], [], [], [], []
    for p, nv, ns, cells in zip(parts, n_vox, n_sub, cell_counts):
        nb = p.neighbors
        neighbors.append(torch.where(nb < nv, nb + vox0, nb - nv + total_vox + sub0))
        lower = torch.where(p.lower_corners >= 0, p.lower_corners + vert0, p.lower_corners)
        vox_cols.append(lower[:, :nv])
        sub_cols.append(lower[:, nv:])
        uppers.append(p.upper_corners + vert0)
        indices.append(p.indices + cell0)
        vert0 += p.num_vertices
        vox0 += nv
        sub0 += ns
        cell0 += cells
    return Voxels(torch.cat([p.vertices for p in parts], dim=0),
                  torch.cat(neighbors, dim=1).contiguous(),
                  torch.cat(vox_cols + sub_cols, dim=1).contiguous(),
                  torch.cat(uppers, dim=1).contiguous(),
                  torch.cat(indices, dim=0).contiguous(),
                  (0, 0, 0))


def weld_vertices(voxels: Voxels, tol: float = WELD_TOLERANCE) -> Voxels:
    """Corner lists re-pointed at the first vertex row of each welded position; negative
    lower corners (missing vertices) are kept. (The reference indexes its map with them,
    tesseract.py:555-560, and torch's wrap-around makes them arbitrary vertex rows: their
    edges then emit triangle vertices that no face uses. Faces and images are the same.)"""
    first = te.index_vertices(voxels.vertices, tol)
    lower = voxels.lower_corners
    welded = first[lower.clamp(min=0).long()]
    return voxels._replace(lower_corners=torch.where(lower >= 0, welded, lower).contiguous(),
                           upper_corners=first[voxels.upper_corners.long()].contiguous())


@torch.no_grad()
def extract_tesseract_voxels(occupancy: Tensor, layout: TesseractLayout, step: int
                             ) -> tuple[Voxels, tuple[int, ...]]:
    """The scene's voxels from the flat packed occupancy (on the GPU), and the first vertex
    row of each component (plus the total) (tesseract.py:346-372)."""
    comps = layout.components(step)
    parts, counts = [], []
    for comp, grid in zip(comps, split_occupancy(occupancy, layout, step)):
        v = te.extract_voxels(grid.contiguous())
        parts.append(v._replace(vertices=comp.map(v.vertices).contiguous()))
        counts.append(comp.shape[0] * comp.shape[1] * comp.shape[2])
    boundaries = [0]
    for p in parts:
        boundaries.append(boundaries[-1] + p.num_vertices)
    vox = concatenate_voxels(parts, counts)
    n = layout.extent(step)
    vox = weld_vertices(vox._replace(grid_shape=(n, n, n)))
    return vox, tuple(boundaries)


# ------------------------------------------------------------------------------------
# SDF parameterisation and world space
# ------------------------------------------------------------------------------------

def _initializer_offset(initializer: Mapping, vertices: Tensor, step: int):
    """What the reference's initializer adds to the raw SDF at ``step`` (None: nothing)."""
    name = initializer.get("name")
    if name == "sphere":  # initializer_sphere.py:24-31
        return torch.linalg.vector_norm(vertices - 0.5, dim=-1) - float(initializer["radius"])
    if name == "point_cloud":  # initializer_point_cloud.py:57-95
        done = int(initializer["interpolate_steps"]) + int(initializer["num_frozen_steps"])
        if step >= done:
            return None
        raise ValueError(f"at step {step} the point-cloud initializer still blends in the "
                         f"distance to the scene's point cloud (until step {done}); the "
                         "point cloud is not part of the checkpoint")
    if name == "none":
        return None
    raise ValueError(f"unknown initializer {initializer!r}: name one of 'sphere' (with its "
                     "radius), 'point_cloud' (with interpolate_steps and num_frozen_steps) or "
                     "'none'")


def sdf_from_raw(raw: Tensor, vertices: Tensor, initializer: Mapping, step: int) -> Tensor:
    """The SDF the reference renders from the stored parameters (tesseract.py:229-236)."""
    off = _initializer_offset(initializer, vertices, step)
    return raw if off is None else raw + off


def raw_from_sdf(sdf: Tensor, vertices: Tensor, initializer: Mapping, step: int) -> Tensor:
    """The stored parameters whose rendered SDF is ``sdf`` (to float32 rounding)."""
    off = _initializer_offset(initializer, vertices, step)
    return sdf if off is None else sdf - off


def level_sets_from_shells(shells, sharpness: float) -> Tensor:
    """SDF level sets f32[S] of shells given as opacity-CDF values, log(c / (1 - c)) / k
    (tessellator.py:14-18, 111-118)."""
    c = torch.as_tensor(shells, dtype=torch.float32)
    return torch.log(c / (1 - c)) / sharpness


@dataclass(frozen=True)
class Contraction:
    """The scene contraction's inverse (contraction.py:10-22, 74-108): unit-cube points to
    world space. Inside the centre cube (Chebyshev radius <= center_extent after mapping to
    [-1, 1]) a uniform scale to ``center_radius``; outside, the radius grows exponentially
    with the arm parameter so that the unit cube's faces reach center_radius *
    exp(2 arm_ratio)."""

    center_radius: float
    arm_ratio: float
    center_extent: float
    disable_exponential: bool = False

    @torch.no_grad()
    def uncontract(self, xyz: Tensor) -> Tensor:
        p = xyz.clip(min=0, max=1) * 2 - 1
        if self.disable_exponential:
            return p * math.exp(2 * self.arm_ratio)
        e, r0 = self.center_extent, self.center_radius
        inside = (p.abs() <= e).all(dim=-1)
        out = torch.empty_like(p)
        out[inside] = p[inside] * r0 / e
        q = p[~inside]
        if q.numel():
            grow = math.exp(2 * self.arm_ratio)
            t_c = (q.abs().max(dim=-1).values - e) / (1 - e)
            t_u = ((2 * t_c * self.arm_ratio).exp() - 1) / (grow - 1)
            num = grow * r0 * t_u + r0 * (1 - t_u)
            den = t_c + e * (1 - t_c)
            out[~inside] = q * (num / den)[..., None]
        return out
