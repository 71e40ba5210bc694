"""Oracle for multi-component tesseract grids (meshtryoshka/model/tesseract.py:85-141,
183-194, 261-372, 555-560), the initializers' SDF offsets (initializer_sphere.py:24-31,
initializer_point_cloud.py:57-95), the shell conversion (tessellator.py:14-18) and the
contraction's inverse (contraction.py:74-108).

Test infrastructure only (see oracle/__init__.py). Each component's corner table is
written out from its rule rather than built by flipping the centre's table; the trilinear
map is evaluated as three nested linear interpolations in float64 (the product sums eight
weighted corners in float32; on dyadic grids both are exact). Negative lower corners
(missing vertices of subvoxel columns) are kept negative through the shift and the weld;
the reference shifts and re-indexes them (torch wraps negative indices).
"""

from __future__ import annotations

import math

import numpy as np

from .extraction import Voxels, extract_voxels
from .subdivision import index_vertices

F32 = np.float32


def _unit(i, j, k):
    return np.array([k, j, i], dtype=np.float64)  # corner (i, j, k) sits at (x, y, z) = (k, j, i)


def component_table(center: int, arm: int, extent: float, disable_tesseract: bool = False):
    """[(shape, corners f64[2, 2, 2, 3])] for the centre and the six arms, in the reference's
    order: arms along grid axis 0, 1, 2, low side before high side (tesseract.py:183-194)."""
    e = 1.0 if disable_tesseract else float(F32(extent))
    lo, hi = 0.5 - 0.5 * e, 0.5 + 0.5 * e

    def centre_corner(i, j, k):
        return lo + (hi - lo) * _unit(i, j, k)

    centre = np.array([[[centre_corner(i, j, k) for k in (0, 1)] for j in (0, 1)] for i in (0, 1)])
    if disable_tesseract:
        return [((center + 2 * arm,) * 3, centre)]
    out = [((center,) * 3, centre)]
    if arm <= 0:
        return out
    for axis in (0, 1, 2):
        for side in (0, 1):
            table = np.empty((2, 2, 2, 3))
            for i in (0, 1):
                for j in (0, 1):
                    for k in (0, 1):
                        idx = [i, j, k]
                        if idx[axis] == side:  # the outer face: on the unit cube
                            table[i, j, k] = _unit(i, j, k)
                        else:  # the inner face: the centre's face on `side`
                            idx[axis] = side
                            table[i, j, k] = centre_corner(*idx)
            shape = [center] * 3
            shape[axis] = arm
            out.append((tuple(shape), table))
    return out


def trilinear(xyz, table) -> np.ndarray:
    """Unit-cube points f32[N, 3] through a corner table, as nested lerps along x, y, z."""
    p = np.asarray(xyz, dtype=np.float64)
    x, y, z = p[:, 0:1], p[:, 1:2], p[:, 2:3]

    def lerp(a, b, t):
        return a + (b - a) * t

    # table[i, j, k]: i along z, j along y, k along x.
    c = [[lerp(table[i, j, 0], table[i, j, 1], x) for j in (0, 1)] for i in (0, 1)]
    d = [lerp(c[i][0], c[i][1], y) for i in (0, 1)]
    return lerp(d[0], d[1], z)


def extract_tesseract(flat_occupancy, center: int, arm: int, extent: float,
                      disable_tesseract: bool = False, tol: float = 1e-6):
    """(Voxels with welded corner lists, vertex boundaries) of a flat packed occupancy."""
    flat = np.asarray(flat_occupancy, dtype=np.int32).reshape(-1)
    parts, cells, word = [], [], 0
    for shape, table in component_table(center, arm, extent, disable_tesseract):
        i, j, k = shape
        n_words = i * j * k // 32
        v = extract_voxels(flat[word : word + n_words].reshape(i, j, k // 32))
        word += n_words
        v.vertices = trilinear(v.vertices, table).astype(F32)
        parts.append(v)
        cells.append(i * j * k)
    assert word == flat.size, (word, flat.size)
    nvox = sum(p.num_voxels for p in parts)
    vertices, neighbors, vox_cols, sub_cols, uppers, indices = [], [], [], [], [], []
    v0 = x0 = s0 = c0 = 0
    bounds = [0]
    for p, n_cells in zip(parts, cells):
        nv = p.num_voxels
        nb = p.neighbors.astype(np.int64)
        nb = np.where(nb < nv, nb + x0, nvox + s0 + (nb - nv))
        neighbors.append(nb)
        lc = p.lower_corners.astype(np.int64)
        lc = np.where(lc >= 0, lc + v0, lc)  # missing corners (negative) stay missing
        vox_cols.append(lc[:, :nv])
        sub_cols.append(lc[:, nv:])
        uppers.append(p.upper_corners.astype(np.int64) + v0)
        indices.append(p.indices.astype(np.int64) + c0)
        vertices.append(p.vertices)
        v0 += p.vertices.shape[0]
        x0 += nv
        s0 += p.lower_corners.shape[1] - nv
        c0 += n_cells
        bounds.append(v0)
    verts = np.concatenate(vertices, axis=0)
    first = index_vertices(verts, tol).astype(np.int64)
    lc = np.concatenate(vox_cols + sub_cols, axis=1)
    lower = np.where(lc >= 0, first[np.maximum(lc, 0)], lc)
    upper = first[np.concatenate(uppers, axis=1)]
    n = center + 2 * arm
    vox = Voxels(verts, np.concatenate(neighbors, axis=1).astype(np.int32),
                 lower.astype(np.int32), upper.astype(np.int32),
                 np.concatenate(indices).astype(np.int32), (n, n, n))
    return vox, tuple(bounds)


def sphere_offset(vertices, radius: float) -> np.ndarray:
    """initializer_sphere.py:24-31: |xyz - 0.5| - radius, in float64."""
    d = np.asarray(vertices, dtype=np.float64) - 0.5
    return np.sqrt((d * d).sum(axis=1)) - radius


def level_sets_from_shells(shells, sharpness: float) -> np.ndarray:
    """tessellator.py:14-18, in float64."""
    c = np.asarray(shells, dtype=np.float64)
    return np.log(c / (1 - c)) / sharpness


def uncontract(xyz, center_radius: float, arm_ratio: float, center_extent: float) -> np.ndarray:
    """contraction.py:74-108 in float64, one point at a time."""
    out = []
    for q in np.clip(np.asarray(xyz, dtype=np.float64), 0, 1) * 2 - 1:
        r = np.abs(q).max()
        if r <= center_extent:
            out.append(q * center_radius / center_extent)
            continue
        t_c = (r - center_extent) / (1 - center_extent)
        far = math.exp(2 * arm_ratio)
        t_u = (math.exp(2 * t_c * arm_ratio) - 1) / (far - 1)
        scale = (far * center_radius * t_u + center_radius * (1 - t_u)) / (t_c + center_extent * (1 - t_c))
        out.append(q * scale)
    return np.array(out).reshape(-1, 3)
