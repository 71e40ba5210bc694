"""Oracle for triangle_extraction: voxel structure and multi-level-set marching cubes.

Test infrastructure only (see oracle/__init__.py). All float arithmetic is IEEE
float32 in the reference's operation order; integer outputs are exact.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product

import numpy as np

from . import tables

F32 = np.float32
INVALID_CORNER = -64  # create_voxels.slang:71, 120-122


# --------------------------------------------------------------------------------------
# Occupancy packing (triangle_extraction/common.py:17-35)
# --------------------------------------------------------------------------------------


def pack_occupancy(occupancy: np.ndarray) -> np.ndarray:
    """bool[i, j, k] -> int32[i, j, k // 32]; bit b of word (i, j, kp) is cell 32 kp + b."""
    i, j, k = occupancy.shape
    assert k % 32 == 0
    bits = occupancy.reshape(i, j, k // 32, 32).astype(np.uint64)
    powers = (np.uint64(1) << np.arange(32, dtype=np.uint64))
    return (bits * powers).sum(axis=-1).astype(np.uint32).view(np.int32)


def unpack_occupancy(packed: np.ndarray) -> np.ndarray:
    words = packed.view(np.uint32)
    powers = (np.uint32(1) << np.arange(32, dtype=np.uint32))
    bits = (words[..., None] & powers) != 0
    i, j, kp = packed.shape
    return bits.reshape(i, j, kp * 32)


# --------------------------------------------------------------------------------------
# Voxel structure (triangle_extraction/__init__.py:31-62, compute_vertex_occupancy.slang,
# count_occupancy.slang, create_voxels.slang)
# --------------------------------------------------------------------------------------


@dataclass
class Voxels:
    vertices: np.ndarray  # f32[V, 3]
    neighbors: np.ndarray  # i32[7, N_vox]
    lower_corners: np.ndarray  # i32[4, N_lc]
    upper_corners: np.ndarray  # i32[4, N_vox]
    indices: np.ndarray  # i32[N_vox]
    grid_shape: tuple[int, int, int]

    @property
    def num_voxels(self) -> int:
        return self.upper_corners.shape[1]


def extract_voxels(packed: np.ndarray) -> Voxels:
    """Rank formulation of create_voxels.slang.

    Every index in the structure is the rank of a lattice position in row-major
    (i, j, k) order: vertices are the positions (i, j, k), 0 <= i <= I, 0 <= j <= J,
    0 <= k <= K, that are a corner of an occupied cell (compute_vertex_occupancy.slang:
    36-61); voxels are the occupied cells; subvoxels are vertex positions that are not
    occupied cells, numbered after the voxels (create_voxels.slang:90-93). A column is
    a voxel or a subvoxel; lower corners are the column's (origin, +x, +y, +z) vertices
    (create_voxels.slang:104-123), upper corners a voxel's (+xyz, +yz, +xz, +xy)
    vertices (:125-131) and neighbors the columns at (+x, +y, +z, +xy, +yz, +xz, +xyz)
    (:134-179, 209-275). x is the packed bit axis k, y is j, z is i.
    """
    I, J, KP = packed.shape
    K = 32 * KP
    occ = unpack_occupancy(packed)
    vocc = np.zeros((I + 1, J + 1, K + 1), dtype=bool)
    for di, dj, dk in product((0, 1), repeat=3):
        vocc[di : di + I, dj : dj + J, dk : dk + K] |= occ
    cell = np.zeros((I + 1, J + 1, K + 1), dtype=bool)
    cell[:I, :J, :K] = occ

    vflat = vocc.reshape(-1)
    vidx = (np.cumsum(vflat, dtype=np.int64) - 1).reshape(vocc.shape)
    num_vertices = int(vflat.sum())
    cflat = cell.reshape(-1)
    num_voxels = int(cflat.sum())
    sub = vocc & ~cell
    col = np.full(vocc.shape, -1, dtype=np.int64)
    col[cell] = np.arange(num_voxels)
    col[sub] = num_voxels + np.arange(int(sub.sum()))

    # Vertex positions: (k / K, j / J, i / I) in float32 (create_voxels.slang:30-48).
    ii, jj, kk = np.nonzero(vocc)
    inv_i = F32(1) / F32(I)
    inv_j = F32(1) / F32(J)
    inv_k = F32(1) / F32(K)
    vertices = np.stack(
        [kk.astype(F32) * inv_k, jj.astype(F32) * inv_j, ii.astype(F32) * inv_i], axis=1
    ).astype(F32)

    # Pad by one on every high side so that +1 lookups never go out of range.
    vocc_p = np.zeros((I + 2, J + 2, K + 2), dtype=bool)
    vocc_p[: I + 1, : J + 1, : K + 1] = vocc
    vidx_p = np.full((I + 2, J + 2, K + 2), -1, dtype=np.int64)
    vidx_p[: I + 1, : J + 1, : K + 1] = vidx
    col_p = np.full((I + 2, J + 2, K + 2), -1, dtype=np.int64)
    col_p[: I + 1, : J + 1, : K + 1] = col

    # Lower corners for every column (voxels first, then subvoxels, in rank order).
    num_columns = num_vertices
    lower = np.full((4, num_columns), INVALID_CORNER, dtype=np.int64)
    ci, cj, ck = np.nonzero(vocc)
    cols = col[ci, cj, ck]
    is_vox = cell[ci, cj, ck]
    lower[0, cols] = vidx[ci, cj, ck]
    for a, (dk, dj, di) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        ni, nj, nk = ci + di, cj + dj, ck + dk
        valid = is_vox | vocc_p[ni, nj, nk]
        lower[a + 1, cols] = np.where(valid, vidx_p[ni, nj, nk], INVALID_CORNER)

    # Upper corners, neighbors and linear indices for voxels.
    vi, vj, vk = np.nonzero(occ)  # row-major == voxel rank order
    upper = np.stack(
        [
            vidx_p[vi + 1, vj + 1, vk + 1],  # +xyz
            vidx_p[vi + 1, vj + 1, vk],  # +yz
            vidx_p[vi + 1, vj, vk + 1],  # +xz
            vidx_p[vi, vj + 1, vk + 1],  # +xy
        ]
    )
    offs = ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1))
    neighbors = np.stack([col_p[vi + dz, vj + dy, vk + dx] for dx, dy, dz in offs])
    indices = K * (J * vi + vj) + vk

    return Voxels(
        vertices=vertices,
        neighbors=neighbors.astype(np.int32),
        lower_corners=lower.astype(np.int32),
        upper_corners=upper.astype(np.int32),
        indices=indices.astype(np.int32),
        grid_shape=(I, J, K),
    )


def extract_voxels_bitserial(packed: np.ndarray) -> Voxels:
    """Literal per-thread transcription of compute_vertex_occupancy.slang:9-68,
    count_occupancy.slang:3-20 and create_voxels.slang:3-276 (pure Python loops).

    Only for tiny grids: it exists to pin :func:`extract_voxels` against the
    reference's own bit-serial algorithm, including its last-bit special cases.
    """
    I, J, KP = packed.shape
    words = packed.view(np.uint32).astype(np.int64)

    def occ_word(i, j, kp):
        return int(words[i, j, kp])

    shape = (I + 1, J + 1, KP + 1)
    vocc = np.zeros(shape, dtype=np.int64)
    vcount = np.zeros(shape, dtype=np.int64)
    for i in range(I + 1):
        for j in range(J + 1):
            for kp in range(KP + 1):
                result = 0
                for di in range(2):
                    for dj in range(2):
                        if (i == 0 and di == 1) or (j == 0 and dj == 1):
                            continue
                        if (i == I and di == 0) or (j == J and dj == 0):
                            continue
                        cur = 0 if kp == KP else occ_word(i - di, j - dj, kp)
                        prev = 0 if kp == 0 else occ_word(i - di, j - dj, kp - 1)
                        previous = (prev >> 31) | ((cur << 1) & 0xFFFFFFFF)
                        result |= cur | previous
                vocc[i, j, kp] = result
                vcount[i, j, kp] = bin(result).count("1")
    voxcount = np.zeros(shape, dtype=np.int64)
    for i in range(I):
        for j in range(J):
            for kp in range(KP):
                voxcount[i, j, kp] = bin(occ_word(i, j, kp)).count("1")
    voff = np.concatenate([[0], np.cumsum(vcount.reshape(-1))])
    num_vertices = int(voff[-1])
    voff = voff[:-1].reshape(shape)
    xoff = np.concatenate([[0], np.cumsum(voxcount.reshape(-1))])
    num_voxels = int(xoff[-1])
    xoff = xoff[:-1].reshape(shape)

    vertices = np.zeros((num_vertices, 3), dtype=F32)
    neighbors = np.zeros((7, num_voxels), dtype=np.int64)
    lower = np.zeros((4, num_vertices), dtype=np.int64)
    upper = np.zeros((4, num_voxels), dtype=np.int64)
    indices = np.zeros((num_voxels,), dtype=np.int64)
    inv_i, inv_j, inv_k = F32(1) / F32(I), F32(1) / F32(J), F32(1) / F32(32 * KP)

    for i in range(I + 1):
        for j in range(J + 1):
            for kp in range(KP + 1):
                vi = int(voff[i, j, kp])
                chunk = int(vocc[i, j, kp])
                k = 32 * kp
                for bit in range(32):
                    if chunk & 1:
                        vertices[vi] = (F32(k) * inv_k, F32(j) * inv_j, F32(i) * inv_i)
                        vi += 1
                    chunk >>= 1
                    k += 1
                lvx = [[0, 0], [0, 0]]
                lvo = [[0, 0], [0, 0]]
                lti = [[0, 0], [0, 0]]
                lto = [[0, 0], [0, 0]]
                lsi = [[0, 0], [0, 0]]
                lso = [[0, 0], [0, 0]]
                for di in range(2):
                    for dj in range(2):
                        ii, jj = i + di, j + dj
                        if ii > I or jj > J:
                            lvx[di][dj] = lti[di][dj] = lsi[di][dj] = -64
                            lvo[di][dj] = lto[di][dj] = lso[di][dj] = 0
                            continue
                        valid = ii < I and jj < J and kp < KP
                        lvx[di][dj] = int(xoff[ii, jj, kp])
                        lvo[di][dj] = occ_word(ii, jj, kp) if valid else 0
                        lti[di][dj] = int(voff[ii, jj, kp])
                        lto[di][dj] = int(vocc[ii, jj, kp])
                        lsi[di][dj] = num_voxels + lti[di][dj] - lvx[di][dj]
                        lso[di][dj] = lto[di][dj] & ~lvo[di][dj]
                k_max = KP * 32
                k = 32 * kp
                for bit in range(32):
                    is_voxel = (lvo[0][0] & 1) != 0
                    is_sub = (lso[0][0] & 1) != 0
                    vox = lvx[0][0]
                    if is_voxel or is_sub:
                        chosen = vox if is_voxel else lsi[0][0]
                        lower[0, chosen] = lti[0][0]
                        xv = is_voxel or ((lto[0][0] & 2) != 0)
                        yv = is_voxel or ((lto[0][1] & 1) != 0)
                        zv = is_voxel or ((lto[1][0] & 1) != 0)
                        lower[1, chosen] = lti[0][0] + 1 if xv else -64
                        lower[2, chosen] = lti[0][1] if yv else -64
                        lower[3, chosen] = lti[1][0] if zv else -64
                    if is_voxel:
                        upper[0, vox] = lti[1][1] + 1
                        upper[1, vox] = lti[1][1]
                        upper[2, vox] = lti[1][0] + 1
                        upper[3, vox] = lti[0][1] + 1
                        indices[vox] = k_max * (J * i + j) + k
                    if is_voxel and bit != 31:
                        x_s = (lso[0][0] & 2) != 0
                        neighbors[0, vox] = lsi[0][0] if x_s else lvx[0][0] + 1
                        y_s = (lso[0][1] & 1) != 0
                        neighbors[1, vox] = lsi[0][1] if y_s else lvx[0][1]
                        z_s = (lso[1][0] & 1) != 0
                        neighbors[2, vox] = lsi[1][0] if z_s else lvx[1][0]
                        xy_s = (lso[0][1] & 2) != 0
                        neighbors[3, vox] = (
                            (lsi[0][1] + 1 if y_s else lsi[0][1])
                            if xy_s
                            else (lvx[0][1] if y_s else lvx[0][1] + 1)
                        )
                        yz_s = (lso[1][1] & 1) != 0
                        neighbors[4, vox] = lsi[1][1] if yz_s else lvx[1][1]
                        xz_s = (lso[1][0] & 2) != 0
                        neighbors[5, vox] = (
                            (lsi[1][0] + 1 if z_s else lsi[1][0])
                            if xz_s
                            else (lvx[1][0] if z_s else lvx[1][0] + 1)
                        )
                        xyz_s = (lso[1][1] & 2) != 0
                        neighbors[6, vox] = (
                            (lsi[1][1] + 1 if yz_s else lsi[1][1])
                            if xyz_s
                            else (lvx[1][1] if yz_s else lvx[1][1] + 1)
                        )
                    if bit != 31:
                        for di in range(2):
                            for dj in range(2):
                                lvx[di][dj] += lvo[di][dj] & 1
                                lvo[di][dj] >>= 1
                                lti[di][dj] += lto[di][dj] & 1
                                lto[di][dj] >>= 1
                                lsi[di][dj] += lso[di][dj] & 1
                                lso[di][dj] >>= 1
                    k += 1
                if (lso[0][0] & 1) != 0:
                    xv = kp < KP and (int(vocc[i, j, kp + 1]) & 1) != 0
                    lower[1, lsi[0][0]] = lti[0][0] + 1 if xv else -64
                if (lvo[0][0] & 1) == 0:
                    continue
                vox = lvx[0][0]
                cx, cy, cz = kp < KP - 1, j < J - 1, i < I - 1
                bit31 = 1 << 31
                x_s = ((occ_word(i, j, kp + 1) & 1) == 0) if cx else True
                neighbors[0, vox] = lsi[0][0] if x_s else lvx[0][0] + 1
                y_s = ((occ_word(i, j + 1, kp) & bit31) == 0) if cy else True
                neighbors[1, vox] = lsi[0][1] if y_s else lvx[0][1]
                z_s = ((occ_word(i + 1, j, kp) & bit31) == 0) if cz else True
                neighbors[2, vox] = lsi[1][0] if z_s else lvx[1][0]
                xy_s = ((occ_word(i, j + 1, kp + 1) & 1) == 0) if (cx and cy) else True
                neighbors[3, vox] = (
                    (lsi[0][1] + 1 if y_s else lsi[0][1])
                    if xy_s
                    else (lvx[0][1] if y_s else lvx[0][1] + 1)
                )
                yz_s = ((occ_word(i + 1, j + 1, kp) & bit31) == 0) if (cy and cz) else True
                neighbors[4, vox] = lsi[1][1] if yz_s else lvx[1][1]
                xz_s = ((occ_word(i + 1, j, kp + 1) & 1) == 0) if (cx and cz) else True
                neighbors[5, vox] = (
                    (lsi[1][0] + 1 if z_s else lsi[1][0])
                    if xz_s
                    else (lvx[1][0] if z_s else lvx[1][0] + 1)
                )
                xyz_s = (
                    ((occ_word(i + 1, j + 1, kp + 1) & 1) == 0) if (cx and cy and cz) else True
                )
                neighbors[6, vox] = (
                    (lsi[1][1] + 1 if yz_s else lsi[1][1])
                    if xyz_s
                    else (lvx[1][1] if yz_s else lvx[1][1] + 1)
                )

    return Voxels(
        vertices=vertices,
        neighbors=neighbors.astype(np.int32),
        lower_corners=lower.astype(np.int32),
        upper_corners=upper.astype(np.int32),
        indices=indices.astype(np.int32),
        grid_shape=(I, J, KP * 32),
    )


# --------------------------------------------------------------------------------------
# Scan (compute_exclusive_cumsum.cu:8-39)
# --------------------------------------------------------------------------------------


def exclusive_cumsum(x: np.ndarray) -> np.ndarray:
    """In-place-style exclusive prefix sum with int32 wrap-around, like CUB on int32."""
    out = np.zeros_like(x)
    if x.size > 1:
        out[1:] = np.cumsum(x[:-1].astype(np.int64)).astype(x.dtype)
    return out


# --------------------------------------------------------------------------------------
# Marching cubes (count_triangle_vertices / create_triangle_vertices / count_triangle_faces
# / create_triangle_faces)
# --------------------------------------------------------------------------------------


def _column_sdfs(sdf: np.ndarray, lower: np.ndarray):
    d0 = sdf[lower[0]]
    d = []
    for a in range(3):
        idx = lower[a + 1]
        valid = idx >= 0
        d.append(np.where(valid, sdf[np.where(valid, idx, 0)], d0))
    return d0, d


def count_triangle_vertices(sdf, lower, level_sets) -> np.ndarray:
    """count_triangle_vertices.slang:3-37. Returns int32[S * N_lc + 1], last entry 0."""
    n = lower.shape[1]
    s_count = level_sets.shape[0]
    d0, d = _column_sdfs(sdf, lower)
    out = np.zeros(s_count * n + 1, dtype=np.int32)
    for s in range(s_count):
        iso = level_sets[s]
        below0 = d0 < iso
        cnt = np.zeros(n, dtype=np.int32)
        for a in range(3):
            cnt += (below0 != (d[a] < iso)).astype(np.int32)
        out[s * n : (s + 1) * n] = cnt
    return out


@dataclass
class TriangleVertexRecord:
    """Where each triangle vertex came from: needed by the adjoint oracle."""

    column: np.ndarray  # int64[V_t]
    level_set: np.ndarray  # int64[V_t]
    axis: np.ndarray  # int64[V_t]
    i0: np.ndarray  # int64[V_t] grid vertex of the column origin
    i1: np.ndarray  # int64[V_t] grid vertex at the axis end
    t: np.ndarray  # f32[V_t]


def create_triangle_vertices(grid_vertices, sdf, sh, lower, level_sets, vertex_offsets):
    """create_triangle_vertices.slang:145-199 (forward).

    t = clamp((iso - d0) / (d1 - d0), 0, 1) (0.5 if d0 == d1, unreachable on a
    crossing), v = v0 + t (v1 - v0), sh = sh0 + t (sh1 - sh0), d = d0 + t (d1 - d0),
    type = axis. Crossing rule is count_triangle_vertices.slang:26-34 so the vertex
    slots always match the counted ones.
    """
    n = lower.shape[1]
    s_count = level_sets.shape[0]
    nsh = sh.shape[0]
    total = int(vertex_offsets[-1])
    d0, d = _column_sdfs(sdf, lower)
    out_v = np.empty((total, 3), dtype=F32)
    out_d = np.empty((total,), dtype=F32)
    out_sh = np.empty((nsh, total, 3), dtype=F32)
    out_type = np.empty((total,), dtype=np.uint8)
    rec_cols, rec_ls, rec_ax, rec_i0, rec_i1, rec_t = [], [], [], [], [], []
    cols = np.arange(n)
    i0 = lower[0].astype(np.int64)
    for s in range(s_count):
        iso = F32(level_sets[s])
        base = vertex_offsets[s * n : (s + 1) * n].astype(np.int64)
        rank = np.zeros(n, dtype=np.int64)
        below0 = d0 < iso
        for a in range(3):
            cross = below0 != (d[a] < iso)
            c = cols[cross]
            idx = base[cross] + rank[cross]
            rank += cross
            da0 = d0[cross]
            da1 = d[a][cross]
            with np.errstate(divide="ignore", invalid="ignore"):
                t = (iso - da0) / (da1 - da0)
            t = np.fmin(np.fmax(t, F32(0)), F32(1)).astype(F32)
            t = np.where(da0 == da1, F32(0.5), t).astype(F32)
            g0 = i0[cross]
            g1 = lower[a + 1][cross].astype(np.int64)
            v0 = grid_vertices[g0]
            v1 = grid_vertices[g1]
            out_v[idx] = v0 + t[:, None] * (v1 - v0)
            out_d[idx] = da0 + t * (da1 - da0)
            sh0 = sh[:, g0]
            sh1 = sh[:, g1]
            out_sh[:, idx] = sh0 + t[None, :, None] * (sh1 - sh0)
            out_type[idx] = a
            rec_cols.append(c)
            rec_ls.append(np.full(c.shape, s))
            rec_ax.append(np.full(c.shape, a))
            rec_i0.append(g0)
            rec_i1.append(g1)
            rec_t.append(t)
    # Reassemble the record in triangle-vertex order.
    if rec_cols:
        cols_all = np.concatenate(rec_cols)
        ls_all = np.concatenate(rec_ls)
        ax_all = np.concatenate(rec_ax)
        rank_key = ls_all * (3 * n) + cols_all * 3 + ax_all
        order = np.argsort(rank_key, kind="stable")
        record = TriangleVertexRecord(
            column=cols_all[order],
            level_set=ls_all[order],
            axis=ax_all[order],
            i0=np.concatenate(rec_i0)[order],
            i1=np.concatenate(rec_i1)[order],
            t=np.concatenate(rec_t)[order],
        )
    else:
        e = np.zeros(0, dtype=np.int64)
        record = TriangleVertexRecord(e, e, e, e, e, np.zeros(0, F32))
    return out_v, out_d, out_sh, out_type, record


def create_triangle_vertices_backward(
    record: TriangleVertexRecord,
    num_grid_vertices: int,
    nsh: int,
    grad_v=None,
    grad_d=None,
    grad_sh=None,
    dtype=np.float64,
):
    """Adjoint of the forward above (create_triangle_vertices.slang:201-291):
    d/d(x0) = (1 - t) g and d/d(x1) = t g for x in {vertex, sdf, sh}; t is detached
    (create_triangle_vertices.slang:92-96) so no term flows through t."""
    t = record.t.astype(dtype)
    w0 = 1 - t
    gv = np.zeros((num_grid_vertices, 3), dtype=dtype)
    gd = np.zeros((num_grid_vertices,), dtype=dtype)
    gs = np.zeros((nsh, num_grid_vertices, 3), dtype=dtype)
    if grad_v is not None:
        g = grad_v.astype(dtype)
        np.add.at(gv, record.i0, w0[:, None] * g)
        np.add.at(gv, record.i1, t[:, None] * g)
    if grad_d is not None:
        g = grad_d.astype(dtype)
        np.add.at(gd, record.i0, w0 * g)
        np.add.at(gd, record.i1, t * g)
    if grad_sh is not None:
        g = grad_sh.astype(dtype)
        for c in range(nsh):
            np.add.at(gs[c], record.i0, w0[:, None] * g[c])
            np.add.at(gs[c], record.i1, t[:, None] * g[c])
    return gv, gd, gs


def cube_sdfs(sdf, lower, upper):
    """Cube corner SDFs in count_triangle_faces.slang:29-38 order."""
    n = upper.shape[1]
    return np.stack(
        [
            sdf[lower[0, :n]],
            sdf[lower[1, :n]],
            sdf[upper[3]],
            sdf[lower[2, :n]],
            sdf[lower[3, :n]],
            sdf[upper[2]],
            sdf[upper[0]],
            sdf[upper[1]],
        ]
    )


def count_triangle_faces(sdf, lower, upper, level_sets):
    """count_triangle_faces.slang:13-50. Returns (face_counts i32[S N_vox + 1], codes u8[S, N_vox]),
    in the positional order the reference returns them (count_triangle_faces.py:63)."""
    case_start, _, _ = tables.load()
    n = upper.shape[1]
    s_count = level_sets.shape[0]
    cs = cube_sdfs(sdf, lower, upper)
    counts = np.zeros(s_count * n + 1, dtype=np.int32)
    codes = np.zeros((s_count, n), dtype=np.uint8)
    for s in range(s_count):
        iso = level_sets[s]
        code = np.zeros(n, dtype=np.int64)
        for c in range(8):
            code |= (cs[c] >= iso).astype(np.int64) << c
        codes[s] = code.astype(np.uint8)
        counts[s * n : (s + 1) * n] = (case_start[code + 1] - case_start[code]).astype(np.int32)
    return counts, codes


def create_triangle_faces(neighbors, indices, vertex_offsets, vertex_types, codes, face_offsets):
    """create_triangle_faces.slang:3-78: walk the case table, find each edge's owner column
    and, among the owner's vertices of this level set, the first one whose type equals the
    edge axis (-1 if none). Winding is reversed: entry j goes to slot 2 - (base + j) % 3."""
    case_start, case_edges, edge_owner = tables.load()
    s_count, n = codes.shape
    n_lc = (vertex_offsets.shape[0] - 1) // s_count
    num_faces = int(face_offsets[-1]) // 3
    faces = np.empty((num_faces, 3), dtype=np.int32)
    voxel_indices = np.empty((num_faces,), dtype=np.int32)
    vo = vertex_offsets.astype(np.int64)
    vt = np.concatenate([vertex_types.astype(np.int64), [255]])
    vox = np.arange(n)
    for s in range(s_count):
        code = codes[s].astype(np.int64)
        base = face_offsets[s * n : (s + 1) * n].astype(np.int64)
        num = case_start[code + 1] - case_start[code]
        for j in range(15):
            m = num > j
            if not m.any():
                break
            v = vox[m]
            e = case_edges[case_start[code[m]] + j]
            nb = edge_owner[e, 0]
            axis = edge_owner[e, 1]
            owner = np.where(nb < 0, v, neighbors[np.maximum(nb, 0), v])
            slot = s * n_lc + owner
            start = vo[slot]
            cnt = vo[slot + 1] - start
            target = np.full(v.shape, -1, dtype=np.int64)
            for kk in range(2, -1, -1):
                has = cnt > kk
                q = np.where(has, start + kk, vt.shape[0] - 1)
                hit = has & (vt[q] == axis)
                target = np.where(hit, start + kk, target)
            pos = base[m] + j
            tri = pos // 3
            off = 2 - pos % 3
            faces[tri, off] = target
            first = off == 0
            voxel_indices[tri[first]] = indices[v[first]]
    return faces, voxel_indices


@dataclass
class Triangles:
    vertices: np.ndarray
    signed_distances: np.ndarray
    spherical_harmonics: np.ndarray
    faces: np.ndarray
    voxel_indices: np.ndarray
    vertex_boundaries: np.ndarray
    face_boundaries: np.ndarray
    # Extra oracle-only outputs.
    vertex_offsets: np.ndarray = None
    face_offsets: np.ndarray = None
    vertex_counts: np.ndarray = None
    face_counts: np.ndarray = None
    cell_codes: np.ndarray = None
    vertex_types: np.ndarray = None
    record: TriangleVertexRecord = None


def tessellate_voxels(sdf, grid_vertices, sh, neighbors, lower, upper, indices, level_sets):
    """triangle_extraction/__init__.py:75-157."""
    s_count = level_sets.shape[0]
    vcounts = count_triangle_vertices(sdf, lower, level_sets)
    voffsets = exclusive_cumsum(vcounts)
    tv, td, tsh, ttype, record = create_triangle_vertices(
        grid_vertices, sdf, sh, lower, level_sets, voffsets
    )
    fcounts, codes = count_triangle_faces(sdf, lower, upper, level_sets)
    foffsets = exclusive_cumsum(fcounts)
    faces, vox_idx = create_triangle_faces(neighbors, indices, voffsets, ttype, codes, foffsets)
    vb = voffsets[1:].reshape(s_count, -1)[:, -1].copy()
    fb = foffsets[1:].reshape(s_count, -1)[:, -1].copy() // 3
    return Triangles(
        tv, td, tsh, faces, vox_idx, vb, fb,
        vertex_offsets=voffsets, face_offsets=foffsets, vertex_counts=vcounts,
        face_counts=fcounts, cell_codes=codes, vertex_types=ttype, record=record,
    )
