"""PyTorch-CPU transcription of the extract -> rasterize -> composite training step, for the
CPU baseline. Test infrastructure only (see oracle/__init__.py): bench.py's cpu_baseline
times it and tests/test_cpu_torch.py checks it against the NumPy oracle.

It follows the oracle's algorithm and FP32 operation order (oracle/extraction.py,
oracle/rasterization.py, which cite the reference lines), but every stage works on whole
arrays, so torch's intra-op thread pool spreads it over the host cores it is given
(torch.set_num_threads):

* extraction, marching-cubes counts / vertices / faces: vectorised over cells and columns,
  one pass per (level set, axis) or (level set, case-table slot);
* projection, tile bounds, keys (the reference's unculled tile boxes), a stable sort and
  the tile runs: vectorised over faces and keys;
* the fused traversal and composite, forward and backward, on dense (tile, pixel, key)
  blocks of tiles with similar key counts. A key is accepted when its pixel passes the
  barycentric test and its shell exceeds every shell accepted before it; that equals
  exceeding the running maximum over all earlier hits (a rejected hit's shell never
  exceeds it), so acceptance is one cummax along the key axis. A pixel accepts at most
  one key per shell, so its accepted keys are packed in order into S slots; transmittance
  is a cumprod and the colour a cumsum over the slots, both in the reference's sequential
  order; the adjoint is the oracle's suffix form in FP32;
* the SH adjoint and the vertex adjoint to the grid as index_add over vertices.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch
from torch import Tensor

from . import tables
from .rasterization import (ALPHA_EPS, DUMMY_OUTER_SDF, EPS_BARY, SH_C0, SH_C1, SH_C2, SH_C3,
                            SUFFIX_EPS, SUM_HI, SUM_LO)

F32 = torch.float32
I64 = torch.int64
INVALID_CORNER = -64


# ------------------------------------------------------------------------------------
# Extraction
# ------------------------------------------------------------------------------------

@dataclass
class Voxels:
    vertices: Tensor  # f32[V, 3]
    neighbors: Tensor  # i64[7, N_vox]
    lower: Tensor  # i64[4, V] (columns: voxels, then subvoxels)
    upper: Tensor  # i64[4, N_vox]
    indices: Tensor  # i64[N_vox]


def unpack(packed: Tensor) -> Tensor:
    bits = torch.arange(32, dtype=torch.int32)
    i, j, kp = packed.shape
    return ((packed[..., None] >> bits) & 1).bool().reshape(i, j, kp * 32)


def extract_voxels(packed: Tensor, offset_i: int = 0, full_i: int | None = None) -> Voxels:
    """Voxels of a packed grid. A slab of a larger grid (layers offset_i.. of full_i along
    i) is extracted as a grid of its own, its vertices placed in the larger grid's unit
    coordinates."""
    occ = unpack(packed)
    I, J, K = occ.shape
    vocc = torch.zeros((I + 1, J + 1, K + 1), dtype=torch.bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                vocc[di : di + I, dj : dj + J, dk : dk + K] |= occ
    cell = torch.zeros_like(vocc)
    cell[:I, :J, :K] = occ
    vidx = (torch.cumsum(vocc.reshape(-1).to(I64), 0) - 1).reshape(vocc.shape)
    nvox = int(cell.sum())
    sub = vocc & ~cell
    col = torch.full(vocc.shape, -1, dtype=I64)
    col[cell] = torch.arange(nvox)
    col[sub] = nvox + torch.arange(int(sub.sum()))
    ci, cj, ck = torch.nonzero(vocc, as_tuple=True)
    inv = [torch.tensor(1.0, dtype=F32) / torch.tensor(float(n), dtype=F32)
           for n in (K, J, I if full_i is None else full_i)]
    vertices = torch.stack([ck.to(F32) * inv[0], cj.to(F32) * inv[1],
                            (ci + offset_i).to(F32) * inv[2]], 1)

    def padded(t, fill):
        out = torch.full((I + 2, J + 2, K + 2), fill, dtype=t.dtype)
        out[: I + 1, : J + 1, : K + 1] = t
        return out

    vocc_p, vidx_p, col_p = padded(vocc, False), padded(vidx, -1), padded(col, -1)
    n = ci.numel()
    cols = col[ci, cj, ck]
    is_vox = cell[ci, cj, ck]
    lower = torch.full((4, n), INVALID_CORNER, dtype=I64)
    lower[0, cols] = vidx[ci, cj, ck]
    for a, (dk, dj, di) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        ni, nj, nk = ci + di, cj + dj, ck + dk
        ok = is_vox | vocc_p[ni, nj, nk]
        lower[a + 1, cols] = torch.where(ok, vidx_p[ni, nj, nk], INVALID_CORNER)
    vi, vj, vk = torch.nonzero(occ, as_tuple=True)
    upper = torch.stack([vidx_p[vi + 1, vj + 1, vk + 1], vidx_p[vi + 1, vj + 1, vk],
                         vidx_p[vi + 1, vj, vk + 1], vidx_p[vi, vj + 1, vk + 1]])
    offs = ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1))
    neighbors = torch.stack([col_p[vi + dz, vj + dy, vk + dx] for dx, dy, dz in offs])
    return Voxels(vertices, neighbors, lower, upper, K * (J * vi + vj) + vk)


# ------------------------------------------------------------------------------------
# Marching cubes over level sets
# ------------------------------------------------------------------------------------

@dataclass
class Tessellation:
    vertices: Tensor  # f32[V_t, 3]
    sdf: Tensor  # f32[V_t]
    sh: Tensor  # f32[NSH, V_t, 3]
    faces: Tensor  # i64[F, 3]
    face_boundaries: Tensor  # i64[S]
    i0: Tensor  # i64[V_t]: grid vertex at the column origin
    i1: Tensor  # i64[V_t]: grid vertex at the axis end
    t: Tensor  # f32[V_t]


def _column_sdfs(sdf, lower):
    d0 = sdf[lower[0]]
    d = [torch.where(lower[a + 1] >= 0, sdf[lower[a + 1].clamp(min=0)], d0) for a in range(3)]
    return d0, d


def _exclusive(counts: Tensor) -> Tensor:
    out = torch.zeros(counts.numel() + 1, dtype=I64)
    torch.cumsum(counts.reshape(-1).to(I64), 0, out=out[1:])
    return out


def tessellate(sdf: Tensor, grid_vertices: Tensor, sh: Tensor, vox: Voxels, level_sets: Tensor
               ) -> Tessellation:
    lower, upper = vox.lower, vox.upper
    n = lower.shape[1]
    S = level_sets.shape[0]
    d0, d = _column_sdfs(sdf, lower)
    below0 = [d0 < level_sets[s] for s in range(S)]
    cross = [[below0[s] != (d[a] < level_sets[s]) for a in range(3)] for s in range(S)]
    counts = torch.stack([cross[s][0].int() + cross[s][1].int() + cross[s][2].int() for s in range(S)])
    voff = _exclusive(counts)  # [S n + 1]
    total = int(voff[-1])
    nsh = sh.shape[0]
    tv = torch.empty((total, 3), dtype=F32)
    td = torch.empty(total, dtype=F32)
    tsh = torch.empty((nsh, total, 3), dtype=F32)
    ttype = torch.empty(total, dtype=I64)
    ti0 = torch.empty(total, dtype=I64)
    ti1 = torch.empty(total, dtype=I64)
    tt = torch.empty(total, dtype=F32)
    for s in range(S):
        iso = level_sets[s]
        base = voff[s * n : (s + 1) * n]
        rank = torch.zeros(n, dtype=I64)
        for a in range(3):
            m = cross[s][a]
            c = torch.nonzero(m, as_tuple=True)[0]
            idx = base[c] + rank[c]
            rank += m
            da0, da1 = d0[c], d[a][c]
            t = ((iso - da0) / (da1 - da0)).clamp(0, 1)
            t = torch.where(da0 == da1, torch.tensor(0.5, dtype=F32), t)
            g0, g1 = lower[0, c], lower[a + 1, c]
            v0, v1 = grid_vertices[g0], grid_vertices[g1]
            tv[idx] = v0 + t[:, None] * (v1 - v0)
            td[idx] = da0 + t * (da1 - da0)
            sh0, sh1 = sh[:, g0], sh[:, g1]
            tsh[:, idx] = sh0 + t[None, :, None] * (sh1 - sh0)
            ttype[idx] = a
            ti0[idx], ti1[idx], tt[idx] = g0, g1, t
    # Faces.
    case_start, case_edges, edge_owner = (torch.from_numpy(x).to(I64) for x in tables.load())
    nv = upper.shape[1]
    corners = torch.stack([sdf[lower[0, :nv]], sdf[lower[1, :nv]], sdf[upper[3]], sdf[lower[2, :nv]],
                           sdf[lower[3, :nv]], sdf[upper[2]], sdf[upper[0]], sdf[upper[1]]])
    weights = (1 << torch.arange(8, dtype=I64))[:, None]
    codes = torch.stack([((corners >= level_sets[s]).to(I64) * weights).sum(0) for s in range(S)])
    fcounts = case_start[codes + 1] - case_start[codes]  # [S, nv], 3 x triangles
    foff = _exclusive(fcounts)
    nf = int(foff[-1]) // 3
    faces = torch.empty((nf, 3), dtype=I64)
    vt = torch.cat([ttype, torch.tensor([255])])
    vox_ids = torch.arange(nv)
    for s in range(S):
        code = codes[s]
        num = fcounts[s]
        fbase = foff[s * nv : (s + 1) * nv]
        for j in range(15):
            v = torch.nonzero(num > j, as_tuple=True)[0]
            if v.numel() == 0:
                break
            e = case_edges[case_start[code[v]] + j]
            nb, axis = edge_owner[e, 0], edge_owner[e, 1]
            owner = torch.where(nb < 0, v, vox.neighbors[nb.clamp(min=0), v])
            slot = s * n + owner
            start = voff[slot]
            cnt = voff[slot + 1] - start
            target = torch.full(v.shape, -1, dtype=I64)
            for kk in (2, 1, 0):
                has = cnt > kk
                q = torch.where(has, start + kk, torch.full_like(start, vt.numel() - 1))
                target = torch.where(has & (vt[q] == axis), start + kk, target)
            pos = fbase[v] + j
            faces[pos // 3, 2 - pos % 3] = target
    fb = foff[1:].reshape(S, -1)[:, -1] // 3
    return Tessellation(tv, td, tsh, faces, fb, ti0, ti1, tt)


# ------------------------------------------------------------------------------------
# Front end: projection, tile bounds, keys, sort, tile runs
# ------------------------------------------------------------------------------------

def project(vertices: Tensor, E: Tensor, K: Tensor):
    x, y, z = vertices[:, 0], vertices[:, 1], vertices[:, 2]
    xy, depth = [], []
    for c in range(E.shape[0]):
        e, k = E[c], K[c]
        cs = [((e[r, 0] * x + e[r, 1] * y) + e[r, 2] * z) + e[r, 3] for r in range(3)]
        xy.append(torch.stack([((k[r, 0] * cs[0] + k[r, 1] * cs[1]) + k[r, 2] * cs[2]) / cs[2]
                               for r in range(2)], -1))
        depth.append(cs[2])
    return torch.stack(xy), torch.stack(depth)


def _f2i(x: Tensor) -> Tensor:
    x = torch.nan_to_num(x.double(), nan=0.0, posinf=2.0**31 - 1, neginf=-(2.0**31))
    return torch.trunc(x).clamp(-(2.0**31), 2.0**31 - 1).to(I64)


def _c_div(a: Tensor, b: int) -> Tensor:
    return torch.div(a, b, rounding_mode="trunc")


def front_end(xy, depth, faces, face_boundaries, tile_shape, image_shape, near_plane):
    """Sorted (key, face) pairs of every camera's unculled tile boxes and the [start, end)
    of every (camera, tile) run [B, T, 2] (fused key layout: camera 6 | tile 20 | depth 32
    | shell 6 bits)."""
    th, tw = tile_shape
    h, w = image_shape
    rows, cols = -(-h // th), -(-w // tw)
    B = xy.shape[0]
    f = faces.shape[0]
    fa, fb, fc = faces[:, 0], faces[:, 1], faces[:, 2]
    shell = torch.searchsorted(face_boundaries, torch.arange(f), right=True)
    keys, vals = [], []
    for c in range(B):
        ax, ay, ad = xy[c, fa, 0], xy[c, fa, 1], depth[c, fa]
        bx, by, bd = xy[c, fb, 0], xy[c, fb, 1], depth[c, fb]
        cx, cy, cd = xy[c, fc, 0], xy[c, fc, 1], depth[c, fc]
        check = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
        culled = (check < 0) | (torch.fmin(torch.fmin(ad, bd), cd) < near_plane)

        def bounds(lo, hi, ts, gmax):
            t0 = _c_div(_f2i(lo), ts)
            t1 = _f2i(torch.ceil(hi / ts))
            return t0.clamp(0, gmax), t1.clamp(0, gmax)

        r0, r1 = bounds(torch.fmin(torch.fmin(ay, by), cy), torch.fmax(torch.fmax(ay, by), cy), th, rows)
        c0, c1 = bounds(torch.fmin(torch.fmin(ax, bx), cx), torch.fmax(torch.fmax(ax, bx), cx), tw, cols)
        width = torch.where(culled, 0, c1 - c0)
        count = torch.where(culled, 0, (r1 - r0) * (c1 - c0))
        face = torch.repeat_interleave(torch.arange(f), count)
        start = torch.repeat_interleave(torch.cumsum(count, 0) - count, count)
        local = torch.arange(face.numel()) - start
        wid = width[face]
        tile = (r0[face] + local // wid) * cols + (c0[face] + local % wid)
        mean = ((ad + bd) + cd) / 3.0
        dbits = mean.view(torch.int32).to(I64) & 0xFFFFFFFF
        keys.append((c << 58) | (tile << 38) | (dbits[face] << 6) | shell[face])
        vals.append(face)
    keys, vals = torch.cat(keys), torch.cat(vals)
    sk, order = torch.sort(keys, stable=True)
    sv = vals[order]
    run = sk >> 38  # camera and tile
    change = torch.ones(sk.numel(), dtype=torch.bool)
    change[1:] = run[1:] != run[:-1]
    starts = torch.nonzero(change, as_tuple=True)[0]
    ends = torch.cat([starts[1:], torch.tensor([sk.numel()])])
    ranges = torch.zeros((B, rows * cols, 2), dtype=I64)
    cam, tile = run[starts] >> 20, run[starts] & ((1 << 20) - 1)
    ranges[cam, tile, 0] = starts
    ranges[cam, tile, 1] = ends
    return sk, sv, ranges


# ------------------------------------------------------------------------------------
# Spherical harmonics
# ------------------------------------------------------------------------------------

def _basis(d: Tensor, deg: int) -> Tensor:
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    out = [torch.full_like(x, SH_C0)]
    if deg > 0:
        out += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if deg > 1:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        out += [SH_C2[0] * xy, SH_C2[1] * yz, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * xz,
                SH_C2[4] * (xx - yy)]
    if deg > 2:
        out += [SH_C3[0] * y * (3 * xx - yy), SH_C3[1] * xy * z, SH_C3[2] * y * (4 * zz - xx - yy),
                SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy), SH_C3[4] * x * (4 * zz - xx - yy),
                SH_C3[5] * z * (xx - yy), SH_C3[6] * x * (xx - 3 * yy)]
    return torch.stack(out, -1)


def _directions(vertices, E):
    cams = torch.linalg.inv(E.double())[:, :3, 3].to(F32)
    d = vertices[None] - cams[:, None]
    return d / torch.linalg.vector_norm(d, dim=-1, keepdim=True)


def sh_forward(vertices, sh, E, deg):
    """colours [B, V, 3] = max(sum_k basis_k sh_k + 0.5, 0), and the pre-activation."""
    nsh = (deg + 1) ** 2
    basis = _basis(_directions(vertices, E), deg)  # [B, V, k]
    pre = torch.einsum("bvk,kvc->bvc", basis, sh[:nsh]) + 0.5
    return pre.clamp(min=0), pre


def sh_backward(vertices, sh, E, deg, pre, grad_colors):
    nsh = (deg + 1) ** 2
    basis = _basis(_directions(vertices, E), deg)
    g = grad_colors * (pre > 0)
    out = torch.zeros_like(sh)
    out[:nsh] = torch.einsum("bvk,bvc->kvc", basis, g)
    return out


# ------------------------------------------------------------------------------------
# Fused traversal and composite on dense tile blocks
# ------------------------------------------------------------------------------------

def _barycentrics(p1x, p1y, p2x, p2y, p3x, p3y, px, py):
    e12x, e12y = p2x - p1x, p2y - p1y
    e23x, e23y = p3x - p2x, p3y - p2y
    e13x, e13y = p3x - p1x, p3y - p1y
    d_12_12 = e12x * e12x + e12y * e12y
    d_12_23 = e12x * e23x + e12y * e23y
    d_12_13 = e12x * e13x + e12y * e13y
    d_13_13 = e13x * e13x + e13y * e13y
    d_13_23 = e13x * e23x + e13y * e23y
    d_23_23 = e23x * e23x + e23y * e23y
    a1x, a1y = px - p1x, py - p1y
    a2x, a2y = px - p2x, py - p2y
    d_1x_12 = a1x * e12x + a1y * e12y
    d_1x_13 = a1x * e13x + a1y * e13y
    d_2x_12 = a2x * e12x + a2y * e12y
    d_2x_13 = a2x * e13x + a2y * e13y
    a, b, c = d_12_13.abs(), d_12_23.abs(), d_13_23.abs()
    den1 = d_12_12 * d_13_13 - d_12_13 * d_12_13
    den2 = d_12_12 * d_23_23 - d_12_23 * d_12_23
    den3 = d_13_13 * d_23_23 - d_13_23 * d_13_23
    nan = torch.tensor(float("nan"), dtype=F32)
    denom = torch.where((a <= b) & (a <= c), den1,
                        torch.where((b <= a) & (b <= c), den2,
                                    torch.where((c <= a) & (c <= b), den3, nan)))
    dd = denom + float(EPS_BARY)
    v = (d_13_13 * d_1x_12 - d_12_13 * d_1x_13) / dd
    w = (d_12_12 * d_1x_13 - d_12_13 * d_1x_12) / dd
    u = (d_12_23 * d_2x_13 - d_13_23 * d_2x_12) / dd
    s = (u + v) + w
    hit = ((u >= 0) & (u <= 1) & (v >= 0) & (v <= 1) & (w >= 0) & (w <= 1)
           & (s >= float(SUM_LO)) & (s <= float(SUM_HI)))  # NaN compares false: a miss
    return u, v, w, hit


def _cdf(s, k):
    return 1 / (1 + torch.exp(-s * k))


@dataclass
class CompositeResult:
    image: Tensor  # f32[B, 4, H, W] (rgb, transmittance) of the composited tiles
    colors_grad: Tensor | None  # f32[B, V, 3]
    sdf_grad: Tensor | None  # f32[V]
    sharpness_grad: float
    loss: float
    pairs: int  # (pixel, key) pairs evaluated


def fused_composite(xy, faces, sdf, colors, sorted_keys, sorted_values, ranges, tile_shape,
                    image_shape, sharpness, background=None, target=None, tiles=None,
                    block_elems: int = 1 << 18) -> CompositeResult:
    """Forward of the fused composite and, with ``target`` f32[B, 3, H, W], the loss
    mean((clip(rgb + T bg, 0, 1) - target)^2) (the mean over all of target's entries; the
    composited tiles' share of it) and its adjoint to colours, SDF and sharpness.
    ``tiles`` (i64[n, 2] of (camera, tile)) restricts the work to those tiles."""
    th, tw = tile_shape
    h, w = image_shape
    B = ranges.shape[0]
    cols = -(-w // tw)
    k = float(sharpness)
    nv = sdf.shape[0]
    image = torch.zeros((B, 4, h, w), dtype=F32)
    image[:, 3] = 1
    backward = target is not None
    cg = torch.zeros((B * nv, 3), dtype=F32) if backward else None
    sg = torch.zeros(nv, dtype=F32) if backward else None
    kg = torch.zeros((), dtype=torch.float64)
    loss = torch.zeros((), dtype=torch.float64)
    whole = tiles is None
    written = torch.zeros((B, h, w), dtype=torch.bool)
    if whole:
        cam, tile = torch.meshgrid(torch.arange(B), torch.arange(ranges.shape[1]), indexing="ij")
        tiles = torch.stack([cam.reshape(-1), tile.reshape(-1)], 1)
    cnt = ranges[tiles[:, 0], tiles[:, 1], 1] - ranges[tiles[:, 0], tiles[:, 1], 0]
    tiles, cnt = tiles[cnt > 0], cnt[cnt > 0]
    order = torch.argsort(cnt, descending=True)
    tiles, cnt = tiles[order], cnt[order]
    n_el = float(3 * B * h * w) if target is None else float(target.numel())  # mean over all
    iy, ix = torch.meshgrid(torch.arange(th), torch.arange(tw), indexing="ij")
    iy, ix = iy.reshape(-1), ix.reshape(-1)
    P = th * tw
    shell_all = sorted_keys & 63
    n_slots = int(shell_all.max()) + 1 if shell_all.numel() else 1
    colors_flat = colors.reshape(-1, 3)
    pairs = 0
    i = 0
    while i < tiles.shape[0]:
        kmax = int(cnt[i])
        g = max(1, min(tiles.shape[0] - i, block_elems // (P * kmax)))
        blk = tiles[i : i + g]
        i += g
        cam, tl = blk[:, 0], blk[:, 1]
        start = ranges[cam, tl, 0]
        n_k = ranges[cam, tl, 1] - start
        kk = torch.arange(kmax)
        kvalid = kk[None] < n_k[:, None]  # [G, K]
        pos = torch.where(kvalid, start[:, None] + kk[None], start[:, None])
        face = sorted_values[pos]
        shell = torch.where(kvalid, shell_all[pos], -1)
        ty, tx = tl // cols, tl % cols
        py_i, px_i = ty[:, None] * th + iy[None], tx[:, None] * tw + ix[None]  # [G, P]
        pvalid = (py_i < h) & (px_i < w)
        px = (px_i.to(F32) + 0.5)[:, :, None]
        py = (py_i.to(F32) + 0.5)[:, :, None]
        corner = [faces[face, q] for q in range(3)]  # [G, K]
        pts = [xy[cam[:, None], corner[q]] for q in range(3)]  # [G, K, 2]
        u, v, ww, hit = _barycentrics(*(pts[q][:, None, :, r] for q in range(3) for r in range(2)), px, py)
        pairs += int((pvalid.sum(1) * n_k).sum())
        valid = hit & kvalid[:, None] & pvalid[:, :, None]
        sh_hit = torch.where(valid, shell[:, None], -1)
        before = torch.cummax(sh_hit, dim=2).values
        before = torch.cat([torch.full_like(before[..., :1], -1), before[..., :-1]], 2)
        acc = valid & (shell[:, None] > before)
        # A pixel accepts at most one key per shell: pack its accepted keys, in key order,
        # into slots [G, P, S] and composite those.
        slot = torch.cumsum(acc, dim=2, dtype=torch.int32) - 1
        g_i, p_i, k_i = acc.nonzero(as_tuple=True)
        s_i = slot[g_i, p_i, k_i].long()
        present = torch.zeros((g, P, n_slots), dtype=torch.bool)
        present[g_i, p_i, s_i] = True
        bary = []
        for src in (u, v, ww):
            dst = torch.zeros((g, P, n_slots), dtype=F32)
            dst[g_i, p_i, s_i] = src[g_i, p_i, k_i]
            bary.append(dst)
        fk = torch.zeros((g, P, n_slots), dtype=I64)
        fk[g_i, p_i, s_i] = face[g_i, k_i]
        cs = [faces[fk, q] for q in range(3)]  # [G, P, S]
        s_in = (bary[0] * sdf[cs[0]] + bary[1] * sdf[cs[1]]) + bary[2] * sdf[cs[2]]
        cidx = [cam[:, None, None] * nv + cs[q] for q in range(3)]
        col = (bary[0][..., None] * colors_flat[cidx[0]] + bary[1][..., None] * colors_flat[cidx[1]]) \
            + bary[2][..., None] * colors_flat[cidx[2]]
        s_out = torch.cat([torch.full_like(s_in[..., :1], DUMMY_OUTER_SDF), s_in[..., :-1]], 2)
        ci, co = _cdf(s_in, k), _cdf(s_out, k)
        den = co + ALPHA_EPS
        x = ((co - ci) + ALPHA_EPS) / den
        zero = torch.zeros((), dtype=F32)
        alpha = torch.where(present, x.clamp(0, 1), zero)
        t_incl = torch.cumprod(1 - alpha, dim=2)
        t_before = torch.cat([torch.ones_like(t_incl[..., :1]), t_incl[..., :-1]], 2)
        contrib = (col * t_before[..., None]) * alpha[..., None]
        c_after = torch.cumsum(contrib, dim=2)
        rgb, trans = c_after[:, :, -1], t_incl[:, :, -1]  # [G, P, 3], [G, P]
        bsel = pvalid.reshape(-1)
        gi = cam[:, None].expand(-1, P).reshape(-1)[bsel]
        yy, xx = py_i.reshape(-1)[bsel], px_i.reshape(-1)[bsel]
        image[gi, 0, yy, xx] = rgb.reshape(-1, 3)[bsel, 0]
        image[gi, 1, yy, xx] = rgb.reshape(-1, 3)[bsel, 1]
        image[gi, 2, yy, xx] = rgb.reshape(-1, 3)[bsel, 2]
        image[gi, 3, yy, xx] = trans.reshape(-1)[bsel]
        written[gi, yy, xx] = True
        if not backward:
            continue
        bg = background.to(F32)
        raw = rgb + trans[..., None] * bg
        img = raw.clamp(0, 1)
        tgt = target[cam[:, None], :, py_i.clamp(max=h - 1), px_i.clamp(max=w - 1)]  # [G, P, 3]
        diff = torch.where(pvalid[..., None], img - tgt, zero)
        loss += (diff.double() ** 2).sum()
        d_img = 2 * diff / n_el * ((raw >= 0) & (raw <= 1))
        dC, dT = d_img, (d_img * bg).sum(-1)
        weight = alpha * t_before
        suffix = rgb[:, :, None] - c_after
        inv = 1 / (1 - alpha + SUFFIX_EPS)
        d_color = weight[..., None] * dC[:, :, None]
        d_alpha = (dC[:, :, None] * (t_before[..., None] * col - suffix * inv[..., None])).sum(-1) \
            - dT[..., None] * trans[..., None] * inv
        inside = present & (x > 0) & (x < 1)
        dci, dco = ci * (1 - ci), co * (1 - co)
        d_in = torch.where(inside, d_alpha * (-1 / den) * k * dci, zero)
        d_out = torch.where(inside, d_alpha * (ci / (den * den)) * k * dco, zero)
        kg += torch.where(inside, d_alpha * ((-1 / den) * s_in * dci + (ci / (den * den)) * s_out * dco),
                          zero).double().sum()
        # The next accepted key of the pixel is the next slot (its outer role).
        d_total = d_in + torch.cat([d_out[..., 1:], torch.zeros_like(d_out[..., :1])], 2)
        sel = present.nonzero(as_tuple=True)
        for q in range(3):
            sg.index_add_(0, cs[q][sel], (bary[q] * d_total)[sel])
            cg.index_add_(0, cidx[q][sel], (bary[q][..., None] * d_color)[sel])
    if backward and whole:  # pixels of tiles without keys show the background
        bg = background.to(F32).clamp(0, 1)
        rest = (bg[None, :, None, None] - target).double() ** 2
        loss += (rest * (~written)[:, None]).sum()
    return CompositeResult(image, None if cg is None else cg.reshape(B, nv, 3), sg, float(kg),
                           float(loss) / n_el, pairs)


# ------------------------------------------------------------------------------------
# Vertex adjoint to the grid
# ------------------------------------------------------------------------------------

def vertex_backward(tess: Tessellation, num_grid: int, grad_sdf: Tensor, grad_sh: Tensor):
    """d/d(grid sdf, grid sh) of the marching-cubes interpolation: (1 - t) g to the origin,
    t g to the axis end (t detached)."""
    w0, w1 = 1 - tess.t, tess.t
    gd = torch.zeros(num_grid, dtype=F32)
    gd.index_add_(0, tess.i0, w0 * grad_sdf)
    gd.index_add_(0, tess.i1, w1 * grad_sdf)
    gs = torch.zeros((grad_sh.shape[0], num_grid, 3), dtype=F32)
    gs.index_add_(1, tess.i0, w0[None, :, None] * grad_sh)
    gs.index_add_(1, tess.i1, w1[None, :, None] * grad_sh)
    return gd, gs
