"""Oracle for triangle_rasterization: tile rasterizer and NeuS-style compositing.

Test infrastructure only (see oracle/__init__.py). FP32 functions follow the
reference's elementwise operation order without fused multiply-add; the hit
decisions, keys and tile ranges they produce are the bit-exact targets. The
continuous outputs can be recomputed in float64 on frozen decisions
(:func:`composite_hits`, ``dtype=np.float64``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

F32 = np.float32
EPS_BARY = F32(1e-8)  # intersect_and_composite.slang:99 (Slang float literal)
SUM_LO = F32(0.999)  # intersect_and_composite.slang:132
SUM_HI = F32(1.001)
ALPHA_EPS = 1e-5  # intersect_and_composite.slang:161
SUFFIX_EPS = 1e-6  # intersect_and_composite.slang:3
DUMMY_OUTER_SDF = 1e10  # intersect_and_composite.slang:318

SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
         0.5462742152960396)
SH_C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
         -0.4570457994644658, 1.445305721320277, -0.5900435899266435)


# --------------------------------------------------------------------------------------
# Keys (keys.py:4-22, keys.slang:1-34)
# --------------------------------------------------------------------------------------


def key_layout(*sequence):
    """MSB-first (name, bits) sequence -> {name: (offset, length)} (keys.py:4-22)."""
    offset = 0
    out = {}
    for name, length in reversed(sequence):
        if name == "depth" and length != 32:
            raise ValueError("Depth must have 32 bits.")
        out[name] = (offset, length)
        offset += length
    assert offset == 64
    return out


def fused_layout(camera_bits=6, shell_bits=6, tile_bits=20):
    return key_layout(("camera", camera_bits), ("tile", tile_bits), ("depth", 32), ("shell", shell_bits))


def explicit_layout(camera_bits=6, shell_bits=6, tile_bits=20):
    return key_layout(("camera", camera_bits), ("shell", shell_bits), ("tile", tile_bits), ("depth", 32))


def pack_keys(layout, camera, shell, tile, depth_f32) -> np.ndarray:
    def field(name, val):
        off, length = layout[name]
        mask = np.uint64((1 << length) - 1)
        return (np.asarray(val).astype(np.uint64) & mask) << np.uint64(off)

    dbits = np.asarray(depth_f32, dtype=F32).view(np.uint32).astype(np.uint64)
    key = field("camera", camera) | field("shell", shell) | field("tile", tile)
    key = key | (dbits << np.uint64(layout["depth"][0]))
    return key.view(np.int64)


def extract_field(layout, keys, name) -> np.ndarray:
    off, length = layout[name]
    k = np.asarray(keys).view(np.uint64)
    return ((k >> np.uint64(off)) & np.uint64((1 << length) - 1)).astype(np.int64)


# --------------------------------------------------------------------------------------
# project_vertices.slang:51-92
# --------------------------------------------------------------------------------------


def project_vertices(vertices, extrinsics, intrinsics):
    """p = E [v, 1] summed left to right, depth = p.z, xy = (K p).xy / depth."""
    v = vertices.astype(F32)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    b = extrinsics.shape[0]
    out_xy = np.empty((b, v.shape[0], 2), dtype=F32)
    out_d = np.empty((b, v.shape[0]), dtype=F32)
    for c in range(b):
        E = extrinsics[c].astype(F32)
        K = intrinsics[c].astype(F32)
        cs = [((E[r, 0] * x + E[r, 1] * y) + E[r, 2] * z) + E[r, 3] for r in range(3)]
        depth = cs[2]
        with np.errstate(divide="ignore", invalid="ignore"):
            for r in range(2):
                p = (K[r, 0] * cs[0] + K[r, 1] * cs[1]) + K[r, 2] * cs[2]
                out_xy[c, :, r] = p / depth
        out_d[c] = depth
    return out_xy, out_d


# --------------------------------------------------------------------------------------
# bound_triangles.slang:62-192
# --------------------------------------------------------------------------------------


def _f2i(x):
    """float -> int32 truncating toward zero, saturating, NaN -> 0 (device cvt semantics)."""
    x = np.asarray(x, dtype=np.float64)
    x = np.where(np.isnan(x), 0.0, x)
    x = np.clip(np.trunc(x), -(2.0**31), 2.0**31 - 1)
    return x.astype(np.int64)


def _c_div(a, b):
    """C integer division (truncation toward zero)."""
    q = np.abs(a) // np.abs(b)
    return np.where((a < 0) != (b < 0), -q, q)


def bound_triangles(xy, depths, faces, tile_shape, grid_shape, near_plane):
    th, tw = tile_shape
    rows, cols = grid_shape
    b = xy.shape[0]
    f = faces.shape[0]
    tmin = np.zeros((b, f, 2), dtype=np.int32)
    tmax = np.zeros((b, f, 2), dtype=np.int32)
    count = np.zeros((b, f), dtype=np.int32)
    fa, fb, fc = faces[:, 0], faces[:, 1], faces[:, 2]
    for c in range(b):
        ax, ay, ad = xy[c, fa, 0], xy[c, fa, 1], depths[c, fa]
        bx, by, bd = xy[c, fb, 0], xy[c, fb, 1], depths[c, fb]
        cx, cy, cd = xy[c, fc, 0], xy[c, fc, 1], depths[c, fc]
        check = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
        min_depth = np.fmin(np.fmin(ad, bd), cd)
        culled = (check < 0) | (min_depth < F32(near_plane))

        def bounds(lo, hi, ts, gmax):
            t0 = _c_div(_f2i(lo), ts)
            with np.errstate(invalid="ignore"):
                t1 = _f2i(np.ceil(hi / F32(ts)))
            return np.clip(t0, 0, gmax), np.clip(t1, 0, gmax)

        ry0, ry1 = bounds(np.fmin(np.fmin(ay, by), cy), np.fmax(np.fmax(ay, by), cy), th, rows)
        rx0, rx1 = bounds(np.fmin(np.fmin(ax, bx), cx), np.fmax(np.fmax(ax, bx), cx), tw, cols)
        z = np.zeros_like(ry0)
        ry0, ry1, rx0, rx1 = (np.where(culled, z, q) for q in (ry0, ry1, rx0, rx1))
        tmin[c, :, 0], tmin[c, :, 1] = ry0, rx0
        tmax[c, :, 0], tmax[c, :, 1] = ry1, rx1
        count[c] = ((ry1 - ry0) * (rx1 - rx0)).astype(np.int32)
    return tmin, tmax, count


# --------------------------------------------------------------------------------------
# generate_keys.py:34-70 / generate_keys.slang:4-68
# --------------------------------------------------------------------------------------


def generate_keys(depths, faces, shell_face_boundaries, tmin, tmax, counts, grid_cols, layout):
    b, f = counts.shape
    offsets = np.cumsum(counts.reshape(-1).astype(np.int64))
    num_keys = int(offsets[-1]) if offsets.size else 0
    keys = np.zeros(num_keys, dtype=np.int64)
    values = np.zeros(num_keys, dtype=np.int32)
    if num_keys == 0:
        return keys, values, offsets
    face_ids = np.arange(f)
    shell = np.searchsorted(shell_face_boundaries.astype(np.int64), face_ids, side="right")
    flat_counts = counts.reshape(-1).astype(np.int64)
    owner = np.repeat(np.arange(b * f), flat_counts)
    start = np.concatenate([[0], offsets[:-1]])
    local = np.arange(num_keys) - start[owner]
    cam = owner // f
    face = owner % f
    r0 = tmin[cam, face, 0].astype(np.int64)
    c0 = tmin[cam, face, 1].astype(np.int64)
    width = (tmax[cam, face, 1] - tmin[cam, face, 1]).astype(np.int64)
    row = r0 + local // width
    col = c0 + local % width
    tile = row * grid_cols + col
    da = depths[cam, faces[face, 0]]
    db = depths[cam, faces[face, 1]]
    dc = depths[cam, faces[face, 2]]
    mean_depth = ((da + db) + dc) / F32(3)
    keys[:] = pack_keys(layout, cam, shell[face], tile, mean_depth.astype(F32))
    values[:] = face.astype(np.int32)
    return keys, values, offsets


def sort_keys(keys, values):
    """Stable ascending sort on signed int64 keys (CUB SortPairs, sort_keys.cu:241-267)."""
    order = np.argsort(keys, kind="stable")
    return keys[order], values[order]


def delineate_tiles(sorted_keys, layout, num_cameras, num_shells, num_tiles, separate_shells):
    """delineate_tiles.slang:12-58: [start, end) of each (camera[, shell], tile) run; empty = (0, 0)."""
    s = num_shells if separate_shells else 1
    out = np.zeros((num_cameras, s, num_tiles, 2), dtype=np.int64)
    n = sorted_keys.shape[0]
    if n == 0:
        return out
    cam = extract_field(layout, sorted_keys, "camera")
    shell = extract_field(layout, sorted_keys, "shell") if separate_shells else np.zeros(n, np.int64)
    tile = extract_field(layout, sorted_keys, "tile")
    change = np.ones(n, dtype=bool)
    change[1:] = (cam[1:] != cam[:-1]) | (shell[1:] != shell[:-1]) | (tile[1:] != tile[:-1])
    starts = np.nonzero(change)[0]
    ends = np.concatenate([starts[1:], [n]])
    out[cam[starts], shell[starts], tile[starts], 0] = np.where(starts == 0, 0, starts)
    out[cam[starts], shell[starts], tile[starts], 1] = ends
    return out


# --------------------------------------------------------------------------------------
# evaluate_spherical_harmonics.slang:124-270
# --------------------------------------------------------------------------------------


def camera_positions(extrinsics):
    """inverse(E)[:, :3, 3] (evaluate_spherical_harmonics.py:38), in float64."""
    return np.linalg.inv(extrinsics.astype(np.float64))[:, :3, 3]


def sh_basis(dirs, sh_degree, dtype=np.float64):
    """Real SH basis values [..., NSH] with the reference's signs (slang:132-154)."""
    x, y, z = (dirs[..., i].astype(dtype) for i in range(3))
    out = [np.full_like(x, SH_C0)]
    if sh_degree > 0:
        out += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if sh_degree > 1:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        out += [SH_C2[0] * xy, SH_C2[1] * yz, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * xz,
                SH_C2[4] * (xx - yy)]
    if sh_degree > 2:
        out += [
            SH_C3[0] * y * (3 * xx - yy),
            SH_C3[1] * xy * z,
            SH_C3[2] * y * (4 * zz - xx - yy),
            SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy),
            SH_C3[4] * x * (4 * zz - xx - yy),
            SH_C3[5] * z * (xx - yy),
            SH_C3[6] * x * (xx - 3 * yy),
        ]
    return np.stack(out, axis=-1)


def evaluate_spherical_harmonics(vertices, sh, extrinsics, sh_degree, dtype=np.float64):
    """colors[b, v] = max(sum_k basis_k(normalize(v - cam_b)) sh[k, v] + 0.5, 0)."""
    nsh = (sh_degree + 1) ** 2
    cams = camera_positions(extrinsics)
    d = vertices.astype(dtype)[None] - cams[:, None].astype(dtype)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    basis = sh_basis(d, sh_degree, dtype)  # [B, V, nsh]
    pre = np.einsum("bvk,kvc->bvc", basis, sh[:nsh].astype(dtype)) + 0.5
    return np.maximum(pre, 0), pre


def evaluate_spherical_harmonics_backward(vertices, sh, extrinsics, sh_degree, grad_colors):
    nsh = (sh_degree + 1) ** 2
    cams = camera_positions(extrinsics)
    d = vertices.astype(np.float64)[None] - cams[:, None]
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    basis = sh_basis(d, sh_degree)
    _, pre = evaluate_spherical_harmonics(vertices, sh, extrinsics, sh_degree)
    g = grad_colors.astype(np.float64) * (pre > 0)
    out = np.zeros(sh.shape, dtype=np.float64)
    out[:nsh] = np.einsum("bvk,bvc->kvc", basis, g)
    return out


# --------------------------------------------------------------------------------------
# Barycentric hit test (intersect_and_composite.slang:65-145, intersect.slang:84-139)
# --------------------------------------------------------------------------------------


def barycentrics(p1, p2, p3, px, py):
    """FP32 barycentrics in the reference's order. p* are (x, y) float32 arrays broadcastable
    against the pixel-centre arrays px, py. Returns (u, v, w, hit)."""
    e12x, e12y = p2[0] - p1[0], p2[1] - p1[1]
    e23x, e23y = p3[0] - p2[0], p3[1] - p2[1]
    e13x, e13y = p3[0] - p1[0], p3[1] - p1[1]
    d_12_12 = e12x * e12x + e12y * e12y
    d_12_23 = e12x * e23x + e12y * e23y
    d_12_13 = e12x * e13x + e12y * e13y
    d_13_13 = e13x * e13x + e13y * e13y
    d_13_23 = e13x * e23x + e13y * e23y
    d_23_23 = e23x * e23x + e23y * e23y
    a1x, a1y = px - p1[0], py - p1[1]
    a2x, a2y = px - p2[0], py - p2[1]
    d_1x_12 = a1x * e12x + a1y * e12y
    d_1x_13 = a1x * e13x + a1y * e13y
    d_2x_12 = a2x * e12x + a2y * e12y
    d_2x_13 = a2x * e13x + a2y * e13y
    a, b, c = np.abs(d_12_13), np.abs(d_12_23), np.abs(d_13_23)
    den1 = d_12_12 * d_13_13 - d_12_13 * d_12_13
    den2 = d_12_12 * d_23_23 - d_12_23 * d_12_23
    den3 = d_13_13 * d_23_23 - d_13_23 * d_13_23
    c1 = (a <= b) & (a <= c)
    c2 = (b <= a) & (b <= c)
    c3 = (c <= a) & (c <= b)
    denom = np.where(c1, den1, np.where(c2, den2, np.where(c3, den3, F32(np.nan))))
    v = d_13_13 * d_1x_12 - d_12_13 * d_1x_13
    w = d_12_12 * d_1x_13 - d_12_13 * d_1x_12
    u = d_12_23 * d_2x_13 - d_13_23 * d_2x_12
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        dd = denom + EPS_BARY
        v = v / dd
        w = w / dd
        u = u / dd
        s = (u + v) + w
        miss = (
            np.isnan(u) | np.isnan(v) | np.isnan(w)
            | (u < 0) | (u > 1) | (v < 0) | (v > 1) | (w < 0) | (w > 1)
            | (s < SUM_LO) | (s > SUM_HI)
        )
    return u.astype(F32), v.astype(F32), w.astype(F32), ~miss


def barycentrics_f64(p1, p2, p3, px, py):
    """The same formula evaluated in float64 (continuous quantities on frozen hits)."""
    p1, p2, p3 = (np.asarray(p, dtype=np.float64) for p in (p1, p2, p3))
    px, py = np.asarray(px, np.float64), np.asarray(py, np.float64)
    e12 = (p2[0] - p1[0], p2[1] - p1[1])
    e23 = (p3[0] - p2[0], p3[1] - p2[1])
    e13 = (p3[0] - p1[0], p3[1] - p1[1])
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1]  # noqa: E731
    d_12_12, d_12_23, d_12_13 = dot(e12, e12), dot(e12, e23), dot(e12, e13)
    d_13_13, d_13_23, d_23_23 = dot(e13, e13), dot(e13, e23), dot(e23, e23)
    a1 = (px - p1[0], py - p1[1])
    a2 = (px - p2[0], py - p2[1])
    d_1x_12, d_1x_13, d_2x_12, d_2x_13 = dot(a1, e12), dot(a1, e13), dot(a2, e12), dot(a2, e13)
    a, b, c = np.abs(d_12_13), np.abs(d_12_23), np.abs(d_13_23)
    den = np.where(
        (a <= b) & (a <= c), d_12_12 * d_13_13 - d_12_13 ** 2,
        np.where((b <= a) & (b <= c), d_12_12 * d_23_23 - d_12_23 ** 2, d_13_13 * d_23_23 - d_13_23 ** 2),
    ) + 1e-8
    v = (d_13_13 * d_1x_12 - d_12_13 * d_1x_13) / den
    w = (d_12_12 * d_1x_13 - d_12_13 * d_1x_12) / den
    u = (d_12_23 * d_2x_13 - d_13_23 * d_2x_12) / den
    return u, v, w


# --------------------------------------------------------------------------------------
# Fused traversal (intersect_and_composite.slang:224-339): per-pixel hit lists
# --------------------------------------------------------------------------------------


@dataclass
class HitList:
    """Accepted hits of every pixel, in traversal order.

    Arrays are parallel; ``pixel`` is the flat index b * H * W + y * W + x and entries
    of one pixel are contiguous and in acceptance order.
    """

    pixel: np.ndarray  # int64
    face: np.ndarray  # int64
    shell: np.ndarray  # int64
    u: np.ndarray  # f32
    v: np.ndarray
    w: np.ndarray
    num_tests: int = 0  # (pixel, key) barycentric evaluations the traversal executes


def _tile_pixels(tile_index, grid_cols, tile_shape, image_shape):
    th, tw = tile_shape
    h, w = image_shape
    ty, tx = divmod(tile_index, grid_cols)
    ys, xs = np.meshgrid(np.arange(th) + ty * th, np.arange(tw) + tx * tw, indexing="ij")
    ys, xs = ys.reshape(-1), xs.reshape(-1)
    inside = (xs < w) & (ys < h)
    return ys[inside], xs[inside]


def fused_hits(xy, faces, sorted_keys, sorted_values, tile_ranges, layout, num_shells,
               tile_shape, image_shape) -> HitList:
    """Which triangles each pixel accepts: in sorted-key order, a triangle is accepted if
    the pixel centre passes the barycentric test and its shell is greater than the last
    accepted shell; traversal stops after shell S - 1 (slang:480-505)."""
    b, t, _ = tile_ranges.shape
    h, w = image_shape
    th, tw = tile_shape
    grid_cols = -(-w // tw)
    shells_all = extract_field(layout, sorted_keys, "shell")
    out = {k: [] for k in ("pixel", "face", "shell", "u", "v", "w")}
    num_tests = 0
    for c in range(b):
        for ti in range(t):
            start, end = (int(q) for q in tile_ranges[c, ti])
            if end <= start:
                continue
            ys, xs = _tile_pixels(ti, grid_cols, tile_shape, image_shape)
            if ys.size == 0:
                continue
            fidx = sorted_values[start:end].astype(np.int64)
            shells = shells_all[start:end]
            pa, pb, pc = (xy[c, faces[fidx, k]] for k in range(3))
            px = (xs.astype(F32) + F32(0.5))[:, None]
            py = (ys.astype(F32) + F32(0.5))[:, None]
            u, v, ww, hit = barycentrics(
                (pa[None, :, 0], pa[None, :, 1]), (pb[None, :, 0], pb[None, :, 1]),
                (pc[None, :, 0], pc[None, :, 1]), px, py,
            )
            npx = ys.size
            last = np.full(npx, -1, dtype=np.int64)
            active = np.ones(npx, dtype=bool)
            rec = []
            for kk in range(end - start):
                num_tests += int(np.count_nonzero(active & (shells[kk] > last)))
                cand = active & hit[:, kk] & (shells[kk] > last)
                if not cand.any():
                    continue
                pix = np.nonzero(cand)[0]
                rec.append((pix, np.full(pix.size, kk)))
                last[pix] = shells[kk]
                if shells[kk] == num_shells - 1:
                    active[pix] = False
            if not rec:
                continue
            pix = np.concatenate([r[0] for r in rec])
            kk = np.concatenate([r[1] for r in rec])
            order = np.lexsort((kk, pix))  # per pixel, in traversal order
            pix, kk = pix[order], kk[order]
            flat = c * h * w + ys[pix].astype(np.int64) * w + xs[pix]
            out["pixel"].append(flat)
            out["face"].append(fidx[kk])
            out["shell"].append(shells[kk])
            out["u"].append(u[pix, kk])
            out["v"].append(v[pix, kk])
            out["w"].append(ww[pix, kk])
    if out["pixel"]:
        arr = {k: np.concatenate(val) for k, val in out.items()}
        order = np.argsort(arr["pixel"], kind="stable")
        arr = {k: val[order] for k, val in arr.items()}
    else:
        arr = {k: np.zeros(0, np.int64 if k in ("pixel", "face", "shell") else F32) for k in out}
    return HitList(**arr, num_tests=num_tests)


def _rank_within_pixel(pixel):
    n = pixel.size
    if n == 0:
        return np.zeros(0, np.int64)
    start = np.ones(n, dtype=bool)
    start[1:] = pixel[1:] != pixel[:-1]
    idx = np.arange(n)
    first = np.maximum.accumulate(np.where(start, idx, 0))
    return idx - first


def sdf_to_cdf(sdf, k):
    with np.errstate(over="ignore"):
        return 1 / (1 + np.exp(-sdf * k))


def composite_hits(hits: HitList, colors, sdf, faces, sharpness, num_pixels, dtype=F32,
                   uvw=None):
    """Front-to-back NeuS compositing along each pixel's hit list (slang:153-176, 480-513).

    Returns (rgb [P, 3], transmittance [P]) over the flat pixel index. ``uvw`` may
    override the stored FP32 barycentrics (the FP64 path recomputes them)."""
    dt = dtype
    k = dt(sharpness)
    rgb = np.zeros((num_pixels, 3), dtype=dt)
    trans = np.ones(num_pixels, dtype=dt)
    if hits.pixel.size == 0:
        return rgb, trans
    u, v, w = (hits.u, hits.v, hits.w) if uvw is None else uvw
    u, v, w = u.astype(dt), v.astype(dt), w.astype(dt)
    fa, fb, fc = (faces[hits.face, i] for i in range(3))
    hw = None
    rank = _rank_within_pixel(hits.pixel)
    outer_sdf = np.full(num_pixels, dt(DUMMY_OUTER_SDF), dtype=dt)
    # Camera of each hit, from the flat pixel index (colors are per camera).
    b = colors.shape[0]
    hw = num_pixels // b
    cam = hits.pixel // hw
    for r in range(int(rank.max()) + 1):
        sel = rank == r
        p = hits.pixel[sel]
        uu, vv, wv = u[sel], v[sel], w[sel]
        a, bb, cc = fa[sel], fb[sel], fc[sel]
        cm = cam[sel]
        s_in = (uu * sdf[a].astype(dt) + vv * sdf[bb].astype(dt)) + wv * sdf[cc].astype(dt)
        col = (uu[:, None] * colors[cm, a].astype(dt) + vv[:, None] * colors[cm, bb].astype(dt)) \
            + wv[:, None] * colors[cm, cc].astype(dt)
        i_cdf = sdf_to_cdf(s_in, k).astype(dt)
        o_cdf = sdf_to_cdf(outer_sdf[p], k).astype(dt)
        alpha = ((o_cdf - i_cdf) + dt(ALPHA_EPS)) / (o_cdf + dt(ALPHA_EPS))
        alpha = np.fmin(np.fmax(alpha, dt(0)), dt(1))
        rgb[p] = rgb[p] + (col * trans[p][:, None]) * alpha[:, None]
        trans[p] = trans[p] * (dt(1) - alpha)
        outer_sdf[p] = s_in
    return rgb, trans


def fused_forward_image(hits, colors, sdf, faces, sharpness, b, image_shape, dtype=F32, uvw=None):
    h, w = image_shape
    rgb, trans = composite_hits(hits, colors, sdf, faces, sharpness, b * h * w, dtype, uvw)
    out = np.empty((b, 4, h, w), dtype=dtype)
    out[:, :3] = rgb.reshape(b, h, w, 3).transpose(0, 3, 1, 2)
    out[:, 3] = trans.reshape(b, h, w)
    return out


def _alpha_and_grads(s_in, s_out, k):
    """alpha and its partials w.r.t. (inner sdf, outer sdf, sharpness), float64.

    The clamp passes gradient only strictly inside (0, 1) (the build's choice at the
    bounds; the sigmoid derivative uses the overflow-free form c (1 - c))."""
    ci = sdf_to_cdf(s_in, k)
    co = sdf_to_cdf(s_out, k)
    den = co + ALPHA_EPS
    x = ((co - ci) + ALPHA_EPS) / den
    alpha = np.clip(x, 0, 1)
    inside = (x > 0) & (x < 1)
    dx_di = -1 / den
    dx_do = ci / (den * den)
    dci = ci * (1 - ci)
    dco = co * (1 - co)
    d_in = np.where(inside, dx_di * k * dci, 0.0)
    d_out = np.where(inside, dx_do * k * dco, 0.0)
    d_k = np.where(inside, dx_di * s_in * dci + dx_do * s_out * dco, 0.0)
    return alpha, d_in, d_out, d_k


def fused_backward(hits: HitList, colors, sdf, faces, sharpness, out_image, grad_image,
                   reference_quirk=False, uvw=None):
    """Suffix-trick adjoint of the fused composite (intersect_and_composite.slang:341-564),
    in float64 on the frozen hit lists.

    Each accepted sample's gradient is complete only after the next hit (its SDF also
    sets the next alpha); it is then scattered to its triangle's vertices with that
    triangle's barycentrics. ``reference_quirk=True`` reproduces the reference, which
    evaluates the barycentrics of the *next* triangle instead (slang:452, 497-513).
    Returns (colors_grad [B, V, 3], sdf_grad [V], sharpness_grad)."""
    b, _, h, w = out_image.shape
    k = float(sharpness)
    cg = np.zeros(colors.shape, dtype=np.float64)
    sg = np.zeros(sdf.shape, dtype=np.float64)
    kg = 0.0
    if hits.pixel.size == 0:
        return cg, sg, kg
    u, v, ww = (hits.u, hits.v, hits.w) if uvw is None else uvw
    u, v, ww = (np.asarray(q, np.float64) for q in (u, v, ww))
    fa, fb, fc = (faces[hits.face, i].astype(np.int64) for i in range(3))
    hw = h * w
    cam = hits.pixel // hw
    pix_in = hits.pixel % hw
    final = out_image.reshape(b, 4, hw).astype(np.float64)
    grad = grad_image.reshape(b, 4, hw).astype(np.float64)
    C_f = final[cam, :3, pix_in]
    T_f = final[cam, 3, pix_in]
    dC = grad[cam, :3, pix_in]
    dT = grad[cam, 3, pix_in]
    s_in = u * sdf[fa] + v * sdf[fb] + ww * sdf[fc]
    col = u[:, None] * colors[cam, fa] + v[:, None] * colors[cam, fb] + ww[:, None] * colors[cam, fc]
    rank = _rank_within_pixel(hits.pixel)
    n = hits.pixel.size
    s_out = np.where(rank == 0, DUMMY_OUTER_SDF, np.concatenate([[0.0], s_in[:-1]]))
    alpha, d_in_a, d_out_a, d_k_a = _alpha_and_grads(s_in, s_out, k)
    # Transmittance before each hit and the accumulated colour after it.
    T_before = np.ones(n)
    C_after = np.zeros((n, 3))
    T_run = None
    C_run = None
    for i in range(n):
        if rank[i] == 0:
            T_run, C_run = 1.0, np.zeros(3)
        T_before[i] = T_run
        C_run = C_run + alpha[i] * T_run * col[i]
        C_after[i] = C_run
        T_run = T_run * (1 - alpha[i])
    weight = alpha * T_before
    suffix = C_f - C_after
    inv = 1 / (1 - alpha + SUFFIX_EPS)
    d_color = weight[:, None] * dC
    d_alpha = np.sum(dC * (T_before[:, None] * col - suffix * inv[:, None]), axis=1) - dT * T_f * inv
    d_sdf_inner = d_alpha * d_in_a
    d_sdf_outer = d_alpha * d_out_a
    kg = float(np.sum(d_alpha * d_k_a))
    # Sample i's SDF gradient also receives the outer-role term of sample i + 1.
    last_of_pixel = np.ones(n, dtype=bool)
    last_of_pixel[:-1] = hits.pixel[1:] != hits.pixel[:-1]
    d_sdf_total = d_sdf_inner.copy()
    nxt = ~last_of_pixel
    d_sdf_total[np.nonzero(nxt)[0]] += d_sdf_outer[np.nonzero(nxt)[0] + 1]
    # Barycentrics used for the scatter of sample i.
    bu, bv, bw = u.copy(), v.copy(), ww.copy()
    if reference_quirk:
        idx = np.nonzero(nxt)[0]
        bu[idx], bv[idx], bw[idx] = u[idx + 1], v[idx + 1], ww[idx + 1]
    for corner, wgt in ((fa, bu), (fb, bv), (fc, bw)):
        np.add.at(sg, corner, wgt * d_sdf_total)
        np.add.at(cg, (cam, corner), wgt[:, None] * d_color)
    return cg, sg, kg


# --------------------------------------------------------------------------------------
# Explicit mode (intersect.slang, interpolate.slang, composite.slang)
# --------------------------------------------------------------------------------------


def explicit_intersect(xy, faces, sorted_values, tile_ranges, tile_shape, image_shape):
    """First hit in depth order per (camera, shell, pixel) (intersect.slang:85-182).
    Returns uv f32[B, S, H, W, 2] and index i32[B, S, H, W] (-1 = miss)."""
    b, s, t, _ = tile_ranges.shape
    h, w = image_shape
    th, tw = tile_shape
    grid_cols = -(-w // tw)
    uv = np.zeros((b, s, h, w, 2), dtype=F32)
    index = np.full((b, s, h, w), -1, dtype=np.int32)
    for c in range(b):
        for sh in range(s):
            for ti in range(t):
                start, end = (int(q) for q in tile_ranges[c, sh, ti])
                if end <= start:
                    continue
                ys, xs = _tile_pixels(ti, grid_cols, tile_shape, image_shape)
                if ys.size == 0:
                    continue
                fidx = sorted_values[start:end].astype(np.int64)
                pa, pb, pc = (xy[c, faces[fidx, q]] for q in range(3))
                px = (xs.astype(F32) + F32(0.5))[:, None]
                py = (ys.astype(F32) + F32(0.5))[:, None]
                u, v, _, hit = barycentrics(
                    (pa[None, :, 0], pa[None, :, 1]), (pb[None, :, 0], pb[None, :, 1]),
                    (pc[None, :, 0], pc[None, :, 1]), px, py,
                )
                anyhit = hit.any(axis=1)
                first = np.argmax(hit, axis=1)
                pi = np.nonzero(anyhit)[0]
                kk = first[pi]
                index[c, sh, ys[pi], xs[pi]] = fidx[kk]
                uv[c, sh, ys[pi], xs[pi], 0] = u[pi, kk]
                uv[c, sh, ys[pi], xs[pi], 1] = v[pi, kk]
    return uv, index


def interpolate(uv, index, faces, sdf, colors, dtype=F32):
    """interpolate.slang:19-76: (rgb, sdf) = u a + v b + (1 - u - v) c, miss -> 0."""
    b, s, h, w, _ = uv.shape
    dt = dtype
    u = uv[..., 0].astype(dt)
    v = uv[..., 1].astype(dt)
    wv = (dt(1) - u) - v
    hit = index >= 0
    fi = np.where(hit, index, 0)
    cam = np.broadcast_to(np.arange(b)[:, None, None, None], index.shape)
    out_sdf = np.zeros(index.shape, dtype=dt)
    out_rgb = np.zeros(index.shape + (3,), dtype=dt)
    va, vb, vc = (faces[fi, q] for q in range(3))
    s_ = (u * sdf[va].astype(dt) + v * sdf[vb].astype(dt)) + wv * sdf[vc].astype(dt)
    c_ = (u[..., None] * colors[cam, va].astype(dt) + v[..., None] * colors[cam, vb].astype(dt)) \
        + wv[..., None] * colors[cam, vc].astype(dt)
    out_sdf[hit] = s_[hit]
    out_rgb[hit] = c_[hit]
    return out_sdf, out_rgb


def interpolate_backward(uv, index, faces, num_vertices, num_cameras, grad_sdf, grad_rgb):
    u = uv[..., 0].astype(np.float64)
    v = uv[..., 1].astype(np.float64)
    wv = 1 - u - v
    hit = index >= 0
    cam = np.broadcast_to(np.arange(num_cameras)[:, None, None, None], index.shape)[hit]
    fi = index[hit]
    sg = np.zeros(num_vertices)
    cg = np.zeros((num_cameras, num_vertices, 3))
    gs = grad_sdf[hit].astype(np.float64)
    gc = grad_rgb[hit].astype(np.float64)
    for q, wgt in enumerate((u[hit], v[hit], wv[hit])):
        corner = faces[fi, q]
        np.add.at(sg, corner, wgt * gs)
        np.add.at(cg, (cam, corner), wgt[:, None] * gc)
    return sg, cg


def composite(sdf_pl, rgb_pl, mask_pl, sharpness, dtype=F32):
    """composite.slang:60-104 per pixel over shells in order, masked."""
    p, s = sdf_pl.shape
    dt = dtype
    k = dt(sharpness)
    color = np.zeros((p, 3), dtype=dt)
    trans = np.ones(p, dtype=dt)
    outer = np.full(p, dt(DUMMY_OUTER_SDF), dtype=dt)
    for sh in range(s):
        m = mask_pl[:, sh]
        s_in = sdf_pl[m, sh].astype(dt)
        ci = sdf_to_cdf(s_in, k).astype(dt)
        co = sdf_to_cdf(outer[m], k).astype(dt)
        alpha = np.fmin(np.fmax(((co - ci) + dt(ALPHA_EPS)) / (co + dt(ALPHA_EPS)), dt(0)), dt(1))
        color[m] = color[m] + (rgb_pl[m, sh].astype(dt) * trans[m][:, None]) * alpha[:, None]
        trans[m] = trans[m] * (dt(1) - alpha)
        outer[m] = s_in
    return color, trans


def composite_backward(sdf_pl, rgb_pl, mask_pl, sharpness, out_color, out_trans, grad_color,
                       grad_trans, reference_quirk=False):
    """Suffix-trick adjoint (composite.slang:106-218), float64.

    The delayed SDF gradient of a shell goes to the previous *unmasked* shell; the
    reference writes it to shell_index - 1 and the last one to num_shells - 1
    (composite.slang:197, 213), which differs when a masked shell sits in between.
    ``reference_quirk=True`` reproduces that."""
    p, s = sdf_pl.shape
    k = float(sharpness)
    gs = np.zeros((p, s))
    gc = np.zeros((p, s, 3))
    gk = 0.0
    Cf = out_color.astype(np.float64)
    Tf = out_trans.astype(np.float64)
    dC = grad_color.astype(np.float64)
    dT = grad_trans.astype(np.float64)
    color = np.zeros((p, 3))
    trans = np.ones(p)
    outer = np.full(p, DUMMY_OUTER_SDF)
    first = np.ones(p, dtype=bool)
    partial = np.zeros(p)
    prev_shell = np.full(p, -1, dtype=np.int64)
    rows = np.arange(p)
    for sh in range(s):
        m = mask_pl[:, sh]
        r = rows[m]
        s_in = sdf_pl[m, sh].astype(np.float64)
        col = rgb_pl[m, sh].astype(np.float64)
        alpha, da_in, da_out, da_k = _alpha_and_grads(s_in, outer[m], k)
        weight = alpha * trans[m]
        color[m] = color[m] + weight[:, None] * col
        suffix = Cf[m] - color[m]
        inv = 1 / (1 - alpha + SUFFIX_EPS)
        gc[m, sh] += weight[:, None] * dC[m]
        d_alpha = np.sum(dC[m] * (trans[m][:, None] * col - suffix * inv[:, None]), axis=1) \
            - dT[m] * Tf[m] * inv
        gk += float(np.sum(d_alpha * da_k))
        nf = ~first[m]
        rr = r[nf]
        tgt = (sh - 1) if reference_quirk else prev_shell[rr]
        gs[rr, tgt] += partial[rr] + (d_alpha * da_out)[nf]
        partial[r] = d_alpha * da_in
        trans[m] = trans[m] * (1 - alpha)
        outer[m] = s_in
        first[m] = False
        prev_shell[r] = sh
    done = ~first
    rr = rows[done]
    tgt = (s - 1) if reference_quirk else prev_shell[rr]
    gs[rr, tgt] += partial[rr]
    return gs, gc, gk
