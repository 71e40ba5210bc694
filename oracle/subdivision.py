"""Oracle for the coarse-to-fine subdivision ops (write / dilate / upscale occupancy,
upscale_parameters, index_vertices).

Test infrastructure only (see oracle/__init__.py). Grids are handled unpacked (bool[i, j, k]);
integer results are exact, the parameter resampling is IEEE float32 in the reference's
operation order (upscale_parameters.slang:54-198).
"""

from __future__ import annotations

import numpy as np

from .extraction import pack_occupancy, unpack_occupancy

F32 = np.float32


def write_occupancy(voxel_indices, packed_shape) -> np.ndarray:
    """write_occupancy.slang:3-20: the listed linear cell indices set (out-of-grid ignored)."""
    i, j, kp = packed_shape
    cells = np.zeros(i * j * kp * 32, dtype=bool)
    idx = np.asarray(voxel_indices, dtype=np.int64)
    idx = idx[(idx >= 0) & (idx < cells.size)]
    cells[idx] = True
    return pack_occupancy(cells.reshape(i, j, kp * 32))


def dilate_occupancy(packed, dilation: int) -> np.ndarray:
    """dilate_occupancy.slang:9-94 as what it computes: a box dilation by ``dilation`` cells
    along every axis, clipped at the grid's borders."""
    occ = unpack_occupancy(np.asarray(packed))
    out = occ.copy()
    for axis in range(3):
        src = out.copy()
        n = src.shape[axis]
        for d in range(1, dilation + 1):
            if d >= n:
                break
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[axis], hi[axis] = slice(0, n - d), slice(d, n)
            out[tuple(hi)] |= src[tuple(lo)]
            out[tuple(lo)] |= src[tuple(hi)]
    return pack_occupancy(out)


def upscale_occupancy(packed, target_shape) -> np.ndarray:
    """upscale_occupancy.slang:1-74: finer cell (i, j, k) is occupied if the coarse cell
    ((c + s) * n_coarse - s) // n_fine (s = 0, 1 per axis) under any corner is."""
    occ = unpack_occupancy(np.asarray(packed))
    I, J, K = occ.shape
    I2, J2, K2 = target_shape
    ii, jj, kk = np.arange(I2), np.arange(J2), np.arange(K2)
    out = np.zeros((I2, J2, K2), dtype=bool)
    for si in (0, 1):
        a = ((ii + si) * I - si) // I2
        for sj in (0, 1):
            b = ((jj + sj) * J - sj) // J2
            for sk in (0, 1):
                c = ((kk + sk) * K - sk) // K2
                out |= occ[np.ix_(a, b, c)]
    return pack_occupancy(out)


def vertex_occupancy(packed) -> tuple[np.ndarray, np.ndarray]:
    """compute_vertex_occupancy.slang: bool[I + 1, J + 1, K + 1] of the corners of occupied
    cells and the row-major rank of each (-1 elsewhere)."""
    occ = unpack_occupancy(np.asarray(packed))
    I, J, K = occ.shape
    v = np.zeros((I + 1, J + 1, K + 1), dtype=bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                v[di : di + I, dj : dj + J, dk : dk + K] |= occ
    rank = np.full(v.shape, -1, dtype=np.int64)
    rank[v] = np.arange(int(v.sum()))
    return v, rank


def upscale_parameters(old_sdf, old_sh, old_packed, to_upscale_packed, new_resolution):
    """upscale_parameters.py:34-99 / .slang:54-198: (new SDF f32[V'], new SH f32[S, V', 3],
    new packed occupancy)."""
    new_packed = upscale_occupancy(np.asarray(old_packed) & np.asarray(to_upscale_packed),
                                   new_resolution)
    old_v, old_rank = vertex_occupancy(old_packed)
    new_v, _ = vertex_occupancy(new_packed)
    OI, OJ, OK = old_v.shape
    NI, NJ, NK = new_v.shape
    # The kernels see the vertex grids padded to whole words along k.
    okp, nkp = (OK - 1) // 32 + 1, (NK - 1) // 32 + 1
    inv_i, inv_j, inv_k = F32(1) / F32(NI - 1), F32(1) / F32(NJ - 1), F32(1) / F32(32 * (nkp - 1))
    sx, sy, sz = F32((okp - 1) * 32), F32(OJ - 1), F32(OI - 1)
    ni, nj, nk = np.nonzero(new_v)  # row-major: the new vertex order
    x = nk.astype(F32) * inv_k
    y = nj.astype(F32) * inv_j
    z = ni.astype(F32) * inv_i
    xu, yu, zu = x * sx, y * sy, z * sz
    ci, cj, ck = zu.astype(np.int64), yu.astype(np.int64), xu.astype(np.int64)
    ti, tj, tk = np.fmod(zu, F32(1)), np.fmod(yu, F32(1)), np.fmod(xu, F32(1))
    old_sdf = np.asarray(old_sdf, dtype=F32)
    old_sh = np.asarray(old_sh, dtype=F32)
    nsh = old_sh.shape[0]
    n = ni.shape[0]
    sh = np.zeros((nsh, n, 3), dtype=F32)
    sd = np.zeros(n, dtype=F32)
    den = np.zeros(n, dtype=F32)
    one = F32(1)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                a, b, c = ci + di, cj + dj, ck + dk
                inside = (a < OI) & (b < OJ) & (c < OK)
                r = np.full(n, -1, dtype=np.int64)
                r[inside] = old_rank[a[inside], b[inside], c[inside]]
                ok = r >= 0
                w = ((one - ti if di == 0 else ti) * (one - tj if dj == 0 else tj)) * (
                    one - tk if dk == 0 else tk)
                w = np.where(ok, w, F32(0)).astype(F32)
                rr = np.where(ok, r, 0)
                sh = np.where(ok[None, :, None], sh + w[None, :, None] * old_sh[:, rr, :], sh).astype(F32)
                sd = np.where(ok, sd + w * old_sdf[rr], sd).astype(F32)
                den = np.where(ok, den + w, den).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = (one / den).astype(F32)
    return (sd * m).astype(F32), (sh * m[None, :, None]).astype(F32), new_packed


def quantize(points, tol: float) -> np.ndarray:
    """index_vertices.py:17, ``torch.round(points / tol).to(torch.int64)`` as torch evaluates
    it on the GPU, where a float32 tensor divided by a host scalar is multiplied in float32 by
    the scalar's reciprocal, taken in double and rounded to float32; np.rint rounds half to
    even, as torch.round does. (On the CPU torch divides in float32; the two differ for about
    one point in a thousand at tol = 1e-5.)"""
    inv = F32(1.0 / tol)
    return np.rint(np.asarray(points, dtype=F32) * inv).astype(np.int64)


def index_keys(keys) -> np.ndarray:
    """index_vertices.py:33-51 by brute force: the input index of each key's first occurrence."""
    first = {}
    keys = np.asarray(keys, dtype=np.int64)
    out = np.empty(keys.shape[0], dtype=np.int32)
    for n, key in enumerate(keys.tolist()):
        out[n] = first.setdefault(key, n)
    return out


def index_vertices(points, tol: float = 1e-5) -> np.ndarray:
    """index_vertices.py:6-51 by brute force: the index of the first point (input order) with
    the same quantised coordinates, compared as coordinate triples (no key packing)."""
    q = quantize(points, tol)
    first = {}
    out = np.empty(q.shape[0], dtype=np.int32)
    for n, key in enumerate(map(tuple, q.tolist())):
        out[n] = first.setdefault(key, n)
    return out
