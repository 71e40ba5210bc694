"""Oracle for the SDF regularizers (Eikonal and curvature losses per voxel).

Test infrastructure only (see oracle/__init__.py). ``compute_sdf_regularizers`` is the
FP32-exact restatement of compute_sdf_regularizers.slang:18-128 (IEEE float32 in the
reference's operation order, no fused multiply-add); ``compute_sdf_regularizers_fp64`` is
the same function in float64 and ``compute_sdf_regularizers_backward`` its analytic
adjoint in float64, the targets for the gradient tolerance.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def _grid_scale(grid_size):
    """compute_sdf_regularizers.py:61-66: voxel_x = k, voxel_y = j, voxel_z = i."""
    i, j, k = grid_size
    return k, j, i


def _corners(sdf, neighbors, lower, upper):
    """The voxel's eight corner values (compute_sdf_regularizers.slang:39-47) and the
    subvoxel-at-(1, 1, 1) lower corners 1..3 (:110-117), -1 where absent."""
    idx = {
        "000": lower[0, : upper.shape[1]], "001": lower[1, : upper.shape[1]],
        "010": lower[2, : upper.shape[1]], "100": lower[3, : upper.shape[1]],
        "111": upper[0], "110": upper[1], "101": upper[2], "011": upper[3],
    }
    sub = neighbors[6].astype(np.int64)
    far = np.stack([lower[i, sub] for i in (1, 2, 3)]).astype(np.int64)
    far = np.where(far >= 0, far, -1)
    return {k: v.astype(np.int64) for k, v in idx.items()}, far


def _centre_gradient(d, half, one):
    """Trilinear gradient at the voxel centre (compute_sdf_regularizers.slang:50-71), terms
    summed left to right."""
    tx = ty = tz = half
    tx1, ty1, tz1 = one - tx, one - ty, one - tz
    dx = ((ty1 * tz1 * (d["001"] - d["000"]) + ty * tz1 * (d["011"] - d["010"]))
          + ty1 * tz * (d["101"] - d["100"])) + ty * tz * (d["111"] - d["110"])
    dy = ((tx1 * tz1 * (d["010"] - d["000"]) + tx * tz1 * (d["011"] - d["001"]))
          + tx1 * tz * (d["110"] - d["100"])) + tx * tz * (d["111"] - d["101"])
    dz = ((tx1 * ty1 * (d["100"] - d["000"]) + tx * ty1 * (d["101"] - d["001"]))
          + tx1 * ty * (d["110"] - d["010"])) + tx * ty * (d["111"] - d["011"])
    return dx, dy, dz


def _stable_norm(a, b, c, dtype):
    """stable_norm (compute_sdf_regularizers.slang:3-16): 0 below a scale of 1e-10."""
    scale = np.fmax(np.fmax(np.abs(a), np.abs(b)), np.abs(c))
    small = scale < dtype(1e-10)
    safe = np.where(small, dtype(1), scale)
    inv = dtype(1) / safe
    sa, sb, sc = inv * a, inv * b, inv * c
    total = (sa * sa + sb * sb) + sc * sc
    return np.where(small, dtype(0), scale * np.sqrt(total))


def _losses(sdf, neighbors, lower, upper, grid_size, dtype):
    d_idx, far = _corners(sdf, neighbors, lower, upper)
    s = sdf.astype(dtype)
    d = {k: s[v] for k, v in d_idx.items()}
    dx, dy, dz = _centre_gradient(d, dtype(0.5), dtype(1))
    vx, vy, vz = (dtype(v) for v in _grid_scale(grid_size))
    delta = _stable_norm(vx * dx, vy * dy, vz * dz, dtype) - dtype(1)
    zero = dtype(0)
    pos = np.where(delta > zero, delta * delta, zero)  # 0 + delta^2, / 1 sample (:72-86)
    neg = np.where(delta < zero, delta * delta, zero)
    total = (d["110"] + d["101"]) + d["011"]  # :105-124
    counted = np.full(total.shape, 3, dtype=np.int64)
    for i in range(3):
        present = far[i] >= 0
        total = np.where(present, total + s[np.where(present, far[i], 0)], total)
        counted += present
    total = total - counted.astype(dtype) * d["111"]
    return pos.astype(dtype), neg.astype(dtype), (total * total).astype(dtype)


def compute_sdf_regularizers(sdf, neighbors, lower, upper, grid_size):
    """FP32-exact (eikonal_loss_pos, eikonal_loss_neg, curvature_loss), each f32[N_vox]
    (compute_sdf_regularizers.slang:18-128)."""
    with np.errstate(all="ignore"):
        return _losses(np.asarray(sdf, dtype=F32), neighbors, lower, upper, grid_size, F32)


def compute_sdf_regularizers_fp64(sdf, neighbors, lower, upper, grid_size):
    """The same losses evaluated in float64."""
    with np.errstate(all="ignore"):
        return _losses(np.asarray(sdf, dtype=np.float64), neighbors, lower, upper, grid_size,
                       np.float64)


def compute_sdf_regularizers_backward(sdf, neighbors, lower, upper, grid_size, g_pos, g_neg,
                                      g_curv):
    """d (sum g_pos * pos + g_neg * neg + g_curv * curvature) / d sdf in float64: the chain
    rule through delta = |(X dx, Y dy, Z dz)| - 1 (d|v| / dv = v / |v|, 0 below the 1e-10
    cut of stable_norm) and through the Laplacian sum."""
    d_idx, far = _corners(sdf, neighbors, lower, upper)
    s = np.asarray(sdf, dtype=np.float64)
    d = {k: s[v] for k, v in d_idx.items()}
    dx, dy, dz = _centre_gradient(d, 0.5, 1.0)
    vx, vy, vz = (float(v) for v in _grid_scale(grid_size))
    a, b, c = vx * dx, vy * dy, vz * dz
    norm = _stable_norm(a, b, c, np.float64)
    delta = norm - 1.0
    g_pos = np.asarray(g_pos, dtype=np.float64)
    g_neg = np.asarray(g_neg, dtype=np.float64)
    g_curv = np.asarray(g_curv, dtype=np.float64)
    g_delta = np.where(delta > 0, 2 * delta * g_pos, 0.0) + np.where(delta < 0, 2 * delta * g_neg, 0.0)
    k = np.where(norm > 0, g_delta / np.where(norm > 0, norm, 1.0), 0.0)
    qx, qy, qz = 0.25 * k * a * vx, 0.25 * k * b * vy, 0.25 * k * c * vz
    total = (d["110"] + d["101"]) + d["011"]
    counted = np.full(total.shape, 3.0)
    for i in range(3):
        present = far[i] >= 0
        total = total + np.where(present, s[np.where(present, far[i], 0)], 0.0)
        counted += present
    total = total - counted * d["111"]
    gc = 2 * total * g_curv
    grad = np.zeros_like(s)
    contrib = {
        "000": -qx - qy - qz, "001": qx - qy - qz, "010": -qx + qy - qz, "100": -qx - qy + qz,
        "011": qx + qy - qz + gc, "101": qx - qy + qz + gc, "110": -qx + qy + qz + gc,
        "111": qx + qy + qz - counted * gc,
    }
    for key, val in contrib.items():
        np.add.at(grad, d_idx[key], val)
    for i in range(3):
        present = far[i] >= 0
        np.add.at(grad, far[i][present], gc[present])
    return grad
