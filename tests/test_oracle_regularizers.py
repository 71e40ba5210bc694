"""Pins the SDF-regularizer oracle (oracle/regularizers.py) by closed forms and by finite
differences: a linear field has Eikonal loss (|g| - 1)^2 in every voxel and no curvature
where all six Laplacian neighbours exist; the quadratic field x^2 has curvature sum 2 / K^2;
the analytic FP64 adjoint matches central differences of the FP64 forward."""

import numpy as np
import pytest

from oracle import extraction as ox
from oracle import regularizers as orz


def _voxels(seed, shape=(8, 12, 32), p=0.35):
    rng = np.random.default_rng(seed)
    occ = rng.random(shape) < p
    occ[2:6, 3:9, 8:24] = True  # a solid block: interior voxels have every neighbour
    return ox.extract_voxels(ox.pack_occupancy(occ))


def test_linear_field_closed_form():
    v = _voxels(0)
    i, j, k = 8, 12, 32
    for g in (np.array([0.3, -0.4, 0.5]), np.array([2.0, 1.0, -2.0])):
        sdf = (v.vertices.astype(np.float64) @ g).astype(np.float32)
        pos, neg, curv = orz.compute_sdf_regularizers(sdf, v.neighbors, v.lower_corners,
                                                      v.upper_corners, (i, j, k))
        want = (np.linalg.norm(g) - 1.0) ** 2
        got = pos + neg
        assert np.allclose(got, want, rtol=1e-4, atol=1e-6)
        assert (pos > 0).all() == (np.linalg.norm(g) > 1) and (neg > 0).all() == (np.linalg.norm(g) < 1)
        # Curvature of a linear field vanishes where all six neighbours are present.
        full = (v.lower_corners[1:, v.neighbors[6]] >= 0).all(axis=0)
        assert full.sum() > 50
        assert np.abs(curv[full]).max() < 1e-8


def test_quadratic_field_curvature():
    v = _voxels(1)
    i, j, k = 8, 12, 32
    x = v.vertices[:, 0].astype(np.float64)
    sdf = (x * x).astype(np.float32)
    _, _, curv = orz.compute_sdf_regularizers_fp64(sdf.astype(np.float64), v.neighbors,
                                                   v.lower_corners, v.upper_corners, (i, j, k))
    full = (v.lower_corners[1:, v.neighbors[6]] >= 0).all(axis=0)
    # Second difference of x^2 along x with spacing 1/K is 2 / K^2; along y and z it is 0.
    assert np.allclose(np.sqrt(curv[full]), 2.0 / k**2, rtol=1e-6)


@pytest.mark.parametrize("seed", [2, 3])
def test_adjoint_matches_central_differences(seed):
    v = _voxels(seed, shape=(4, 5, 32), p=0.25)
    i, j, k = 4, 5, 32
    rng = np.random.default_rng(seed)
    sdf = rng.normal(0.0, 0.05, v.vertices.shape[0])
    n = v.num_voxels
    g_pos, g_neg, g_curv = rng.random(n), rng.random(n), rng.random(n)

    def objective(s):
        p, q, c = orz.compute_sdf_regularizers_fp64(s, v.neighbors, v.lower_corners,
                                                    v.upper_corners, (i, j, k))
        return float((g_pos * p).sum() + (g_neg * q).sum() + (g_curv * c).sum())

    grad = orz.compute_sdf_regularizers_backward(sdf, v.neighbors, v.lower_corners,
                                                 v.upper_corners, (i, j, k), g_pos, g_neg, g_curv)
    h = 1e-6
    picks = rng.choice(sdf.shape[0], size=min(60, sdf.shape[0]), replace=False)
    for idx in picks:
        e = np.zeros_like(sdf)
        e[idx] = h
        fd = (objective(sdf + e) - objective(sdf - e)) / (2 * h)
        assert abs(fd - grad[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, grad[idx])


def test_fp32_forward_tracks_fp64():
    v = _voxels(4)
    rng = np.random.default_rng(4)
    sdf = rng.normal(0.0, 0.05, v.vertices.shape[0]).astype(np.float32)
    a = orz.compute_sdf_regularizers(sdf, v.neighbors, v.lower_corners, v.upper_corners, (8, 12, 32))
    b = orz.compute_sdf_regularizers_fp64(sdf.astype(np.float64), v.neighbors, v.lower_corners,
                                          v.upper_corners, (8, 12, 32))
    for x, y in zip(a, b):
        assert x.dtype == np.float32
        assert np.allclose(x, y, rtol=1e-4, atol=1e-7 * np.abs(y).max())
