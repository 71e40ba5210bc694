"""Pins the marching-cubes oracle with closed forms and mesh invariants."""

import numpy as np
import torch

from oracle import extraction as ox

F32 = np.float32


def _dense_grid(n, k=32):
    """Fully occupied n x n x k grid (k, the packed axis, a multiple of 32)."""
    packed = ox.pack_occupancy(np.ones((n, n, k), dtype=bool))
    v = ox.extract_voxels(packed)
    return v, (2 * v.vertices - 1).astype(F32)


def _tessellate(v, world, sdf, sh, level_sets):
    return ox.tessellate_voxels(sdf, world, sh, v.neighbors, v.lower_corners, v.upper_corners,
                                v.indices, np.asarray(level_sets, F32))


def test_exclusive_cumsum_closed_form():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 100, 1000).astype(np.int32)
    out = ox.exclusive_cumsum(x)
    assert out[0] == 0 and np.array_equal(out[1:], np.cumsum(x[:-1]))


def test_sphere_meshes_are_watertight_and_consistently_oriented():
    v, world = _dense_grid(32)
    sdf = (np.linalg.norm(world, axis=1) - 0.55).astype(F32)  # closed, inside the grid
    sh = np.zeros((1, world.shape[0], 3), F32)
    ls = [0.1, 0.0, -0.1]
    tri = _tessellate(v, world, sdf, sh, ls)
    assert (tri.faces >= 0).all()
    starts = np.concatenate([[0], tri.face_boundaries[:-1]])
    for s0, s1 in zip(starts, tri.face_boundaries):
        f = tri.faces[s0:s1].astype(np.int64)
        # Every directed edge appears once and its reverse once: closed, oriented manifold.
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = e[:, 0] * 10**7 + e[:, 1]
        rev = e[:, 1] * 10**7 + e[:, 0]
        assert np.unique(key).size == key.size
        assert np.array_equal(np.sort(key), np.sort(rev))


def test_vertices_lie_on_level_sets_and_grid_edges():
    v, world = _dense_grid(16)
    rng = np.random.default_rng(1)
    sdf = (np.linalg.norm(world, axis=1) - 0.5 + 0.05 * rng.standard_normal(world.shape[0])).astype(F32)
    sh = rng.random((4, world.shape[0], 3)).astype(F32)
    ls = np.array([0.05, 0.0, -0.05], F32)
    tri = _tessellate(v, world, sdf, sh, ls)
    vb = np.concatenate([[0], tri.vertex_boundaries])
    for s in range(3):
        d = tri.signed_distances[vb[s] : vb[s + 1]].astype(np.float64)
        assert np.abs(d - ls[s]).max() < 1e-6
    rec = tri.record
    p0 = world[rec.i0].astype(np.float64)
    p1 = world[rec.i1].astype(np.float64)
    t = rec.t.astype(np.float64)
    np.testing.assert_allclose(tri.vertices, p0 + t[:, None] * (p1 - p0), atol=1e-6)
    # Interpolated SH uses the same weight.
    np.testing.assert_allclose(tri.spherical_harmonics,
                               sh[:, rec.i0] + t[None, :, None] * (sh[:, rec.i1] - sh[:, rec.i0]),
                               atol=1e-6)
    # Exactly one coordinate differs from the column origin (the edge axis).
    diff = np.abs(tri.vertices - world[rec.i0]) > 0
    assert (diff.sum(axis=1) <= 1).all()
    assert np.array_equal(tri.vertex_types, rec.axis.astype(np.uint8))


def test_plane_gives_two_triangles_per_crossed_cell():
    n = 16
    v, world = _dense_grid(n)
    # Plane z = 0.03 (not on a lattice plane): crosses one layer of n * n cells.
    sdf = (world[:, 2] - 0.03).astype(F32)
    sh = np.zeros((1, world.shape[0], 3), F32)
    tri = _tessellate(v, world, sdf, sh, [0.0])
    assert tri.faces.shape[0] == 2 * n * 32
    assert tri.vertices.shape[0] == (n + 1) * (32 + 1)
    assert (tri.faces >= 0).all()


def test_cube_codes_closed_form():
    v, world = _dense_grid(8)
    rng = np.random.default_rng(2)
    sdf = rng.standard_normal(world.shape[0]).astype(F32)
    counts, codes = ox.count_triangle_faces(sdf, v.lower_corners, v.upper_corners, np.array([0.1], F32))
    # Corner (dx, dy, dz) of the voxel at lattice position p is vertex p + (dx, dy, dz).
    n_vox = v.num_voxels
    pos = np.rint(v.vertices * np.array([32, 8, 8])).astype(np.int64)
    lut = {tuple(p): i for i, p in enumerate(pos)}
    o = pos[v.lower_corners[0, :n_vox]]
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    want = np.zeros(n_vox, np.int64)
    for bit, off in enumerate(order):
        idx = np.array([lut[tuple(p + off)] for p in o])
        want |= (sdf[idx] >= F32(0.1)).astype(np.int64) << bit
    np.testing.assert_array_equal(codes[0], want.astype(np.uint8))


def test_vertex_adjoint_matches_torch_autograd():
    v, world = _dense_grid(8)
    rng = np.random.default_rng(3)
    sdf = (np.linalg.norm(world, axis=1) - 0.5 + 0.05 * rng.standard_normal(world.shape[0])).astype(F32)
    sh = rng.random((4, world.shape[0], 3)).astype(F32)
    tri = _tessellate(v, world, sdf, sh, [0.02, -0.02])
    rec = tri.record
    gv = rng.standard_normal(tri.vertices.shape)
    gd = rng.standard_normal(tri.signed_distances.shape)
    gs = rng.standard_normal(tri.spherical_harmonics.shape)
    W = torch.tensor(world, dtype=torch.float64, requires_grad=True)
    D = torch.tensor(sdf, dtype=torch.float64, requires_grad=True)
    SH = torch.tensor(sh, dtype=torch.float64, requires_grad=True)
    t = torch.tensor(rec.t, dtype=torch.float64)  # detached, as in the reference
    i0, i1 = torch.tensor(rec.i0), torch.tensor(rec.i1)
    ov = W[i0] + t[:, None] * (W[i1] - W[i0])
    od = D[i0] + t * (D[i1] - D[i0])
    osh = SH[:, i0] + t[None, :, None] * (SH[:, i1] - SH[:, i0])
    loss = (ov * torch.tensor(gv)).sum() + (od * torch.tensor(gd)).sum() + (osh * torch.tensor(gs)).sum()
    loss.backward()
    wv, wd, ws = ox.create_triangle_vertices_backward(rec, world.shape[0], 4, gv, gd, gs)
    np.testing.assert_allclose(wv, W.grad.numpy(), atol=1e-10)
    np.testing.assert_allclose(wd, D.grad.numpy(), atol=1e-10)
    np.testing.assert_allclose(ws, SH.grad.numpy(), atol=1e-10)
