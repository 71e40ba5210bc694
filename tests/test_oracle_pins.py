"""Known-answer pins of the rasterization oracle's front end and explicit-mode adjoint.

Each expected value below is worked out by hand from the reference's definition (file
and line cited per test), not by running the oracle's own formula again:

* ``project_vertices`` (project_vertices.slang:61-69): a rotated, translated camera with
  fx != fy and cx != cy, chosen so that every intermediate is exact in float32. A
  transposed K (no principal point), a transposed E, swapped focal lengths or dividing by
  (K p).z instead of the camera-space z each change at least one expected number.
* ``bound_triangles`` (bound_triangles.slang:62-104): hand-placed triangles on a grid of
  16-row x 8-column tiles: both windings, a degenerate triangle, corners behind and on the
  near plane, negative and out-of-image coordinates, and maxima exactly on a tile edge.
* ``generate_keys`` (generate_keys.slang:4-68, keys.slang:6-22): literal 64-bit keys of
  hand-placed faces in (camera, face, row, col) order, in both key layouts.
* ``interpolate`` / ``interpolate_backward`` (interpolate.slang:19-76): vertex values at
  the barycentric corners, misses = 0, and the adjoint against central differences of
  the forward (it is linear in SDF and colours, so the differences are exact).
* The end-to-end render adjoints of :mod:`oracle.render` against central differences of
  the float64 forward on frozen hits, in both modes, with and without SSAA.
"""

import numpy as np
import pytest

from oracle import rasterization as orr
from oracle import render as ore
from tests.helpers import OracleScene, small_sphere_spec

F32 = np.float32


# --------------------------------------------------------------------------------------
# project_vertices
# --------------------------------------------------------------------------------------


def _camera():
    # World-to-camera: camera = R world + t with the cyclic permutation
    # R = [[0,0,1],[1,0,0],[0,1,0]] (det +1, not symmetric) and t = (0.5, -0.25, 3).
    E = np.array([[0, 0, 1, 0.5], [1, 0, 0, -0.25], [0, 1, 0, 3], [0, 0, 0, 1]], F32)
    # fx = 100, fy = 50, cx = 32, cy = 16.
    K = np.array([[100, 0, 32], [0, 50, 16], [0, 0, 1]], F32)
    return E, K


def test_project_vertices_known_answers():
    E, K = _camera()
    world = np.array([[1, 2, 4], [-1, 0.5, 1], [0, 0, 0]], F32)
    xy, depth = orr.project_vertices(world, E[None], K[None])
    # (1, 2, 4): camera = (4 + 0.5, 1 - 0.25, 2 + 3) = (4.5, 0.75, 5)
    #   -> u = 100 * 4.5 / 5 + 32 = 122, v = 50 * 0.75 / 5 + 16 = 23.5, depth 5.
    # (-1, 0.5, 1): camera = (1.5, -1.25, 3.5)
    #   -> u = 100 * 1.5 / 3.5 + 32 = 74.857142..., v = 50 * -1.25 / 3.5 + 16 = -1.857142...
    # (0, 0, 0): camera = t = (0.5, -0.25, 3) -> u = 100 / 6 + 32, v = -12.5 / 3 + 16.
    want_xy = np.array([[122.0, 23.5], [32 + 150 / 3.5, 16 - 62.5 / 3.5],
                        [32 + 50 / 3.0, 16 - 12.5 / 3.0]])
    want_depth = np.array([5.0, 3.5, 3.0])
    np.testing.assert_allclose(xy[0], want_xy, rtol=2e-7, atol=0)
    np.testing.assert_array_equal(depth[0], want_depth.astype(F32))
    assert xy[0, 0, 0] == F32(122) and xy[0, 0, 1] == F32(23.5)


def test_project_vertices_divides_by_camera_z_not_projected_z():
    # A third intrinsics row other than (0, 0, 1) changes (K p).z but not the reference's
    # xy = (K p).xy / p.z (project_vertices.slang:66-68).
    E, K = _camera()
    K2 = K.copy()
    K2[2] = (0, 0, 2)
    xy, depth = orr.project_vertices(np.array([[1, 2, 4]], F32), E[None], K2[None])
    assert xy[0, 0, 0] == F32(122) and xy[0, 0, 1] == F32(23.5) and depth[0, 0] == F32(5)


def test_project_vertices_general_rotation_against_float64():
    rng = np.random.default_rng(0)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    t = np.array([0.3, -0.2, 4.0])
    E = np.eye(4)
    E[:3, :3], E[:3, 3] = R, t
    E = E.astype(F32)
    K = np.array([[321.5, 0, 200.25], [0, 287.0, 150.75], [0, 0, 1]], F32)
    # World points from chosen camera-space points: world = R^T (camera - t).
    cam = np.stack([rng.uniform(-1, 1, 500), rng.uniform(-1, 1, 500), rng.uniform(2, 6, 500)], 1)
    world = ((cam - t) @ R).astype(F32)
    xy, depth = orr.project_vertices(world, E[None], K[None])
    c64 = world.astype(np.float64) @ E[:3, :3].astype(np.float64).T + E[:3, 3].astype(np.float64)
    want = np.stack([321.5 * c64[:, 0] / c64[:, 2] + 200.25, 287.0 * c64[:, 1] / c64[:, 2] + 150.75], 1)
    np.testing.assert_allclose(xy[0], want, rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(depth[0], c64[:, 2], rtol=1e-6)


# --------------------------------------------------------------------------------------
# bound_triangles
# --------------------------------------------------------------------------------------


def _bound(tris, depths=None, tile_shape=(16, 8), grid=(4, 6), near=0.2):
    """Each triangle gets its own three vertices; one camera."""
    tris = np.asarray(tris, F32)
    n = tris.shape[0]
    xy = tris.reshape(1, 3 * n, 2)
    d = np.ones((1, 3 * n), F32) if depths is None else np.asarray(depths, F32).reshape(1, 3 * n)
    faces = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return orr.bound_triangles(xy, d, faces, tile_shape, grid, near)


def test_bound_triangles_hand_placed():
    # Tiles are 16 rows x 8 columns; the grid is 4 rows x 6 columns (a 64 x 48 image).
    tris = [
        # check = (20 - 3.5) (30 - 2) - (2 - 2) (3.5 - 3.5) = 462 > 0: kept.
        # x in [3.5, 20]: columns int(3.5) / 8 = 0 .. ceil(20 / 8) = 3.
        # y in [2, 30]: rows 2 / 16 = 0 .. ceil(30 / 16) = 2.  2 x 3 = 6 tiles.
        [(3.5, 2), (20, 2), (3.5, 30)],
        # The same corners in the other order: check = -462 < 0, culled.
        [(3.5, 2), (3.5, 30), (20, 2)],
        # Collinear: check = 4 * 8 - 4 * 8 = 0, not < 0, so kept: x, y in [1, 9]:
        # columns 0 .. ceil(9 / 8) = 2, rows 0 .. ceil(9 / 16) = 1.
        [(1, 1), (5, 5), (9, 9)],
        # Negative coordinates: check = 30.5 * 20 = 610 > 0. x in [-20.5, 10]:
        # int(-20.5) = -20, -20 / 8 = -2 (C division) -> clamped to 0; ceil(10 / 8) = 2.
        # y in [-3, 17]: int(-3) / 16 = 0; ceil(17 / 16) = 2.  2 x 2 = 4 tiles.
        [(-20.5, -3), (10, -3), (-20.5, 17)],
        # Maxima exactly on tile edges: x in [8, 16] -> columns 1 .. ceil(16 / 8) = 2;
        # y in [16, 32] -> rows 1 .. ceil(32 / 16) = 2: exactly one tile.
        [(8, 16), (16, 16), (8, 32)],
        # Just past the edges: x max 16.5 -> ceil(2.0625) = 3; y max 32.5 -> 3: 2 x 2.
        [(8, 16), (16.5, 16), (8, 32.5)],
        # Entirely right of and below the image: minima 100 / 8 = 12 -> 6 and
        # 200 / 16 = 12 -> 4, maxima clamped to 6 and 4: empty ranges, 0 tiles.
        [(100, 200), (110, 200), (100, 210)],
        # Entirely left of the image: x max -3 -> ceil(-0.375) = 0: columns [0, 0).
        [(-10, 5), (-3, 5), (-10, 9)],
    ]
    tmin, tmax, cnt = _bound(tris)
    want_min = [(0, 0), (0, 0), (0, 0), (0, 0), (1, 1), (1, 1), (4, 6), (0, 0)]
    want_max = [(2, 3), (0, 0), (1, 2), (2, 2), (2, 2), (3, 3), (4, 6), (1, 0)]
    want_cnt = [6, 0, 2, 4, 1, 4, 0, 0]
    np.testing.assert_array_equal(tmin[0], want_min)
    np.testing.assert_array_equal(tmax[0], want_max)
    np.testing.assert_array_equal(cnt[0], want_cnt)


def test_bound_triangles_near_plane():
    tri = [(3.5, 2), (20, 2), (3.5, 30)]
    # One corner closer than the near plane (0.1 < 0.2): culled. A corner exactly on it
    # (min_depth < near_plane is false): kept. Behind the camera (negative depth): culled.
    tmin, tmax, cnt = _bound([tri, tri, tri], depths=[[1, 0.1, 1], [0.2, 1, 1], [-1, 2, 2]])
    np.testing.assert_array_equal(cnt[0], [0, 6, 0])
    np.testing.assert_array_equal(tmax[0], [(0, 0), (2, 3), (0, 0)])


def test_bound_triangles_tile_axes_not_swapped():
    # With 8-row x 16-column tiles the same triangle covers rows 0 .. ceil(30 / 8) = 4 and
    # columns 0 .. ceil(20 / 16) = 2.
    tmin, tmax, cnt = _bound([[(3.5, 2), (20, 2), (3.5, 30)]], tile_shape=(8, 16), grid=(8, 3))
    np.testing.assert_array_equal(tmax[0], [(4, 2)])
    np.testing.assert_array_equal(cnt[0], [8])


# --------------------------------------------------------------------------------------
# generate_keys
# --------------------------------------------------------------------------------------


def test_generate_keys_known_answers():
    # Two cameras, three faces; shell boundaries (cumulative face ends) [1, 3]: face 0 is
    # in shell 0, face 1 (== boundary 0) and face 2 in shell 1 (face < boundary, slang:25-30).
    faces = np.arange(9, dtype=np.int32).reshape(3, 3)
    depths = np.array([[1, 2, 3, 1, 1, 2, 4, 4, 4],
                       [2, 2, 2, 0.5, 0.5, 0.5, 8, 8, 8]], F32)
    tmin = np.array([[[0, 1], [2, 0], [0, 0]], [[1, 1], [0, 0], [0, 0]]], np.int32)
    tmax = np.array([[[1, 3], [3, 1], [0, 0]], [[2, 2], [0, 0], [0, 0]]], np.int32)
    counts = np.array([[2, 1, 0], [1, 0, 0]], np.int32)
    grid_cols = 5
    bounds = np.array([1, 3], np.int32)
    # Mean depths: (1 + 2 + 3) / 3 = 2.0 -> 0x40000000; (1 + 1 + 2) / 3 = 4/3 rounded to
    # float32 -> 0x3FAAAAAB; camera 1, face 0: 2.0.
    d_two, d_four_thirds = 0x40000000, 0x3FAAAAAB
    # Fused layout, MSB -> LSB camera 6 | tile 20 | depth 32 | shell 6.
    fused = [
        (0 << 58) | (1 << 38) | (d_two << 6) | 0,        # camera 0, face 0, tile (0, 1) = 1
        (0 << 58) | (2 << 38) | (d_two << 6) | 0,        # tile (0, 2) = 2
        (0 << 58) | (10 << 38) | (d_four_thirds << 6) | 1,  # face 1, tile (2, 0) = 10
        (1 << 58) | (6 << 38) | (d_two << 6) | 0,        # camera 1, face 0, tile (1, 1) = 6
    ]
    keys, values, _ = orr.generate_keys(depths, faces, bounds, tmin, tmax, counts, grid_cols,
                                        orr.fused_layout())
    np.testing.assert_array_equal(keys.view(np.uint64), np.array(fused, np.uint64))
    np.testing.assert_array_equal(values, [0, 0, 1, 0])
    # Explicit layout, MSB -> LSB camera 6 | shell 6 | tile 20 | depth 32.
    explicit = [
        (0 << 58) | (0 << 52) | (1 << 32) | d_two,
        (0 << 58) | (0 << 52) | (2 << 32) | d_two,
        (0 << 58) | (1 << 52) | (10 << 32) | d_four_thirds,
        (1 << 58) | (0 << 52) | (6 << 32) | d_two,
    ]
    keys, values, _ = orr.generate_keys(depths, faces, bounds, tmin, tmax, counts, grid_cols,
                                        orr.explicit_layout())
    np.testing.assert_array_equal(keys.view(np.uint64), np.array(explicit, np.uint64))


# --------------------------------------------------------------------------------------
# interpolate and its adjoint
# --------------------------------------------------------------------------------------


def _interp_case(seed=0, b=2, s=3, h=5, w=4, v=12, f=7):
    rng = np.random.default_rng(seed)
    faces = rng.integers(0, v, (f, 3)).astype(np.int32)
    index = rng.integers(-1, f, (b, s, h, w)).astype(np.int32)
    uv = rng.uniform(0, 1, (b, s, h, w, 2)).astype(F32)
    sdf = rng.standard_normal(v).astype(F32)
    colors = rng.random((b, v, 3)).astype(F32)
    return uv, index, faces, sdf, colors


def test_interpolate_known_answers():
    faces = np.array([[0, 1, 2]], np.int32)
    sdf = np.array([10.0, 20.0, 40.0], F32)
    colors = np.array([[[1, 0, 0], [0, 1, 0], [0, 0, 1]]], F32)
    # (u, v) = (1, 0) -> vertex a; (0, 1) -> b; (0, 0) -> c; (0.25, 0.5) ->
    # 0.25 * 10 + 0.5 * 20 + 0.25 * 40 = 22.5; index -1 -> 0.
    uv = np.array([[[[[1, 0], [0, 1], [0, 0], [0.25, 0.5], [0.3, 0.3]]]]], F32)
    index = np.array([[[[0, 0, 0, 0, -1]]]], np.int32)
    s_img, c_img = orr.interpolate(uv, index, faces, sdf, colors)
    np.testing.assert_array_equal(s_img[0, 0, 0], [10, 20, 40, 22.5, 0])
    np.testing.assert_array_equal(c_img[0, 0, 0], [[1, 0, 0], [0, 1, 0], [0, 0, 1],
                                                  [0.25, 0.5, 0.25], [0, 0, 0]])


def test_interpolate_backward_matches_finite_differences():
    uv, index, faces, sdf, colors = _interp_case()
    rng = np.random.default_rng(1)
    gs = rng.standard_normal(index.shape)
    gc = rng.standard_normal(index.shape + (3,))
    sg, cg = orr.interpolate_backward(uv, index, faces, sdf.shape[0], colors.shape[0], gs, gc)

    def loss(sd, co):
        a, c = orr.interpolate(uv, index, faces, sd, co, dtype=np.float64)
        return float((a * gs).sum() + (c * gc).sum())

    h = 0.5  # the loss is linear in sdf and colours: any step gives the exact derivative
    s64, c64 = sdf.astype(np.float64), colors.astype(np.float64)
    for i in range(sdf.shape[0]):
        e = np.zeros_like(s64)
        e[i] = h
        fd = (loss(s64 + e, c64) - loss(s64 - e, c64)) / (2 * h)
        assert abs(fd - sg[i]) <= 1e-9 * max(1.0, abs(fd))
    for bi in range(colors.shape[0]):
        for i in range(colors.shape[1]):
            for ch in range(3):
                e = np.zeros_like(c64)
                e[bi, i, ch] = h
                fd = (loss(s64, c64 + e) - loss(s64, c64 - e)) / (2 * h)
                assert abs(fd - cg[bi, i, ch]) <= 1e-9 * max(1.0, abs(fd))


# --------------------------------------------------------------------------------------
# End-to-end render adjoints (oracle/render.py)
# --------------------------------------------------------------------------------------


@pytest.fixture(scope="module")
def small():
    return OracleScene(small_sphere_spec(resolution=32, image=20, sh_degree=1))


@pytest.mark.parametrize("mode", ["fused", "explicit"])
@pytest.mark.parametrize("ssaa", [1, 2])
def test_render_adjoint_matches_finite_differences(small, mode, ssaa):
    sc = small
    t = sc.tri
    bg = np.array([0.2, 0.5, 0.9], F32)
    args = (sc.E, sc.K, (sc.H, sc.W), t.vertices, t.faces, t.face_boundaries)
    img, st = ore.render(*args, t.signed_distances, t.spherical_harmonics, 1, sc.spec.sharpness,
                         bg, ssaa=ssaa, mode=mode)
    assert img.shape == (sc.B, 3, sc.H, sc.W)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(img.shape)
    sg, shg = ore.render_backward(st, t.vertices, t.faces, t.signed_distances, t.spherical_harmonics,
                                  sc.E, 1, sc.spec.sharpness, bg, g)
    k = sc.spec.sharpness

    def loss(sdf, sh):
        """The same image in float64 on the forward's frozen hits."""
        col = orr.evaluate_spherical_harmonics(t.vertices, sh, sc.E, 1)[0]
        hh, ww = st.image_shape
        if mode == "fused":
            raw = orr.fused_forward_image(st.hits, col, sdf, t.faces, k, sc.B, (hh, ww),
                                          dtype=np.float64)
            im = ore.ssaa_mean(raw[:, :3] + raw[:, 3:4] * bg[None, :, None, None], ssaa)
            im = np.where((st.pre_clip >= 0) & (st.pre_clip <= 1), im, 0.0)
        else:
            s_img, c_img = orr.interpolate(st.uv, st.index, t.faces, sdf, col, dtype=np.float64)
            s = st.index.shape[1]
            cl, tr = orr.composite(s_img.transpose(0, 2, 3, 1).reshape(-1, s),
                                   c_img.transpose(0, 2, 3, 1, 4).reshape(-1, s, 3),
                                   st.planes[2], k, dtype=np.float64)
            flat = cl + bg[None] * tr[:, None]
            im = ore.ssaa_mean(flat.reshape(sc.B, hh, ww, 3).transpose(0, 3, 1, 2), ssaa)
        return float((im * g).sum())

    s64 = t.signed_distances.astype(np.float64)
    h64 = t.spherical_harmonics.astype(np.float64)
    hit_faces = st.hits.face if mode == "fused" else st.index[st.index >= 0]
    used = np.unique(t.faces[hit_faces])
    # Tolerance 1e-3 relative (the gradient bar): the suffix trick divides by
    # 1 - alpha + 1e-6 (intersect_and_composite.slang:3, composite.slang) and the adjoint
    # reads the FP32 forward image, both ~1e-4 relative where alpha is close to 1.
    h = 1e-6
    checked = 0
    for i in rng.choice(used, 12, replace=False):
        e = np.zeros_like(s64)
        e[i] = h
        fd = (loss(s64 + e, h64) - loss(s64 - e, h64)) / (2 * h)
        assert abs(fd - sg[i]) <= 1e-3 * max(abs(fd), 1e-2), (i, fd, sg[i])
        checked += abs(fd) > 1e-6
    for i in rng.choice(used, 12, replace=False):
        c, ch = rng.integers(0, 4), rng.integers(0, 3)
        e = np.zeros_like(h64)
        e[c, i, ch] = h
        fd = (loss(s64, h64 + e) - loss(s64, h64 - e)) / (2 * h)
        assert abs(fd - shg[c, i, ch]) <= 1e-3 * max(abs(fd), 1e-2), (c, i, ch, fd, shg[c, i, ch])
        checked += abs(fd) > 1e-6
    assert checked >= 6


def test_render_ssaa_intrinsics_and_mean():
    K = np.array([[[100, 0, 32], [0, 50, 16], [0, 0, 1]]], F32)
    np.testing.assert_array_equal(ore.ssaa_intrinsics(K, 2)[0], [[200, 0, 64], [0, 100, 32], [0, 0, 1]])
    im = np.arange(16, dtype=F32).reshape(1, 1, 4, 4)
    np.testing.assert_array_equal(ore.ssaa_mean(im, 2)[0, 0], [[2.5, 4.5], [10.5, 12.5]])


def test_fused_empty_scene_is_background():
    E = np.eye(4, dtype=F32)[None]
    K = np.array([[[8, 0, 4], [0, 8, 4], [0, 0, 1]]], F32)
    img, _ = ore.render(E, K, (6, 5), np.zeros((0, 3), F32), np.zeros((0, 3), np.int32),
                        np.zeros(2, np.int32), np.zeros(0, F32), np.zeros((1, 0, 3), F32), 0, 50.0,
                        np.array([0.1, 0.2, 0.3], F32))
    assert img.shape == (1, 3, 6, 5)
    np.testing.assert_array_equal(img[0, :, 3, 2], np.array([0.1, 0.2, 0.3], F32))


# --------------------------------------------------------------------------------------
# Explicit first-hit traversal and explicit compositing
# --------------------------------------------------------------------------------------


def test_explicit_intersect_matches_brute_force(small):
    """Per (shell, pixel), over *all* of the shell's triangles (no tiles): the first
    triangle that passes the barycentric test in (mean-depth bits, face) order, which is
    the explicit key order within one (camera, shell, tile) (intersect.slang:85-182)."""
    sc = small
    t = sc.tri
    uv, index = ore.rasterize(t.vertices, t.faces, t.face_boundaries, sc.E, sc.K, (sc.H, sc.W),
                              (8, 8))
    f = t.faces
    d = sc.depth[0, f]
    md = ((d[:, 0] + d[:, 1]) + d[:, 2]) / F32(3)
    shell = np.searchsorted(t.face_boundaries, np.arange(f.shape[0]), side="right")
    keep = (sc.counts[0] > 0)
    ys, xs = np.meshgrid(np.arange(sc.H), np.arange(sc.W), indexing="ij")
    px = (xs.reshape(-1).astype(F32) + F32(0.5))[:, None]
    py = (ys.reshape(-1).astype(F32) + F32(0.5))[:, None]
    for s in range(sc.S):
        fidx = np.nonzero(keep & (shell == s))[0]
        fidx = fidx[np.lexsort((fidx, md[fidx].view(np.uint32)))]
        pa, pb, pc = (sc.xy[0, f[fidx, k]] for k in range(3))
        u, v, _, hit = orr.barycentrics((pa[None, :, 0], pa[None, :, 1]), (pb[None, :, 0], pb[None, :, 1]),
                                        (pc[None, :, 0], pc[None, :, 1]), px, py)
        first = np.argmax(hit, axis=1)
        any_hit = hit.any(axis=1)
        want = np.where(any_hit, fidx[first], -1).reshape(sc.H, sc.W)
        np.testing.assert_array_equal(index[0, s], want)
        pix = np.nonzero(any_hit)[0]
        np.testing.assert_array_equal(uv[0, s].reshape(-1, 2)[pix, 0], u[pix, first[pix]])
        np.testing.assert_array_equal(uv[0, s].reshape(-1, 2)[pix, 1], v[pix, first[pix]])
    assert (index >= 0).sum() > 100


def test_explicit_and_fused_agree_on_nested_planes():
    """For well-nested shells the two modes composite the same samples in the same order
    (README.md:57; they differ otherwise, Q3), through independently written code paths:
    the per-layer composite and the hit-list traversal."""
    S, H, W = 3, 16, 16
    verts, faces = [], []
    for s in range(S):
        z = 1.0 + s
        base = len(verts)
        verts += [(-10, -10, z), (10, -10, z), (10, 10, z), (-10, 10, z)]
        faces += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    verts = np.array(verts, F32)
    faces = np.array(faces, np.int32)
    bounds = np.array([2, 4, 6], np.int32)
    sdf = np.repeat(np.array([0.04, 0.0, -0.04], F32), 4)
    sh = np.repeat(np.array([[0.9, -0.5, 0.1], [-0.3, 0.6, 0.2], [0.1, 0.2, -0.8]], F32), 4, axis=0)[None]
    E = np.eye(4, dtype=F32)[None]
    K = np.array([[[8, 0, 8], [0, 8, 8], [0, 0, 1]]], F32)
    bg = np.array([0.3, 0.3, 0.3], F32)
    fused, _ = ore.render(E, K, (H, W), verts, faces, bounds, sdf, sh, 0, 40.0, bg, mode="fused")
    explicit, st = ore.render(E, K, (H, W), verts, faces, bounds, sdf, sh, 0, 40.0, bg, mode="explicit")
    assert (st.index >= 0).all()
    assert 0.05 < explicit.min() and explicit.max() < 0.95
    np.testing.assert_allclose(explicit, fused, atol=2e-6)
