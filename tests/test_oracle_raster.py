"""Pins the rasterization oracle: brute force over all triangles, closed-form
compositing cases, key/sort/delineation invariants, tile-shape invariance."""

import numpy as np
import pytest

from oracle import rasterization as orr
from tests.helpers import OracleScene, small_sphere_spec

F32 = np.float32


@pytest.fixture(scope="module")
def scene():
    return OracleScene(small_sphere_spec(resolution=32, image=48, sh_degree=1))


def test_key_pack_roundtrip_and_signed_order():
    lay = orr.fused_layout()
    cams = np.array([0, 1, 31, 32, 63])
    keys = orr.pack_keys(lay, cams, np.array([1, 2, 3, 4, 5]), np.array([7, 8, 9, 10, 11]),
                         np.array([1.5, 2.0, 0.25, 3.0, 9.0], F32))
    assert np.array_equal(orr.extract_field(lay, keys, "camera"), cams)
    assert np.array_equal(orr.extract_field(lay, keys, "shell"), [1, 2, 3, 4, 5])
    assert np.array_equal(orr.extract_field(lay, keys, "tile"), [7, 8, 9, 10, 11])
    # Cameras >= 32 set bit 63 and sort first in signed order (as CUB's int64 sort).
    sk, sv = orr.sort_keys(keys, np.arange(5, dtype=np.int32))
    assert list(orr.extract_field(lay, sk, "camera")) == [32, 63, 0, 1, 31]


def test_generate_keys_matches_python_loop(scene):
    sc = scene
    f = sc.tri.faces
    want_k, want_v = [], []
    for c in range(sc.B):
        for fi in range(f.shape[0]):
            n = sc.counts[c, fi]
            if n == 0:
                continue
            shell = int(np.searchsorted(sc.tri.face_boundaries, fi, side="right"))
            d = sc.depth[c, f[fi]]
            md = ((d[0] + d[1]) + d[2]) / F32(3)
            for r in range(sc.tmin[c, fi, 0], sc.tmax[c, fi, 0]):
                for q in range(sc.tmin[c, fi, 1], sc.tmax[c, fi, 1]):
                    want_k.append(orr.pack_keys(sc.layout, np.array([c]), np.array([shell]),
                                                np.array([r * sc.cols + q]), np.array([md], F32))[0])
                    want_v.append(fi)
    assert np.array_equal(sc.keys, np.array(want_k, np.int64))
    assert np.array_equal(sc.values, np.array(want_v, np.int32))


def test_delineation_matches_run_scan(scene):
    sc = scene
    cam = orr.extract_field(sc.layout, sc.sorted_keys, "camera")
    tile = orr.extract_field(sc.layout, sc.sorted_keys, "tile")
    want = np.zeros_like(sc.ranges)
    for i in range(sc.sorted_keys.size):
        if i == 0 or (cam[i], tile[i]) != (cam[i - 1], tile[i - 1]):
            want[cam[i], 0, tile[i], 0] = i
        want[cam[i], 0, tile[i], 1] = i + 1
    assert np.array_equal(sc.ranges, want)


def test_barycentrics_agree_with_exact_solve():
    rng = np.random.default_rng(0)
    n = 2000
    p = rng.uniform(0, 50, (n, 3, 2)).astype(F32)
    px = rng.uniform(0, 50, n).astype(F32)
    py = rng.uniform(0, 50, n).astype(F32)
    u, v, w, hit = orr.barycentrics(p[:, 0].T, p[:, 1].T, p[:, 2].T, px, py)
    a = p.astype(np.float64)
    m = np.stack([a[:, 0] - a[:, 2], a[:, 1] - a[:, 2]], axis=-1)  # columns: a - c, b - c
    rhs = np.stack([px - a[:, 2, 0], py - a[:, 2, 1]], axis=-1)
    det = np.linalg.det(m)
    ok = np.abs(det) > 1.0
    sol = np.linalg.solve(m[ok], rhs[ok][..., None])[..., 0]
    eu, ev = sol[:, 0], sol[:, 1]
    ew = 1 - eu - ev
    near = (np.abs(eu) < 2) & (np.abs(ev) < 2) & (np.abs(ew) < 2)
    assert near.sum() > 300
    assert np.abs(u[ok][near] - eu[near]).max() < 1e-3 and np.abs(v[ok][near] - ev[near]).max() < 1e-3
    assert np.abs(w[ok][near] - ew[near]).max() < 1e-3
    inside = (eu > 1e-3) & (ev > 1e-3) & (ew > 1e-3)
    outside = (eu < -1e-3) | (ev < -1e-3) | (ew < -1e-3)
    assert hit[ok][inside].all() and not hit[ok][outside].any()


def test_fused_hits_match_brute_force_over_all_triangles(scene):
    """For every pixel walk *all* of the camera's triangles in global sorted-key order
    (ignoring tiles): the tile delineation and bounds must not drop or add hits."""
    sc = scene
    f = sc.tri.faces
    # Within a tile the traversal order is (depth bits, shell, face): the key order with
    # the tile field removed, ties kept in generation (face) order by the stable sort.
    keep = sc.counts[0] > 0
    fidx = np.nonzero(keep)[0]
    d = sc.depth[0, f[fidx]]
    md = ((d[:, 0] + d[:, 1]) + d[:, 2]) / F32(3)
    shell_of = np.searchsorted(sc.tri.face_boundaries, fidx, side="right")
    order = np.lexsort((fidx, shell_of, md.view(np.uint32)))
    fidx, sh = fidx[order], shell_of[order]
    pa, pb, pc = (sc.xy[0, f[fidx, k]] for k in range(3))
    ys, xs = np.meshgrid(np.arange(sc.H), np.arange(sc.W), indexing="ij")
    px = (xs.reshape(-1).astype(F32) + F32(0.5))[:, None]
    py = (ys.reshape(-1).astype(F32) + F32(0.5))[:, None]
    u, v, w, hit = orr.barycentrics((pa[None, :, 0], pa[None, :, 1]), (pb[None, :, 0], pb[None, :, 1]),
                                    (pc[None, :, 0], pc[None, :, 1]), px, py)
    got = {}
    for pix, face in zip(sc.hits.pixel, sc.hits.face):
        got.setdefault(int(pix), []).append(int(face))
    for p in range(sc.H * sc.W):
        last, seq = -1, []
        for k in np.nonzero(hit[p])[0]:
            if sh[k] > last:
                seq.append(int(fidx[k]))
                last = sh[k]
                if last == sc.S - 1:
                    break
        assert got.get(p, []) == seq, p


def test_tile_shape_invariance():
    imgs = []
    for ts in [(16, 16), (8, 8), (8, 32), (32, 16)]:
        sc = OracleScene(small_sphere_spec(resolution=32, image=48), tile_shape=ts)
        imgs.append(sc.image)
    for im in imgs[1:]:
        assert np.array_equal(im, imgs[0])


def _single_quad_scene(sdf_values, colors, k):
    """S full-screen quads (two triangles each), one per shell, at increasing depth."""
    S = len(sdf_values)
    H = W = 16
    verts, faces = [], []
    for s in range(S):
        z = 1.0 + s
        base = len(verts)
        # With identity extrinsics the corners land at screen top-left, top-right,
        # bottom-right and bottom-left (screen y grows downwards). (TL, TR, BR) and
        # (TL, BR, BL) have check = (b.x - a.x)(c.y - a.y) - (b.y - a.y)(c.x - a.x) > 0,
        # which bound_triangles.slang:95-98 keeps.
        verts += [(-10, -10, z), (10, -10, z), (10, 10, z), (-10, 10, z)]
        faces += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    verts = np.array(verts, F32)
    faces = np.array(faces, np.int32)
    E = np.eye(4, dtype=F32)[None]
    K = np.array([[[8, 0, 8], [0, 8, 8], [0, 0, 1]]], F32)
    xy, d = orr.project_vertices(verts, E, K)
    tmin, tmax, cnt = orr.bound_triangles(xy, d, faces, (16, 16), (1, 1), 0.2)
    assert (cnt == 1).all()
    # The opposite winding is back-facing: every triangle culled.
    _, _, cnt_rev = orr.bound_triangles(xy, d, faces[:, [0, 2, 1]].copy(), (16, 16), (1, 1), 0.2)
    assert (cnt_rev == 0).all()
    bounds = np.arange(2, 2 * S + 1, 2, dtype=np.int32)
    lay = orr.fused_layout()
    keys, vals, _ = orr.generate_keys(d, faces, bounds, tmin, tmax, cnt, 1, lay)
    sk, sv = orr.sort_keys(keys, vals)
    ranges = orr.delineate_tiles(sk, lay, 1, S, 1, False)
    hits = orr.fused_hits(xy, faces, sk, sv, ranges[:, 0], lay, S, (16, 16), (H, W))
    sdf = np.repeat(np.array(sdf_values, F32), 4)
    col = np.repeat(np.array(colors, F32), 4, axis=0)[None]
    img = orr.fused_forward_image(hits, col, sdf, faces, k, 1, (H, W), dtype=np.float64)
    return img, hits


def test_nested_planes_closed_form():
    sdfs = [0.05, 0.0, -0.05]
    cols = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    k = 30.0
    img, hits = _single_quad_scene(sdfs, cols, k)
    sig = lambda x: 1 / (1 + np.exp(-x * k))  # noqa: E731
    T, C, outer = 1.0, np.zeros(3), 1.0
    for s, c in zip(sdfs, cols):
        a = np.clip((outer - sig(s) + 1e-5) / (outer + 1e-5), 0, 1)
        C = C + np.array(c) * T * a
        T = T * (1 - a)
        outer = sig(s)
    assert np.allclose(img[0, :3].reshape(3, -1).T, C, atol=1e-6)
    assert np.allclose(img[0, 3], T, atol=1e-6)
    assert hits.pixel.size == 3 * 16 * 16
