"""GPU parity of the rasterization kernels against the CPU oracle.

Bit-exact: projected positions and depths, tile bounds, keys, sorted keys and values,
tile ranges, explicit hit indices and barycentrics. Images: within 1e-4 max-abs of the
oracle (FP32 and FP64-continuous). Gradients: within 1e-3 relative of the oracle adjoint.
"""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import _native, config, workloads
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import triangle_rasterization as tr
from meshtryoshka_amd.triangle_rasterization.keys import create_key_defines, to_layout
from meshtryoshka_amd.triangle_rasterization.common import TileGrid
from oracle import rasterization as orr
from tests.helpers import OracleScene, small_sphere_spec, to_dev

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3


@pytest.fixture(scope="module")
def scene():
    return OracleScene(small_sphere_spec(resolution=32, image=64, sh_degree=2))


@pytest.fixture(scope="module")
def scene_explicit():
    return OracleScene(small_sphere_spec(resolution=32, image=64, sh_degree=1), explicit=True)


def _grid(sc):
    return TileGrid(sc.tile_shape, (sc.H, sc.W), sc.B, sc.S)


def _key_defines(explicit=False):
    if explicit:
        return tr.create_key_defines(("camera", 6), ("shell", 6), ("tile", 20), ("depth", 32))
    return tr.create_key_defines(("camera", 6), ("tile", 20), ("depth", 32), ("shell", 6))


def _front_end(sc, explicit=False):
    t = sc.tri
    verts = to_dev(t.vertices)
    faces = to_dev(t.faces)
    xy, depth = tr.project_vertices(verts, to_dev(sc.E), to_dev(sc.K))
    b = tr.bound_triangles(xy, depth, faces, _grid(sc), sc.spec.near_plane)
    kd = _key_defines(explicit)
    pk = tr.generate_keys(depth, faces, to_dev(t.face_boundaries), b.tile_minima, b.tile_maxima,
                          b.num_tiles_overlapped, _grid(sc), kd)
    spk = tr.sort_keys(pk, _grid(sc))
    ranges = tr.delineate_tiles(spk, _grid(sc), kd, explicit)
    return xy, depth, b, pk, spk, ranges


@pytest.mark.parametrize("v_count,live,shift_in,shift_out", [
    (1001, None, 0, 0), (1001, 777, 1, 0), (1000, 998, 0, 1), (3, 2, 1, 1), (4096, 4096, 0, 0)])
def test_projection_live_and_unaligned(scene, v_count, live, shift_in, shift_out):
    """The projection kernel's 16-byte paths and their fallbacks: vertex counts that are not
    a multiple of four, vertex / output arrays that start off a 16-byte boundary (views one
    row / one entry into a buffer), odd camera offsets (3 cameras, odd V), and the live-count
    entry (entries past the live count stay untouched). Bit-exact against the oracle."""
    rng = np.random.default_rng(v_count + 7 * (live or 0))
    E = np.tile(scene.E[:1], (3, 1, 1)).astype(np.float32)
    K = np.tile(scene.K[:1], (3, 1, 1)).astype(np.float32)
    E[1, :3, 3] += np.float32([0.25, -0.5, 0.125])
    K[2, 0, 0] *= np.float32(1.5)
    verts = rng.uniform(-1.0, 1.0, (v_count, 3)).astype(np.float32)
    want_xy, want_d = orr.project_vertices(verts, E, K)
    B = E.shape[0]
    vbuf = torch.zeros((v_count + 1, 3), device="cuda")
    vbuf[shift_in:shift_in + v_count] = to_dev(verts)
    vv = vbuf[shift_in:shift_in + v_count]
    xbuf = torch.full((B * v_count * 2 + 2,), -7.0, device="cuda")
    dbuf = torch.full((B * v_count + 1,), -7.0, device="cuda")
    xy = xbuf[2 * shift_out:2 * shift_out + B * v_count * 2]
    d = dbuf[shift_out:shift_out + B * v_count]
    st = torch.cuda.current_stream().cuda_stream
    E_d, K_d = to_dev(E), to_dev(K)  # kept alive until the kernel has read them
    if live is None:
        _native.call("mt_project_vertices", vv.data_ptr(), v_count, E_d.data_ptr(), K_d.data_ptr(),
                     B, xy.data_ptr(), d.data_ptr(), st)
        n = v_count
    else:
        nl = torch.tensor([live], dtype=torch.int32, device="cuda")
        _native.call("mt_project_vertices_live", vv.data_ptr(), v_count, nl.data_ptr(),
                     E_d.data_ptr(), K_d.data_ptr(), B, xy.data_ptr(), d.data_ptr(), st)
        n = live
    got_xy = xy.view(B, v_count, 2).cpu().numpy()
    got_d = d.view(B, v_count).cpu().numpy()
    assert np.array_equal(got_xy[:, :n].view(np.uint32), want_xy[:, :n].view(np.uint32))
    assert np.array_equal(got_d[:, :n].view(np.uint32), want_d[:, :n].view(np.uint32))
    assert (got_xy[:, n:] == -7.0).all() and (got_d[:, n:] == -7.0).all()


def test_front_end_bit_exact(scene):
    xy, depth, b, pk, spk, ranges = _front_end(scene)
    assert np.array_equal(xy.cpu().numpy().view(np.uint32), scene.xy.view(np.uint32))
    assert np.array_equal(depth.cpu().numpy().view(np.uint32), scene.depth.view(np.uint32))
    np.testing.assert_array_equal(b.tile_minima.cpu().numpy(), scene.tmin)
    np.testing.assert_array_equal(b.tile_maxima.cpu().numpy(), scene.tmax)
    np.testing.assert_array_equal(b.num_tiles_overlapped.cpu().numpy(), scene.counts)
    np.testing.assert_array_equal(pk.keys.cpu().numpy(), scene.keys)
    np.testing.assert_array_equal(pk.triangle_indices.cpu().numpy(), scene.values)
    np.testing.assert_array_equal(spk.keys.cpu().numpy(), scene.sorted_keys)
    np.testing.assert_array_equal(spk.triangle_indices.cpu().numpy(), scene.sorted_values)
    np.testing.assert_array_equal(ranges.cpu().numpy(), scene.ranges)


def test_keys_from_geometry_match_oracle(scene):
    """The batched step's key path (block counts -> scan -> keys recomputing each face's
    tile rectangle) writes exactly the oracle's keys and values; the device face count
    cuts the faces off like mt_bound_triangles does."""
    from meshtryoshka_amd import _native
    from meshtryoshka_amd.triangle_rasterization.keys import to_layout

    t = scene.tri
    xy, depth = to_dev(scene.xy), to_dev(scene.depth)
    faces = to_dev(t.faces)
    v, f = xy.shape[1], faces.shape[0]
    b = scene.B
    th, tw = scene.tile_shape
    st = torch.cuda.current_stream().cuda_stream
    nblk = -(-f // 256)
    layout = to_layout(_key_defines())
    for live in (f, f // 2):
        live_x3 = torch.tensor([3 * live], dtype=torch.int64, device="cuda")
        counts = torch.empty(b * nblk, dtype=torch.int32, device="cuda")
        _native.call("mt_key_block_counts", xy.data_ptr(), depth.data_ptr(), v, faces.data_ptr(), f,
                     live_x3.data_ptr(), b, th, tw, 0, 0, scene.rows, scene.cols,
                     float(scene.spec.near_plane), None, 0, counts.data_ptr(), st)
        offs = torch.empty(b * nblk, dtype=torch.int64, device="cuda")
        ws = _native.workspace(_native.query("mt_scan_workspace_bytes", b * nblk), xy.device)
        _native.call("mt_inclusive_scan_i32_to_i64", counts.data_ptr(), offs.data_ptr(), b * nblk,
                     ws.data_ptr(), ws.numel(), st)
        cnt = scene.counts.reshape(b, f).copy()
        cnt[:, live:] = 0
        want_n = int(cnt.sum())
        assert int(offs[-1]) == want_n
        keys = torch.full((want_n + 7,), -1, dtype=torch.int64, device="cuda")
        vals = torch.full((want_n + 7,), -1, dtype=torch.int32, device="cuda")
        over = torch.zeros(1, dtype=torch.int32, device="cuda")
        fb = to_dev(t.face_boundaries)  # kept alive until the kernel has read it
        _native.call("mt_generate_keys_from_geometry", xy.data_ptr(), depth.data_ptr(), v,
                     faces.data_ptr(), f, live_x3.data_ptr(), fb.data_ptr(),
                     scene.S, b, th, tw, 0, 0, scene.rows, scene.cols, float(scene.spec.near_plane),
                     None, 0, offs.data_ptr(), scene.cols, layout, keys.data_ptr(), vals.data_ptr(),
                     keys.numel(), over.data_ptr(), st)
        assert int(over) == 0
        # Oracle keys of the live faces, in (camera, face, tile) order.
        alive = np.broadcast_to(np.arange(f)[None, :] < live, (b, f)).reshape(-1)
        keep = np.repeat(alive, scene.counts.reshape(-1))
        np.testing.assert_array_equal(keys[:want_n].cpu().numpy(), scene.keys[keep])
        np.testing.assert_array_equal(vals[:want_n].cpu().numpy(), scene.values[keep])
        assert bool((keys[want_n:] == -1).all())


@pytest.mark.parametrize("n", [1, 2, 100, 4096, 4097, 65_537, 1_000_003])
def test_sort_matches_stable_argsort(n):
    rng = np.random.default_rng(n)
    # Few distinct keys (many ties), both signs (camera >= 32 sets bit 63), varying digits.
    base = rng.integers(-(2**62), 2**62, size=max(1, n // 7), dtype=np.int64)
    keys = base[rng.integers(0, base.size, size=n)]
    keys[rng.random(n) < 0.1] |= np.int64(-(2**63))
    values = np.arange(n, dtype=np.int32)
    out = tr.sort_keys(tr.PairedKeys(to_dev(keys), to_dev(values)), None)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(out.keys.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(out.triangle_indices.cpu().numpy(), values[order])


def test_sort_all_equal_and_structured_keys():
    n = 50_000
    keys = np.full(n, 123456789, dtype=np.int64)
    values = np.arange(n, dtype=np.int32)[::-1].copy()
    out = tr.sort_keys(tr.PairedKeys(to_dev(keys), to_dev(values)), None)
    np.testing.assert_array_equal(out.triangle_indices.cpu().numpy(), values)
    # Realistic fused keys: only a few digit positions vary.
    rng = np.random.default_rng(3)
    lay = orr.fused_layout()
    depth = rng.uniform(0.5, 4.0, n).astype(np.float32)
    keys = orr.pack_keys(lay, rng.integers(0, 4, n), rng.integers(0, 5, n), rng.integers(0, 4056, n), depth)
    out = tr.sort_keys(tr.PairedKeys(to_dev(keys), to_dev(values)), None)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(out.keys.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(out.triangle_indices.cpu().numpy(), values[order])


@pytest.mark.parametrize("mask", [0x5555_5555_5555_5555, -0x5555_5555_5555_5556,
                                  (1 << 0) | (1 << 9) | (1 << 33) | (1 << 62), -(2**63) | 0xF0])
def test_sort_scattered_varying_bits(mask):
    """Digits are taken from the varying key bits only; scattered and alternating bit
    patterns (up to 32 separate runs, with and without the sign bit) keep exact order."""
    n = 300_001
    rng = np.random.default_rng(abs(mask) % 1000)
    keys = rng.integers(-(2**63), 2**63 - 1, size=n, dtype=np.int64) & np.int64(mask)
    values = rng.permutation(n).astype(np.int32)
    out = tr.sort_keys(tr.PairedKeys(to_dev(keys), to_dev(values)), None)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(out.keys.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(out.triangle_indices.cpu().numpy(), values[order])


@pytest.mark.parametrize("depths", ["few", "one", "uniform"])
def test_tile_bucketed_sort_equals_full_sort(depths):
    """mt_sort_keys_high over the (camera, tile) bits, then mt_sort_key_ranges over each
    tile's range by the (depth, shell) bits, gives exactly the stable sort of the whole
    keys (mt_sort_keys' order). Ranges of every size class: up to 1024, 2048, 4096, 8192
    and 16384 keys (one workgroup of 64, 128, 256, 512 or 1024 threads), longer (through the
    temporary arrays), each class's bounds on both sides,
    with both signs of the key, ties, one, few or many varying low bits (odd and even pass
    counts), empty ranges and slots that hold no keys."""
    from meshtryoshka_amd import _native

    rng = np.random.default_rng({"few": 1, "one": 2, "uniform": 3}[depths])
    lay = orr.fused_layout()
    sizes = [1, 2, 5, 300, 1024, 1025, 2048, 2049, 4095, 4096, 4097, 8192, 8193, 9000, 16383, 16384, 16385,
             40000, 70001]
    sizes += [int(x) for x in rng.integers(1, 3000, 120)]
    cam, tile = [], []
    for i, m in enumerate(sizes):
        cam.append(np.full(m, (i * 13) % 64))  # cameras >= 32 set the key's sign bit
        tile.append(np.full(m, 1 + i * 7))
    cam, tile = np.concatenate(cam), np.concatenate(tile)
    n = cam.size
    if depths == "few":
        depth = rng.choice(np.float32([0.5, 1.25, 2.0, 2.0000002, 3.75]), n)
    elif depths == "one":
        depth = np.full(n, 1.5, np.float32)
    else:
        depth = rng.uniform(0.5, 4.0, n).astype(np.float32)
    shell = rng.integers(0, 5, n) if depths != "one" else np.zeros(n, np.int64)
    keys = orr.pack_keys(lay, cam, shell, tile, depth)
    perm = rng.permutation(n)
    keys = keys[perm]
    values = rng.permutation(n).astype(np.int32)
    order = np.argsort(keys, kind="stable")
    low = lay["tile"][0]  # camera | tile above, depth | shell below
    st = torch.cuda.current_stream().cuda_stream
    k_in, v_in = to_dev(keys), to_dev(values)
    k_out = torch.empty_like(k_in)
    v_out = torch.empty_like(v_in)
    ws = _native.workspace(_native.query("mt_sort_workspace_bytes", n), k_in.device)
    _native.call("mt_sort_keys_high", k_in.data_ptr(), v_in.data_ptr(), n, k_out.data_ptr(),
                 v_out.data_ptr(), ws.data_ptr(), ws.numel(), None, low, st)
    # Grouped by the high bits, in input order inside each group.
    hi = k_out.cpu().numpy() >> np.int64(low)
    want_hi = keys[np.argsort(keys >> np.int64(low), kind="stable")]
    np.testing.assert_array_equal(k_out.cpu().numpy(), want_hi)
    # Ranges of equal high bits, listed in a shuffled slot table with empty slots.
    starts = np.flatnonzero(np.r_[True, hi[1:] != hi[:-1]])
    ends = np.r_[starts[1:], n]
    table = np.zeros((starts.size + 40, 2), np.int64)
    slots = rng.permutation(table.shape[0])[: starts.size]
    table[slots, 0], table[slots, 1] = starts, ends
    tmp_k = torch.full_like(k_in, -1)
    tmp_v = torch.full_like(v_in, -1)
    rbytes = _native.query("mt_sort_key_ranges_workspace_bytes", table.shape[0])
    rws = torch.full((rbytes,), 0xA5, dtype=torch.uint8, device=k_in.device)  # no initial contents
    t_dev = to_dev(table)
    _native.call("mt_sort_key_ranges", k_out.data_ptr(), v_out.data_ptr(), n, t_dev.data_ptr(),
                 table.shape[0], low, tmp_k.data_ptr(), tmp_v.data_ptr(), rws.data_ptr(), rbytes, st)
    np.testing.assert_array_equal(k_out.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(v_out.cpu().numpy(), values[order])
    # Overlapping scratch is refused on the host.
    with pytest.raises(RuntimeError, match="separate"):
        _native.call("mt_sort_key_ranges", k_out.data_ptr(), v_out.data_ptr(), n, t_dev.data_ptr(),
                     table.shape[0], low, k_out.data_ptr(), tmp_v.data_ptr(), rws.data_ptr(), rbytes,
                     st)
    with pytest.raises(RuntimeError, match="workspace"):
        _native.call("mt_sort_key_ranges", k_out.data_ptr(), v_out.data_ptr(), n, t_dev.data_ptr(),
                     table.shape[0], low, tmp_k.data_ptr(), tmp_v.data_ptr(), rws.data_ptr(),
                     rbytes - 1, st)


def test_spherical_harmonics(scene):
    t = scene.tri
    verts = to_dev(t.vertices)
    sh = to_dev(t.spherical_harmonics).requires_grad_(True)
    col = tr.evaluate_spherical_harmonics(verts, sh, to_dev(scene.E), scene.spec.sh_degree)
    want, _ = orr.evaluate_spherical_harmonics(t.vertices, t.spherical_harmonics, scene.E, scene.spec.sh_degree)
    assert np.abs(col.detach().cpu().numpy() - want).max() < 1e-5
    g = np.random.default_rng(1).standard_normal(col.shape).astype(np.float32)
    col.backward(to_dev(g))
    wg = orr.evaluate_spherical_harmonics_backward(t.vertices, t.spherical_harmonics, scene.E,
                                                   scene.spec.sh_degree, g)
    got = sh.grad.cpu().numpy()
    assert np.abs(got - wg).max() / (np.abs(wg).max() + 1e-12) < 1e-5


def _fused(sc, chunk=0):
    t = sc.tri
    xy, depth, b, pk, spk, ranges = _front_end(sc)
    colors = to_dev(sc.colors).requires_grad_(True)
    sdf = to_dev(t.signed_distances).requires_grad_(True)
    sharp = torch.tensor(sc.spec.sharpness, device="cuda").requires_grad_(True)
    prev = config.composite_chunk
    config.set_composite_chunk(chunk)
    try:
        out = tr.intersect_and_composite(xy, colors, sdf, to_dev(t.faces), sharp, spk.keys,
                                         spk.triangle_indices, ranges.view(sc.B, sc.T, 2), sc.S,
                                         _grid(sc), _key_defines())
    finally:
        config.set_composite_chunk(prev)
    return out, colors, sdf, sharp


@pytest.mark.parametrize("chunk", [0, 32, 100])
def test_fused_forward_matches_oracle(scene, chunk):
    out, *_ = _fused(scene, chunk)
    got = out.detach().cpu().numpy()
    assert np.abs(got - scene.image).max() < IMG_TOL
    # FP64-continuous reference on the same hit lists.
    t = scene.tri
    img64 = orr.fused_forward_image(scene.hits, scene.colors.astype(np.float64),
                                    t.signed_distances.astype(np.float64), t.faces,
                                    scene.spec.sharpness, scene.B, (scene.H, scene.W), dtype=np.float64)
    assert np.abs(got - img64).max() < IMG_TOL


@pytest.mark.parametrize("quirk", [False, True])
def test_fused_backward_matches_oracle(scene, quirk):
    out, colors, sdf, sharp = _fused(scene)
    g = np.random.default_rng(2).standard_normal(out.shape).astype(np.float32)
    prev = config.reference_gradient_quirks
    config.set_reference_gradient_quirks(quirk)
    try:
        out.backward(to_dev(g))
    finally:
        config.set_reference_gradient_quirks(prev)
    t = scene.tri
    wc, ws, wk = orr.fused_backward(scene.hits, scene.colors.astype(np.float64),
                                    t.signed_distances.astype(np.float64), t.faces,
                                    scene.spec.sharpness, out.detach().cpu().numpy(), g,
                                    reference_quirk=quirk)
    for got, want in ((colors.grad, wc), (sdf.grad, ws)):
        gg = got.cpu().numpy().astype(np.float64)
        assert np.abs(gg - want).max() / (np.abs(want).max() + 1e-12) < GRAD_TOL
    assert abs(float(sharp.grad) - wk) / (abs(wk) + 1e-12) < GRAD_TOL


def test_explicit_mode_matches_oracle(scene_explicit):
    sc = scene_explicit
    t = sc.tri
    xy, depth, b, pk, spk, ranges = _front_end(sc, explicit=True)
    np.testing.assert_array_equal(spk.keys.cpu().numpy(), sc.sorted_keys)
    np.testing.assert_array_equal(ranges.cpu().numpy(), sc.ranges)
    uv, index = tr.intersect(xy, to_dev(t.faces), spk.triangle_indices, ranges, _grid(sc), _key_defines(True))
    want_uv, want_index = orr.explicit_intersect(sc.xy, t.faces, sc.sorted_values, sc.ranges,
                                                 sc.tile_shape, (sc.H, sc.W))
    np.testing.assert_array_equal(index.cpu().numpy(), want_index)
    assert np.array_equal(uv.cpu().numpy().view(np.uint32), want_uv.view(np.uint32))
    colors = to_dev(sc.colors).requires_grad_(True)
    sdf = to_dev(t.signed_distances).requires_grad_(True)
    s_img, c_img = tr.interpolate(uv, index, to_dev(t.faces), sdf, colors)
    ws_img, wc_img = orr.interpolate(want_uv, want_index, t.faces, t.signed_distances, sc.colors)
    assert np.abs(s_img.detach().cpu().numpy() - ws_img).max() < 1e-6
    assert np.abs(c_img.detach().cpu().numpy() - wc_img).max() < 1e-6
    B, S, H, W = want_index.shape
    sdf_pl = s_img.permute(0, 2, 3, 1).reshape(-1, S)
    col_pl = c_img.permute(0, 2, 3, 1, 4).reshape(-1, S, 3)
    mask = (index != -1).permute(0, 2, 3, 1).reshape(-1, S).contiguous()
    sharp = torch.tensor(sc.spec.sharpness, device="cuda").requires_grad_(True)
    from meshtryoshka_amd.triangle_rasterization.composite import Composite
    oc, ot = Composite.apply(sdf_pl.contiguous(), col_pl.contiguous(), mask, sharp)
    wcol, wtr = orr.composite(ws_img.transpose(0, 2, 3, 1).reshape(-1, S),
                              wc_img.transpose(0, 2, 3, 1, 4).reshape(-1, S, 3),
                              (want_index != -1).transpose(0, 2, 3, 1).reshape(-1, S), sc.spec.sharpness)
    assert np.abs(oc.detach().cpu().numpy() - wcol).max() < IMG_TOL
    assert np.abs(ot.detach().cpu().numpy() - wtr).max() < IMG_TOL


@pytest.mark.parametrize("quirk", [False, True])
def test_composite_backward_matches_oracle(quirk):
    rng = np.random.default_rng(5)
    P, S = 5000, 5
    sdf = (rng.standard_normal((P, S)) * 0.05).astype(np.float32)
    col = rng.random((P, S, 3)).astype(np.float32)
    mask = rng.random((P, S)) < 0.7
    k = 40.0
    ts, tc = to_dev(sdf).requires_grad_(True), to_dev(col).requires_grad_(True)
    sharp = torch.tensor(k, device="cuda").requires_grad_(True)
    from meshtryoshka_amd.triangle_rasterization.composite import Composite
    oc, ot = Composite.apply(ts, tc, to_dev(mask), sharp)
    gc = rng.standard_normal((P, 3)).astype(np.float32)
    gt = rng.standard_normal(P).astype(np.float32)
    prev = config.reference_gradient_quirks
    config.set_reference_gradient_quirks(quirk)
    try:
        torch.autograd.backward([oc, ot], [to_dev(gc), to_dev(gt)])
    finally:
        config.set_reference_gradient_quirks(prev)
    ws, wc, wk = orr.composite_backward(sdf, col, mask, k, oc.detach().cpu().numpy(),
                                        ot.detach().cpu().numpy(), gc, gt, reference_quirk=quirk)
    for got, want in ((ts.grad, ws), (tc.grad, wc)):
        gg = got.cpu().numpy().astype(np.float64)
        assert np.abs(gg - want).max() / (np.abs(want).max() + 1e-12) < GRAD_TOL
    assert abs(float(sharp.grad) - wk) / (abs(wk) + 1e-12) < GRAD_TOL


def test_render_fused_end_to_end(scene):
    t = scene.tri
    img = tr.render(to_dev(scene.E), to_dev(scene.K), (scene.H, scene.W), to_dev(t.vertices),
                    to_dev(t.faces), to_dev(t.face_boundaries), to_dev(t.signed_distances),
                    to_dev(t.spherical_harmonics), scene.spec.sh_degree,
                    torch.tensor(scene.spec.sharpness, device="cuda"),
                    torch.ones(3, device="cuda"), mode="fused")
    want = np.clip(scene.image[:, :3] + scene.image[:, 3:4], 0, 1)
    assert np.abs(img.cpu().numpy() - want).max() < IMG_TOL


def test_tile_shape_invariance(scene):
    t = scene.tri
    imgs = []
    for ts in [(16, 16), (8, 8), (8, 32), (32, 16)]:
        imgs.append(tr.render(to_dev(scene.E), to_dev(scene.K), (scene.H, scene.W), to_dev(t.vertices),
                              to_dev(t.faces), to_dev(t.face_boundaries), to_dev(t.signed_distances),
                              to_dev(t.spherical_harmonics), scene.spec.sh_degree,
                              torch.tensor(scene.spec.sharpness, device="cuda"),
                              torch.ones(3, device="cuda"), tile_shape=ts).cpu().numpy())
    for im in imgs[1:]:
        assert np.array_equal(im, imgs[0])


@pytest.mark.parametrize("cull", [0, 1])
def test_single_pass_keys_match_three_pass(cull):
    """mt_generate_keys_single_pass (each unit's keys counted and written into a reserved
    range, then moved to unit order) writes the same total as mt_key_block_counts + scan +
    mt_generate_keys_from_geometry and, per camera, the same keys and values in the same
    order (cameras are interleaved by face block), so that the stable sort gives identical
    results; with and without culling, for several cameras (thousands of units)."""
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=4, num_spheres=10, resolution=192, num_cameras=5,
                                  image_shape=(150, 230), device=dev)
    vox = te.extract_voxels(spec.occupancy.to(dev))
    world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    with torch.no_grad():
        tri = te.tessellate_voxels(sdf.to(dev), world.to(dev), sh.to(dev), vox.neighbors,
                                   vox.lower_corners, vox.upper_corners, vox.indices,
                                   spec.level_sets.to(dev))
    cams = spec.cameras.to(dev)
    b = cams.extrinsics.shape[0]
    xy, depth = tr.project_vertices(tri.vertices, cams.extrinsics, cams.intrinsics)
    xy, depth = xy.contiguous(), depth.contiguous()
    v, f = tri.vertices.shape[0], tri.faces.shape[0]
    th, tw = 8, 8
    rows, cols = -(-150 // th), -(-230 // tw)
    st = torch.cuda.current_stream().cuda_stream
    layout = to_layout(create_key_defines(("camera", 6), ("tile", 20), ("depth", 32), ("shell", 6)))
    nblk = -(-f // 256)
    counts = torch.empty(b * nblk, dtype=torch.int32, device=dev)
    faces = tri.faces.contiguous()
    bounds = tri.face_boundaries.contiguous()
    _native.call("mt_key_block_counts", xy.data_ptr(), depth.data_ptr(), v, faces.data_ptr(), f, None,
                 b, th, tw, 0, 0, rows, cols, 0.2, None, cull, counts.data_ptr(), st)
    offs = torch.empty(b * nblk, dtype=torch.int64, device=dev)
    ws = _native.workspace(_native.query("mt_scan_workspace_bytes", b * nblk), dev)
    _native.call("mt_inclusive_scan_i32_to_i64", counts.data_ptr(), offs.data_ptr(), b * nblk,
                 ws.data_ptr(), ws.numel(), st)
    k = int(offs[-1])
    assert k > 0
    keys = torch.empty(k, dtype=torch.int64, device=dev)
    vals = torch.empty(k, dtype=torch.int32, device=dev)
    over = torch.zeros(1, dtype=torch.int32, device=dev)
    _native.call("mt_generate_keys_from_geometry", xy.data_ptr(), depth.data_ptr(), v,
                 faces.data_ptr(), f, None, bounds.data_ptr(), bounds.numel(), b, th, tw, 0, 0, rows,
                 cols, 0.2, None, cull, offs.data_ptr(), cols, layout, keys.data_ptr(),
                 vals.data_ptr(), k, over.data_ptr(), st)
    keys2 = torch.full((k + 7,), -1, dtype=torch.int64, device=dev)
    vals2 = torch.full((k + 7,), -1, dtype=torch.int32, device=dev)
    n_keys = torch.zeros(1, dtype=torch.int64, device=dev)
    ws2 = _native.workspace(_native.query("mt_generate_keys_single_pass_workspace_bytes", f, b), dev)
    scratch_k = torch.empty(k + 7, dtype=torch.int64, device=dev)
    scratch_v = torch.empty(k + 7, dtype=torch.int32, device=dev)
    _native.call("mt_generate_keys_single_pass", xy.data_ptr(), depth.data_ptr(), v, faces.data_ptr(),
                 f, None, bounds.data_ptr(), bounds.numel(), b, th, tw, 0, 0, rows, cols, 0.2, None,
                 cull, cols, layout, keys2.data_ptr(), vals2.data_ptr(), k + 7, n_keys.data_ptr(),
                 over.data_ptr(), scratch_k.data_ptr(), scratch_v.data_ptr(), ws2.data_ptr(),
                 ws2.numel(), st)
    torch.cuda.synchronize()
    assert int(over) == 0
    assert int(n_keys) == k
    cam, cam2 = (keys >> 58) & 63, (keys2[:k] >> 58) & 63
    assert int(cam.max()) == b - 1
    for c in range(b):
        assert torch.equal(keys2[:k][cam2 == c], keys[cam == c])
        assert torch.equal(vals2[:k][cam2 == c], vals[cam == c])
    # The stable sort of both arrangements is the same.
    order, order2 = torch.sort(keys, stable=True).indices, torch.sort(keys2[:k], stable=True).indices
    assert torch.equal(keys[order], keys2[:k][order2]) and torch.equal(vals[order], vals2[:k][order2])
    assert bool((keys2[k:] == -1).all())
    # Too small a capacity: the count is still exact, the overflow flag is raised, nothing
    # is written past the capacity, and every key below it is the right key or key 0 with
    # value 0 (a safe stand-in for keys that did not fit).
    small = torch.full((k // 2 + 1,), -1, dtype=torch.int64, device=dev)
    small_v = torch.full((k // 2 + 1,), -1, dtype=torch.int32, device=dev)
    scratch_small_k = torch.full((k // 2 + 1,), -1, dtype=torch.int64, device=dev)
    scratch_small_v = torch.full((k // 2 + 1,), -1, dtype=torch.int32, device=dev)
    _native.call("mt_generate_keys_single_pass", xy.data_ptr(), depth.data_ptr(), v, faces.data_ptr(),
                 f, None, bounds.data_ptr(), bounds.numel(), b, th, tw, 0, 0, rows, cols, 0.2, None,
                 cull, cols, layout, small.data_ptr(), small_v.data_ptr(), k // 2, n_keys.data_ptr(),
                 over.data_ptr(), scratch_small_k.data_ptr(), scratch_small_v.data_ptr(),
                 ws2.data_ptr(), ws2.numel(), st)
    torch.cuda.synchronize()
    assert int(over) == 1 and int(n_keys) == k
    got, got_v = small[: k // 2], small_v[: k // 2]
    right = (got == keys2[: k // 2]) & (got_v == vals2[: k // 2])
    stand_in = (got == 0) & (got_v == 0)
    assert bool((right | stand_in).all()) and bool(right.any())
    assert int(small[k // 2]) == -1 and int(scratch_small_k[k // 2]) == -1
    # Scratch arrays that overlap the outputs are refused.
    with pytest.raises(RuntimeError, match="separate device arrays"):
        _native.call("mt_generate_keys_single_pass", xy.data_ptr(), depth.data_ptr(), v,
                     faces.data_ptr(), f, None, bounds.data_ptr(), bounds.numel(), b, th, tw, 0, 0,
                     rows, cols, 0.2, None, cull, cols, layout, keys2.data_ptr(), vals2.data_ptr(),
                     k + 7, n_keys.data_ptr(), over.data_ptr(), keys2.data_ptr(), scratch_v.data_ptr(),
                     ws2.data_ptr(), ws2.numel(), st)
