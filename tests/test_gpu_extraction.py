"""GPU parity of the extraction kernels against the CPU oracle (bit-exact for integer
outputs; vertex attributes within 1e-5 relative, expected bit-exact)."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import triangle_extraction as te
from oracle import extraction as ox
from tests.helpers import OracleScene, small_sphere_spec, to_dev

pytestmark = pytest.mark.gpu


def _random_packed(shape, density, seed):
    rng = np.random.default_rng(seed)
    i, j, k = shape
    occ = rng.random((i, j, k)) < density
    return ox.pack_occupancy(occ)


GRIDS = [
    ((4, 4, 32), 0.3, 0),
    ((5, 7, 64), 0.2, 1),
    ((16, 16, 32), 0.05, 2),
    ((9, 3, 96), 0.5, 3),
    ((8, 8, 32), 1.0, 4),
    ((6, 5, 64), 0.0, 5),
]


@pytest.mark.parametrize("shape,density,seed", GRIDS)
def test_extract_voxels_bit_exact(shape, density, seed):
    packed = _random_packed(shape, density, seed)
    want = ox.extract_voxels(packed)
    got = te.extract_voxels(to_dev(packed))
    assert got.grid_shape == want.grid_shape
    np.testing.assert_array_equal(got.vertices.cpu().numpy().view(np.uint32), want.vertices.view(np.uint32))
    for name in ("neighbors", "lower_corners", "upper_corners", "indices"):
        np.testing.assert_array_equal(getattr(got, name).cpu().numpy(), getattr(want, name), err_msg=name)


def test_extract_voxels_single_cells_at_word_edges():
    # Cells on bit 0 / bit 31 of a word and on the grid's high faces exercise the
    # last-bit branches of the structure.
    occ = np.zeros((4, 4, 64), dtype=bool)
    for (i, j, k) in [(0, 0, 0), (0, 0, 31), (1, 2, 32), (3, 3, 63), (2, 0, 31), (3, 1, 32)]:
        occ[i, j, k] = True
    packed = ox.pack_occupancy(occ)
    want = ox.extract_voxels(packed)
    got = te.extract_voxels(to_dev(packed))
    for name in ("neighbors", "lower_corners", "upper_corners", "indices"):
        np.testing.assert_array_equal(getattr(got, name).cpu().numpy(), getattr(want, name), err_msg=name)


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 100_003, 3_000_001])
def test_exclusive_scan(n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 7, size=n, dtype=np.int32)
    t = to_dev(x)
    te.compute_exclusive_cumsum(t)
    want = np.zeros_like(x)
    want[1:] = np.cumsum(x[:-1], dtype=np.int64).astype(np.int32)
    np.testing.assert_array_equal(t.cpu().numpy(), want)


def test_exclusive_scan_wraps_like_int32():
    x = np.full(10_000, 2**20, dtype=np.int32)  # total 1.0e10 > 2^31
    t = to_dev(x)
    te.compute_exclusive_cumsum(t)
    want = (np.arange(10_000, dtype=np.int64) * 2**20).astype(np.int32)
    np.testing.assert_array_equal(t.cpu().numpy(), want)


@pytest.fixture(scope="module")
def scene():
    return OracleScene(small_sphere_spec(resolution=32, image=48))


def _gpu_tessellate(sc):
    v = sc.vox
    dev = lambda a: to_dev(a)  # noqa: E731
    return te.tessellate_voxels(dev(sc.sdf), dev(sc.world), dev(sc.sh), dev(v.neighbors),
                                dev(v.lower_corners), dev(v.upper_corners), dev(v.indices),
                                dev(sc.level_sets))


def test_tessellate_matches_oracle(scene):
    got = _gpu_tessellate(scene)
    want = scene.tri
    np.testing.assert_array_equal(got.faces.cpu().numpy(), want.faces)
    np.testing.assert_array_equal(got.voxel_indices.cpu().numpy(), want.voxel_indices)
    np.testing.assert_array_equal(got.vertex_boundaries.cpu().numpy(), want.vertex_boundaries)
    np.testing.assert_array_equal(got.face_boundaries.cpu().numpy(), want.face_boundaries)
    gv = got.vertices.detach().cpu().numpy()
    np.testing.assert_allclose(gv, want.vertices, rtol=1e-5, atol=1e-7)
    # Expected bit-exact: same IEEE operations in the same order, no contraction.
    assert np.array_equal(gv.view(np.uint32), want.vertices.view(np.uint32))
    np.testing.assert_array_equal(got.signed_distances.detach().cpu().numpy(), want.signed_distances)
    np.testing.assert_array_equal(got.spherical_harmonics.detach().cpu().numpy(), want.spherical_harmonics)


def test_stage_outputs_match_oracle(scene):
    v = scene.vox
    sdf, lower, upper = to_dev(scene.sdf), to_dev(v.lower_corners), to_dev(v.upper_corners)
    ls = to_dev(scene.level_sets)
    counts = te.count_triangle_vertices(sdf, lower, ls)
    np.testing.assert_array_equal(counts.cpu().numpy(), scene.tri.vertex_counts)
    te.compute_exclusive_cumsum(counts)
    np.testing.assert_array_equal(counts.cpu().numpy(), scene.tri.vertex_offsets)
    fc, codes = te.count_triangle_faces(sdf, lower, upper, ls)
    np.testing.assert_array_equal(fc.cpu().numpy(), scene.tri.face_counts)
    np.testing.assert_array_equal(codes.cpu().numpy(), scene.tri.cell_codes)


def test_faces_from_crossing_axes_match_oracle(scene):
    """The batched step's faces pass (mt_count_triangle_vertices_axes +
    mt_create_triangle_faces_axes_i64: each corner's vertex from the owner column's
    crossing-axis bits) gives the oracle's faces; the bits agree with the counts and with
    the vertex types, and a vertex capacity below the vertex count raises the flag."""
    from meshtryoshka_amd import _native

    v = scene.vox
    sdf, lower, upper = to_dev(scene.sdf), to_dev(v.lower_corners), to_dev(v.upper_corners)
    ls = to_dev(scene.level_sets)
    n, s, nvox = lower.shape[1], ls.shape[0], upper.shape[1]
    st = torch.cuda.current_stream().cuda_stream
    counts = torch.zeros(s * n + 1, dtype=torch.int32, device="cuda")
    axes = torch.full((s * n,), 255, dtype=torch.uint8, device="cuda")
    _native.call("mt_count_triangle_vertices_axes", sdf.data_ptr(), lower.data_ptr(), n, ls.data_ptr(),
                 s, counts.data_ptr(), axes.data_ptr(), st)
    np.testing.assert_array_equal(counts.cpu().numpy(), scene.tri.vertex_counts)
    bits = axes.to(torch.int32)
    pop = (bits & 1) + ((bits >> 1) & 1) + ((bits >> 2) & 1)
    assert bool((bits < 8).all()) and torch.equal(pop, counts[:-1])
    # The k-th vertex of a column has the type of its k-th set bit (the oracle's types).
    offsets = torch.as_tensor(scene.tri.vertex_offsets.astype(np.int64))
    types = torch.as_tensor(scene.tri.vertex_types.astype(np.int64))
    b = bits.cpu().to(torch.int64)
    for a in range(3):
        has = ((b >> a) & 1).bool()
        rank = ((b & ((1 << a) - 1)) & 1) + (((b & ((1 << a) - 1)) >> 1) & 1)
        idx = offsets[:-1][has] + rank[has]
        assert bool((types[idx] == a).all())
    fc, codes = te.count_triangle_faces(sdf, lower, upper, ls)
    foff = torch.zeros(s * nvox + 1, dtype=torch.int64, device="cuda")
    foff[1:] = torch.cumsum(fc[:-1].to(torch.int64), 0)  # counts carry one trailing entry
    nf = int(foff[-1]) // 3
    voff = counts.clone()
    te.compute_exclusive_cumsum(voff)
    nv = int(voff[-1])
    faces = torch.full((nf, 3), -7, dtype=torch.int32, device="cuda")
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    nbr = to_dev(v.neighbors)  # kept alive across the calls
    _native.call("mt_create_triangle_faces_axes_i64", nbr.data_ptr(), nvox, n,
                 voff.data_ptr(), axes.data_ptr(), codes.data_ptr(), s, foff.data_ptr(),
                 faces.data_ptr(), nf, nv, flag.data_ptr(), st)
    assert int(flag) == 0
    np.testing.assert_array_equal(faces.cpu().numpy(), scene.tri.faces)
    _native.call("mt_create_triangle_faces_axes_i64", nbr.data_ptr(), nvox, n,
                 voff.data_ptr(), axes.data_ptr(), codes.data_ptr(), s, foff.data_ptr(),
                 faces.data_ptr(), nf, nv // 2, flag.data_ptr(), st)
    assert int(flag) == 1
    # Column words ((first vertex << 3) | crossing axes), written by the vertex pass, and the
    # faces read from them: the same faces; a capacity below the vertex count raises the flag.
    world, sh = to_dev(scene.world), to_dev(scene.sh)
    words = torch.full((s * n,), 0xFFFFFFF, dtype=torch.int32, device="cuda")
    tv = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    td = torch.empty(nv, dtype=torch.float32, device="cuda")
    tt = torch.empty(nv, dtype=torch.uint8, device="cuda")
    _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(), sdf.data_ptr(),
                 sh.data_ptr(), world.shape[0], sh.shape[0], 0, lower.data_ptr(), n, ls.data_ptr(),
                 s, voff.data_ptr(), nv, None, 0, tv.data_ptr(), td.data_ptr(), tt.data_ptr(), None,
                 flag.data_ptr(), 1, 0, words.data_ptr(), st)
    want = (voff[:-1].to(torch.int64) << 3) | axes.to(torch.int64)
    assert torch.equal(words.to(torch.int64) & 0xFFFFFFFF, want)
    flag.zero_()
    faces.fill_(-7)
    _native.call("mt_create_triangle_faces_words_i64", nbr.data_ptr(), nvox, n, words.data_ptr(),
                 codes.data_ptr(), s, foff.data_ptr(), faces.data_ptr(), nf, nv, flag.data_ptr(), st)
    assert int(flag) == 0
    np.testing.assert_array_equal(faces.cpu().numpy(), scene.tri.faces)
    _native.call("mt_create_triangle_faces_words_i64", nbr.data_ptr(), nvox, n, words.data_ptr(),
                 codes.data_ptr(), s, foff.data_ptr(), faces.data_ptr(), nf, nv // 2, flag.data_ptr(),
                 st)
    assert int(flag) == 1
    with pytest.raises(RuntimeError, match="2\\^29"):
        _native.call("mt_create_triangle_faces_words_i64", nbr.data_ptr(), nvox, n,
                     words.data_ptr(), codes.data_ptr(), s, foff.data_ptr(), faces.data_ptr(), nf,
                     (1 << 29) + 1, flag.data_ptr(), st)


def test_create_vertices_backward_matches_oracle(scene):
    v = scene.vox
    grid_v = to_dev(scene.world).requires_grad_(True)
    grid_d = to_dev(scene.sdf).requires_grad_(True)
    grid_sh = to_dev(scene.sh).requires_grad_(True)
    counts = te.count_triangle_vertices(grid_d.detach(), to_dev(v.lower_corners), to_dev(scene.level_sets))
    te.compute_exclusive_cumsum(counts)
    out = te.create_triangle_vertices(grid_v, grid_d, grid_sh, to_dev(v.lower_corners),
                                      to_dev(scene.level_sets), counts)
    rng = np.random.default_rng(7)
    gv = rng.standard_normal(out.vertices.shape).astype(np.float32)
    gd = rng.standard_normal(out.signed_distances.shape).astype(np.float32)
    gs = rng.standard_normal(out.spherical_harmonics.shape).astype(np.float32)
    torch.autograd.backward([out.vertices, out.signed_distances, out.spherical_harmonics],
                            [to_dev(gv), to_dev(gd), to_dev(gs)])
    wv, wd, ws = ox.create_triangle_vertices_backward(scene.tri.record, scene.sdf.shape[0],
                                                      scene.sh.shape[0], gv, gd, gs)
    for got, want in ((grid_v.grad, wv), (grid_d.grad, wd), (grid_sh.grad, ws)):
        g = got.cpu().numpy().astype(np.float64)
        scale = np.abs(want).max() + 1e-12
        assert np.abs(g - want).max() / scale < 1e-5


def test_vertices_backward_with_shared_origins_uses_scatter_path(scene):
    """Columns whose origins repeat (not a permutation of the grid vertices) make the
    gather path's inverse maps collide; the device then runs the atomic scatter instead."""
    v = scene.vox
    lower = np.concatenate([v.lower_corners, v.lower_corners[:, -50:]], axis=1)
    ls = scene.level_sets
    counts = ox.count_triangle_vertices(scene.sdf, lower, ls)
    offsets = ox.exclusive_cumsum(counts)
    _, _, _, _, record = ox.create_triangle_vertices(scene.world, scene.sdf, scene.sh, lower, ls, offsets)
    grid_d = to_dev(scene.sdf).requires_grad_(True)
    grid_sh = to_dev(scene.sh).requires_grad_(True)
    t_off = to_dev(offsets)
    out = te.create_triangle_vertices(to_dev(scene.world), grid_d, grid_sh, to_dev(lower), to_dev(ls), t_off)
    rng = np.random.default_rng(11)
    gd = rng.standard_normal(out.signed_distances.shape).astype(np.float32)
    gs = rng.standard_normal(out.spherical_harmonics.shape).astype(np.float32)
    torch.autograd.backward([out.signed_distances, out.spherical_harmonics], [to_dev(gd), to_dev(gs)])
    _, wd, ws = ox.create_triangle_vertices_backward(record, scene.sdf.shape[0], scene.sh.shape[0],
                                                     None, gd, gs)
    for got, want in ((grid_d.grad, wd), (grid_sh.grad, ws)):
        g = got.cpu().numpy().astype(np.float64)
        assert np.abs(g - want).max() / (np.abs(want).max() + 1e-12) < 1e-5
