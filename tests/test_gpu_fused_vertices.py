"""Fused vertex + colour kernels of the batched step against the separate kernels they
replace: mt_create_triangle_vertices_colors_forward == mt_create_triangle_vertices_forward
followed by mt_evaluate_sh_forward (bit for bit), and the fused adjoint == mt_evaluate_sh_backward
followed by mt_create_triangle_vertices_backward (equal up to float summation order), plus the
FP64 oracle adjoint within 1e-3 relative."""

import numpy as np
import pytest
import torch

from meshtryoshka_amd import _native
from meshtryoshka_amd import triangle_extraction as te
from meshtryoshka_amd import workloads
from oracle import extraction as ox
from oracle import rasterization as orr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[3, 1])
def case(request):
    dev = torch.device("cuda")
    spec = workloads.sphere_field(seed=5, num_spheres=6, resolution=64, num_cameras=3,
                                  image_shape=(48, 64), sh_degree=request.param, device=dev)
    vox = te.extract_voxels(spec.occupancy.to(dev))
    world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    ls = spec.level_sets.to(dev)
    cams = spec.cameras.to(dev)
    return spec, vox, world.to(dev), sdf.to(dev), sh.to(dev), ls, cams


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _offsets(vox, sdf, ls):
    n = vox.lower_corners.shape[1]
    s = ls.shape[0]
    off = torch.zeros(s * n + 1, dtype=torch.int32, device=sdf.device)
    _native.call("mt_count_triangle_vertices", sdf.data_ptr(), vox.lower_corners.data_ptr(), n,
                 ls.data_ptr(), s, off.data_ptr(), _stream())
    ws = _native.workspace(_native.query("mt_scan_workspace_bytes", off.numel()), sdf.device)
    _native.call("mt_exclusive_scan_i32", off.data_ptr(), off.numel(), ws.data_ptr(), ws.numel(),
                 _stream())
    return off


def test_fused_forward_and_adjoint_match_separate_kernels(case):
    spec, vox, world, sdf, sh, ls, cams = case
    dev = sdf.device
    n, s, ng, nsh, deg = vox.lower_corners.shape[1], ls.shape[0], world.shape[0], sh.shape[0], spec.sh_degree
    b = cams.extrinsics.shape[0]
    off = _offsets(vox, sdf, ls)
    V = int(off[-1])
    f32 = dict(dtype=torch.float32, device=dev)
    campos = torch.empty((b, 3), **f32)
    _native.call("mt_camera_positions", cams.extrinsics.contiguous().data_ptr(), b, campos.data_ptr(), _stream())
    # Separate kernels.
    tv, td = torch.empty((V, 3), **f32), torch.empty(V, **f32)
    tsh, tt = torch.empty((nsh, V, 3), **f32), torch.empty(V, dtype=torch.uint8, device=dev)
    ov = torch.zeros(1, dtype=torch.int32, device=dev)
    _native.call("mt_create_triangle_vertices_forward", world.data_ptr(), sdf.data_ptr(), sh.data_ptr(),
                 ng, nsh, vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V,
                 tv.data_ptr(), td.data_ptr(), tsh.data_ptr(), tt.data_ptr(), ov.data_ptr(), _stream())
    col = torch.empty((b, V, 3), **f32)
    _native.call("mt_evaluate_sh_forward", tv.data_ptr(), V, tsh.data_ptr(), nsh, deg,
                 campos.data_ptr(), b, col.data_ptr(), _stream())
    # Fused kernel.
    tv2, td2 = torch.empty((V, 3), **f32), torch.empty(V, **f32)
    tt2, col2 = torch.empty(V, dtype=torch.uint8, device=dev), torch.empty((b, V, 3), **f32)
    _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(), sdf.data_ptr(),
                 sh.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(), n, ls.data_ptr(), s,
                 off.data_ptr(), V, campos.data_ptr(), b, tv2.data_ptr(), td2.data_ptr(),
                 tt2.data_ptr(), col2.data_ptr(), ov.data_ptr(), 3, 0, None, _stream())
    assert int(ov) == 0
    assert torch.equal(tv, tv2) and torch.equal(td, td2) and torch.equal(tt, tt2)
    assert torch.equal(col, col2)
    # Split in two calls (vertices, then colours): the same outputs, and each call leaves the
    # other's outputs untouched.
    tv3, td3 = torch.full((V, 3), 7.0, **f32), torch.full((V,), 7.0, **f32)
    tt3 = torch.full((V,), 7, dtype=torch.uint8, device=dev)
    col3 = torch.full((b, V, 3), 7.0, **f32)
    _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(), sdf.data_ptr(),
                 sh.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(), n, ls.data_ptr(), s,
                 off.data_ptr(), V, campos.data_ptr(), b, tv3.data_ptr(), td3.data_ptr(),
                 tt3.data_ptr(), col3.data_ptr(), ov.data_ptr(), 1, 0, None, _stream())
    assert torch.equal(tv3, tv) and torch.equal(td3, td) and torch.equal(tt3, tt)
    assert bool((col3 == 7.0).all())
    tv4 = torch.full((V, 3), 7.0, **f32)
    _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(), sdf.data_ptr(),
                 sh.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(), n, ls.data_ptr(), s,
                 off.data_ptr(), V, campos.data_ptr(), b, tv4.data_ptr(), td3.data_ptr(),
                 tt3.data_ptr(), col3.data_ptr(), ov.data_ptr(), 2, 0, None, _stream())
    assert torch.equal(col3, col) and bool((tv4 == 7.0).all())
    # SH rows: a permutation of grid_sh's first (deg+1)^2 planes, zero-padded; the colour
    # pass reading them gives the same bits.
    kn = (deg + 1) ** 2
    rf = _native.query("mt_sh_row_floats", deg)
    assert rf % 4 == 0 and 3 * kn <= rf < 3 * kn + 4
    rows = torch.full((ng, rf), 7.0, **f32)
    _native.call("mt_pack_sh_rows", sh.data_ptr(), ng, nsh, deg, rows.data_ptr(), _stream())
    assert torch.equal(rows[:, : 3 * kn], sh[:kn].permute(1, 0, 2).reshape(ng, 3 * kn))
    assert bool((rows[:, 3 * kn :] == 0).all())
    for parts in (3 | 4, 2 | 4):
        col5 = torch.full((b, V, 3), 7.0, **f32)
        tv5, td5 = torch.full((V, 3), 7.0, **f32), torch.full((V,), 7.0, **f32)
        tt5 = torch.full((V,), 7, dtype=torch.uint8, device=dev)
        _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(),
                     sdf.data_ptr(), rows.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(),
                     n, ls.data_ptr(), s, off.data_ptr(), V, campos.data_ptr(), b, tv5.data_ptr(),
                     td5.data_ptr(), tt5.data_ptr(), col5.data_ptr(), ov.data_ptr(), parts,
                     0, None, _stream())
        assert torch.equal(col5, col)
        if parts & 1:
            assert torch.equal(tv5, tv) and torch.equal(td5, td) and torch.equal(tt5, tt)
        else:
            assert bool((tv5 == 7.0).all())

    # Vertex-major colours (colors_stride >= b): the same values, each vertex's cameras
    # together; a larger stride leaves the unused camera slots untouched.
    for stride in (b, b + 2):
        col6 = torch.full((V, stride, 3), 7.0, **f32)
        _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(),
                     sdf.data_ptr(), rows.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(),
                     n, ls.data_ptr(), s, off.data_ptr(), V, campos.data_ptr(), b, None, None,
                     None, col6.data_ptr(), ov.data_ptr(), 2 | 4, stride, None, _stream())
        assert torch.equal(col6[:, :b].permute(1, 0, 2), col)
        assert bool((col6[:, b:] == 7.0).all())
    with pytest.raises(RuntimeError, match="colors_stride"):
        _native.call("mt_create_triangle_vertices_colors_forward", world.data_ptr(),
                     sdf.data_ptr(), rows.data_ptr(), ng, nsh, deg, vox.lower_corners.data_ptr(),
                     n, ls.data_ptr(), s, off.data_ptr(), V, campos.data_ptr(), b, None, None,
                     None, col3.data_ptr(), ov.data_ptr(), 2 | 4, b - 1, None, _stream())

    # Adjoint: random upstream gradients for SDF and colours.
    gen = torch.Generator("cpu").manual_seed(11)
    g_d = torch.randn(V, generator=gen).to(dev)
    g_col = torch.randn((b, V, 3), generator=gen).to(dev)
    g_tsh = torch.zeros((nsh, V, 3), **f32)
    _native.call("mt_evaluate_sh_backward", tv.data_ptr(), V, tsh.data_ptr(), nsh, deg,
                 campos.data_ptr(), b, g_col.data_ptr(), g_tsh.data_ptr(), _stream())
    ws = _native.workspace(_native.query("mt_create_triangle_vertices_backward_workspace_bytes", ng), dev)
    gg_d, gg_sh = torch.zeros(ng, **f32), torch.zeros((nsh, ng, 3), **f32)
    _native.call("mt_create_triangle_vertices_backward", sdf.data_ptr(), ng, nsh,
                 vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V, None,
                 g_d.data_ptr(), g_tsh.data_ptr(), None, gg_d.data_ptr(), gg_sh.data_ptr(),
                 ws.data_ptr(), ws.numel(), _stream())
    ws2 = _native.workspace(_native.query("mt_create_triangle_vertices_colors_backward_workspace_bytes",
                                          ng, n, deg), dev)
    gg_d2 = torch.full((ng,), float("nan"), **f32)  # fully overwritten
    gg_sh2 = torch.full((nsh, ng, 3), float("nan"), **f32)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    _native.call("mt_create_triangle_vertices_colors_backward", sdf.data_ptr(), ng, nsh, deg,
                 vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V, tv2.data_ptr(),
                 col2.data_ptr(), campos.data_ptr(), b, g_d.data_ptr(), g_col.data_ptr(),
                 gg_d2.data_ptr(), gg_sh2.data_ptr(), ws2.data_ptr(), ws2.numel(), 0, None, 0, 0,
                 status.data_ptr(), 0, _stream())
    assert int(status) == 0
    # Same terms, summed per column and axis first: equal up to float summation order.
    torch.testing.assert_close(gg_d2, gg_d, rtol=1e-5, atol=1e-6 * float(gg_d.abs().max()))
    torch.testing.assert_close(gg_sh2, gg_sh, rtol=1e-5, atol=1e-6 * float(gg_sh.abs().max()))

    # Gradients only at some vertices, flagged as the composite backward flags them: the
    # adjoint that skips unflagged vertices gives the same sums; with clear_gradients it
    # also leaves the vertex gradients all zero. The flagged calls reuse the grid-vertex maps
    # that the workspace holds from the call before (maps_ready = 1).
    flags = (torch.rand(V, generator=gen) < 0.1).to(dev)
    g_d_s = torch.where(flags, g_d, torch.zeros_like(g_d))
    g_col_s = torch.where(flags[None, :, None], g_col, torch.zeros_like(g_col))
    outs = []
    for touched, clear in ((None, 0), (flags.to(torch.uint8), 0), (flags.to(torch.uint8), 1)):
        gd_in, gc_in = g_d_s.clone(), g_col_s.clone()
        gd_o = torch.full((ng,), float("nan"), **f32)
        gs_o = torch.full((nsh, ng, 3), float("nan"), **f32)
        _native.call("mt_create_triangle_vertices_colors_backward", sdf.data_ptr(), ng, nsh, deg,
                     vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V,
                     tv2.data_ptr(), col2.data_ptr(), campos.data_ptr(), b, gd_in.data_ptr(),
                     gc_in.data_ptr(), gd_o.data_ptr(), gs_o.data_ptr(), ws2.data_ptr(),
                     ws2.numel(), 0, None if touched is None else touched.data_ptr(), clear,
                     int(touched is not None), status.data_ptr(), 0, _stream())
        outs.append((gd_o, gs_o))
        if clear:
            assert not bool(gd_in.any()) and not bool(gc_in.any())
        else:
            assert torch.equal(gd_in, g_d_s) and torch.equal(gc_in, g_col_s)
    assert int(status) == 0
    for o in outs[1:]:
        assert torch.equal(outs[0][0], o[0]) and torch.equal(outs[0][1], o[1])
    # The same flagged, clearing call on vertex-major colours and colour gradients (stride
    # b + 1): the same sums, bit for bit, and the gradients left zero.
    stride = b + 1
    col_vm = torch.zeros((V, stride, 3), **f32)
    col_vm[:, :b] = col2.permute(1, 0, 2)
    gc_vm = torch.full((V, stride, 3), 0.0, **f32)
    gc_vm[:, :b] = g_col_s.permute(1, 0, 2)
    gd_in = g_d_s.clone()
    gd_o = torch.full((ng,), float("nan"), **f32)
    gs_o = torch.full((nsh, ng, 3), float("nan"), **f32)
    _native.call("mt_create_triangle_vertices_colors_backward", sdf.data_ptr(), ng, nsh, deg,
                 vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V,
                 tv2.data_ptr(), col_vm.data_ptr(), campos.data_ptr(), b, gd_in.data_ptr(),
                 gc_vm.data_ptr(), gd_o.data_ptr(), gs_o.data_ptr(), ws2.data_ptr(), ws2.numel(), 0,
                 flags.to(torch.uint8).data_ptr(), 1, 1, status.data_ptr(), stride, _stream())
    assert int(status) == 0
    assert torch.equal(gd_o, outs[0][0]) and torch.equal(gs_o, outs[0][1])
    assert not bool(gd_in.any()) and not bool(gc_vm.any())

    # Accumulating mode adds to what the grid gradients hold: splitting the cameras into
    # two calls (the second accumulating) gives the one-call result.
    gg_d3 = torch.full((ng,), float("nan"), **f32)
    gg_sh3 = torch.full((nsh, ng, 3), float("nan"), **f32)
    for i, (c0, c1) in enumerate([(0, 1), (1, b)]):
        gd_part = g_d if i == 0 else torch.zeros_like(g_d)
        _native.call("mt_create_triangle_vertices_colors_backward", sdf.data_ptr(), ng, nsh, deg,
                     vox.lower_corners.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V,
                     tv2.data_ptr(), col2[c0:].data_ptr(), campos[c0:].data_ptr(), c1 - c0,
                     gd_part.data_ptr(), g_col[c0:].data_ptr(), gg_d3.data_ptr(), gg_sh3.data_ptr(),
                     ws2.data_ptr(), ws2.numel(), i, None, 0, i, status.data_ptr(), 0, _stream())
    assert int(status) == 0
    torch.testing.assert_close(gg_d3, gg_d2, rtol=1e-5, atol=1e-6 * float(gg_d.abs().max()))
    torch.testing.assert_close(gg_sh3, gg_sh2, rtol=1e-5, atol=1e-6 * float(gg_sh.abs().max()))

    # FP64 oracle adjoint of the same frozen pipeline (vertex adjoint of the SH adjoint).
    vnp = ox.extract_voxels(spec.occupancy.cpu().numpy())
    tri = ox.tessellate_voxels(sdf.cpu().numpy(), world.cpu().numpy(), sh.cpu().numpy(),
                               vnp.neighbors, vnp.lower_corners, vnp.upper_corners, vnp.indices,
                               ls.cpu().numpy())
    g_sh_ref = orr.evaluate_spherical_harmonics_backward(tri.vertices, tri.spherical_harmonics,
                                                         cams.extrinsics.cpu().numpy(), deg,
                                                         g_col.cpu().numpy())
    _, gd_ref, gs_ref = ox.create_triangle_vertices_backward(tri.record, ng, nsh, None,
                                                             g_d.cpu().numpy(), g_sh_ref)
    got_sh = gg_sh2.cpu().numpy().astype(np.float64)
    assert np.abs(got_sh - gs_ref).max() <= 1e-3 * np.abs(gs_ref).max()
    got_d = gg_d2.cpu().numpy().astype(np.float64)
    assert np.abs(got_d - gd_ref).max() <= 1e-3 * np.abs(gd_ref).max()


def test_fused_adjoint_with_shared_origins(case):
    """Columns that share grid vertices (as after welding a multi-component grid) cannot be
    gathered: the rows are added with atomics instead. Overwriting, accumulating and flagged
    calls all equal the separate SH and vertex adjoints on the same columns."""
    spec, vox, world, sdf, sh, ls, cams = case
    dev = sdf.device
    lower = vox.lower_corners.clone()
    lower[0, 1:400:2] = lower[0, 0:399:2]  # 200 columns take their neighbour's origin
    lower[1, 5] = lower[1, 3]  # and one +x end is shared
    n, s, ng, nsh, deg = lower.shape[1], ls.shape[0], world.shape[0], sh.shape[0], spec.sh_degree
    off = _offsets(vox._replace(lower_corners=lower), sdf, ls)
    V = int(off[-1])
    b = cams.extrinsics.shape[0]
    f32 = dict(dtype=torch.float32, device=dev)
    campos = torch.empty((b, 3), **f32)
    _native.call("mt_camera_positions", cams.extrinsics.contiguous().data_ptr(), b, campos.data_ptr(), _stream())
    tv, td = torch.empty((V, 3), **f32), torch.empty(V, **f32)
    tsh, tt = torch.empty((nsh, V, 3), **f32), torch.empty(V, dtype=torch.uint8, device=dev)
    ov = torch.zeros(1, dtype=torch.int32, device=dev)
    _native.call("mt_create_triangle_vertices_forward", world.data_ptr(), sdf.data_ptr(), sh.data_ptr(),
                 ng, nsh, lower.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V,
                 tv.data_ptr(), td.data_ptr(), tsh.data_ptr(), tt.data_ptr(), ov.data_ptr(), _stream())
    col = torch.empty((b, V, 3), **f32)
    _native.call("mt_evaluate_sh_forward", tv.data_ptr(), V, tsh.data_ptr(), nsh, deg,
                 campos.data_ptr(), b, col.data_ptr(), _stream())
    gen = torch.Generator("cpu").manual_seed(12)
    g_d = torch.randn(V, generator=gen).to(dev)
    g_col = torch.randn((b, V, 3), generator=gen).to(dev)
    # The separate adjoints (the vertex adjoint scatters with atomics on shared origins).
    g_tsh = torch.zeros((nsh, V, 3), **f32)
    _native.call("mt_evaluate_sh_backward", tv.data_ptr(), V, tsh.data_ptr(), nsh, deg,
                 campos.data_ptr(), b, g_col.data_ptr(), g_tsh.data_ptr(), _stream())
    ws = _native.workspace(_native.query("mt_create_triangle_vertices_backward_workspace_bytes", ng), dev)
    want_d, want_sh = torch.zeros(ng, **f32), torch.zeros((nsh, ng, 3), **f32)
    _native.call("mt_create_triangle_vertices_backward", sdf.data_ptr(), ng, nsh, lower.data_ptr(), n,
                 ls.data_ptr(), s, off.data_ptr(), V, None, g_d.data_ptr(), g_tsh.data_ptr(), None,
                 want_d.data_ptr(), want_sh.data_ptr(), ws.data_ptr(), ws.numel(), _stream())
    ws2 = _native.workspace(_native.query("mt_create_triangle_vertices_colors_backward_workspace_bytes",
                                          ng, n, deg), dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)

    def fused(gd_out, gs_out, c0, c1, gd_in, accumulate, touched=None, maps_ready=0):
        _native.call("mt_create_triangle_vertices_colors_backward", sdf.data_ptr(), ng, nsh, deg,
                     lower.data_ptr(), n, ls.data_ptr(), s, off.data_ptr(), V, tv.data_ptr(),
                     col[c0:].data_ptr(), campos[c0:].data_ptr(), c1 - c0, gd_in.data_ptr(),
                     g_col[c0:].contiguous().data_ptr(), gd_out.data_ptr(), gs_out.data_ptr(),
                     ws2.data_ptr(), ws2.numel(), accumulate,
                     None if touched is None else touched.data_ptr(), 0, maps_ready,
                     status.data_ptr(), 0, _stream())

    def close(got_d, got_sh):
        torch.testing.assert_close(got_d, want_d, rtol=1e-5, atol=1e-6 * float(want_d.abs().max()))
        torch.testing.assert_close(got_sh, want_sh, rtol=1e-5, atol=1e-6 * float(want_sh.abs().max()))

    # Overwriting call: NaN-filled outputs are cleared, then the rows added.
    gd1 = torch.full((ng,), float("nan"), **f32)
    gs1 = torch.full((nsh, ng, 3), float("nan"), **f32)
    fused(gd1, gs1, 0, b, g_d, 0)
    close(gd1, gs1)
    # Two accumulating calls over the cameras, reusing the maps.
    gd2, gs2 = torch.zeros(ng, **f32), torch.zeros((nsh, ng, 3), **f32)
    fused(gd2, gs2, 0, 1, g_d, 1, maps_ready=1)
    fused(gd2, gs2, 1, b, torch.zeros_like(g_d), 1, maps_ready=1)
    close(gd2, gs2)
    # All vertices flagged: the same sums through the flagged path.
    gd3, gs3 = torch.zeros(ng, **f32), torch.zeros((nsh, ng, 3), **f32)
    fused(gd3, gs3, 0, b, g_d, 1, touched=torch.ones(V, dtype=torch.uint8, device=dev), maps_ready=1)
    close(gd3, gs3)
    torch.cuda.synchronize()
    assert int(status) == 0
