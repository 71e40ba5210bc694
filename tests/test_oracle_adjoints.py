"""Pins the oracle adjoints against torch float64 autograd of differentiable
restatements of the forward (frozen discrete decisions), and against central finite
differences through the whole frozen-decision pipeline."""

import numpy as np
import pytest
import torch

from oracle import extraction as ox
from oracle import rasterization as orr
from tests.helpers import OracleScene, small_sphere_spec

F64 = torch.float64


@pytest.fixture(scope="module")
def scene():
    return OracleScene(small_sphere_spec(resolution=32, image=48, sh_degree=1))


def torch_fused_composite(hits, colors, sdf, k, faces, num_pixels, uvw=None):
    """Differentiable FP64 front-to-back compositing along frozen hit lists."""
    u, v, w = (hits.u, hits.v, hits.w) if uvw is None else uvw
    u, v, w = (torch.tensor(np.asarray(q, np.float64)) for q in (u, v, w))
    fa, fb, fc = (torch.tensor(faces[hits.face, i].astype(np.int64)) for i in range(3))
    pix = torch.tensor(hits.pixel)
    hw = num_pixels // colors.shape[0]
    cam = pix // hw
    s_in = u * sdf[fa] + v * sdf[fb] + w * sdf[fc]
    col = u[:, None] * colors[cam, fa] + v[:, None] * colors[cam, fb] + w[:, None] * colors[cam, fc]
    rank = torch.tensor(orr._rank_within_pixel(hits.pixel))
    rgb = torch.zeros((num_pixels, 3), dtype=F64)
    T = torch.ones(num_pixels, dtype=F64)
    outer = torch.full((num_pixels,), 1e10, dtype=F64)
    for r in range(int(rank.max()) + 1):
        sel = rank == r
        p = pix[sel]
        ci = torch.sigmoid(k * s_in[sel])
        co = torch.sigmoid(k * outer[p])
        alpha = torch.clamp((co - ci + 1e-5) / (co + 1e-5), 0, 1)
        rgb = rgb.index_put((p,), rgb[p] + col[sel] * T[p][:, None] * alpha[:, None])
        T = T.index_put((p,), T[p] * (1 - alpha))
        outer = outer.index_put((p,), s_in[sel])
    return rgb, T


def test_fused_adjoint_matches_autograd(scene):
    sc = scene
    t = sc.tri
    P = sc.B * sc.H * sc.W
    colors = torch.tensor(sc.colors.astype(np.float64), requires_grad=True)
    sdf = torch.tensor(t.signed_distances.astype(np.float64), requires_grad=True)
    k = torch.tensor(sc.spec.sharpness, dtype=F64, requires_grad=True)
    rgb, T = torch_fused_composite(sc.hits, colors, sdf, k, t.faces, P)
    rng = np.random.default_rng(0)
    g = rng.standard_normal((sc.B, 4, sc.H, sc.W))
    gr = torch.tensor(g[:, :3].transpose(0, 2, 3, 1).reshape(P, 3))
    gt = torch.tensor(g[:, 3].reshape(P))
    ((rgb * gr).sum() + (T * gt).sum()).backward()
    out = np.empty((sc.B, 4, sc.H, sc.W))
    out[:, :3] = rgb.detach().numpy().reshape(sc.B, sc.H, sc.W, 3).transpose(0, 3, 1, 2)
    out[:, 3] = T.detach().numpy().reshape(sc.B, sc.H, sc.W)
    cg, sg, kg = orr.fused_backward(sc.hits, sc.colors.astype(np.float64),
                                    t.signed_distances.astype(np.float64), t.faces,
                                    sc.spec.sharpness, out, g)
    # The suffix trick divides by (1 - alpha + 1e-6): relative deviation ~1e-6 / (1 - alpha).
    for a, b in ((cg, colors.grad.numpy()), (sg, sdf.grad.numpy())):
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-4
    assert abs(kg - float(k.grad)) / abs(float(k.grad)) < 1e-4


def test_reference_quirk_moves_but_preserves_gradient_mass(scene):
    sc = scene
    t = sc.tri
    g = np.random.default_rng(1).standard_normal(sc.image.shape)
    args = (sc.hits, sc.colors.astype(np.float64), t.signed_distances.astype(np.float64), t.faces,
            sc.spec.sharpness, sc.image.astype(np.float64), g)
    cg0, sg0, kg0 = orr.fused_backward(*args, reference_quirk=False)
    cg1, sg1, kg1 = orr.fused_backward(*args, reference_quirk=True)
    assert np.abs(sg0 - sg1).max() > 1e-6  # the quirk does change where gradient goes
    # Barycentrics sum to ~1, so the total per sample is (nearly) unchanged.
    assert abs(sg0.sum() - sg1.sum()) < 1e-3 * np.abs(sg0).sum()
    assert kg0 == kg1


def test_composite_adjoint_matches_autograd():
    rng = np.random.default_rng(2)
    P, S, k = 300, 5, 25.0
    sdf = rng.standard_normal((P, S)) * 0.05
    col = rng.random((P, S, 3))
    mask = rng.random((P, S)) < 0.7
    ts = torch.tensor(sdf, requires_grad=True)
    tc = torch.tensor(col, requires_grad=True)
    tk = torch.tensor(k, dtype=F64, requires_grad=True)
    rgb = torch.zeros((P, 3), dtype=F64)
    T = torch.ones(P, dtype=F64)
    outer = torch.full((P,), 1e10, dtype=F64)
    m = torch.tensor(mask)
    for s in range(S):
        ci = torch.sigmoid(tk * ts[:, s])
        co = torch.sigmoid(tk * outer)
        a = torch.clamp((co - ci + 1e-5) / (co + 1e-5), 0, 1)
        a = torch.where(m[:, s], a, torch.zeros_like(a))
        rgb = rgb + tc[:, s] * T[:, None] * a[:, None]
        T = T * (1 - a)
        outer = torch.where(m[:, s], ts[:, s], outer)
    gc = rng.standard_normal((P, 3))
    gt = rng.standard_normal(P)
    ((rgb * torch.tensor(gc)).sum() + (T * torch.tensor(gt)).sum()).backward()
    oc, ot = orr.composite(sdf, col, mask, k, dtype=np.float64)
    ws, wc, wk = orr.composite_backward(sdf, col, mask, k, oc, ot, gc, gt)
    assert np.abs(ws - ts.grad.numpy()).max() / np.abs(ts.grad.numpy()).max() < 1e-4
    assert np.abs(wc - tc.grad.numpy()).max() / np.abs(tc.grad.numpy()).max() < 1e-4
    assert abs(wk - float(tk.grad)) / abs(float(tk.grad)) < 1e-4
    # The reference's flush targets differ only where a masked layer sits in between.
    ws_q, _, _ = orr.composite_backward(sdf, col, mask, k, oc, ot, gc, gt, reference_quirk=True)
    gaps = np.array([np.any(np.diff(np.nonzero(r)[0]) > 1) or (r.any() and not r[-1]) for r in mask])
    np.testing.assert_allclose(ws_q[~gaps], ws[~gaps], rtol=1e-12, atol=1e-15)
    assert np.abs(ws_q[gaps] - ws[gaps]).max() > 0


def test_sh_adjoint_matches_autograd():
    rng = np.random.default_rng(3)
    V, deg = 200, 3
    verts = rng.uniform(-1, 1, (V, 3))
    sh = rng.standard_normal((16, V, 3)) * 0.3
    E = np.stack([np.eye(4), np.eye(4)])
    E[0, :3, 3] = [0.1, -0.2, 3.0]
    E[1, :3, 3] = [-0.5, 0.3, 2.5]
    g = rng.standard_normal((2, V, 3))
    tsh = torch.tensor(sh, requires_grad=True)
    cams = torch.tensor(orr.camera_positions(E))
    d = torch.tensor(verts)[None] - cams[:, None]
    d = d / d.norm(dim=-1, keepdim=True)
    basis = torch.tensor(orr.sh_basis(d.numpy(), deg))
    pre = torch.einsum("bvk,kvc->bvc", basis, tsh) + 0.5
    (torch.clamp(pre, min=0) * torch.tensor(g)).sum().backward()
    want = orr.evaluate_spherical_harmonics_backward(verts, sh, E, deg, g)
    np.testing.assert_allclose(want, tsh.grad.numpy(), atol=1e-12)


def test_sh_basis_is_orthonormal():
    # Monte Carlo check of the real SH basis (with the reference's sign convention).
    rng = np.random.default_rng(4)
    d = rng.standard_normal((400_000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = orr.sh_basis(d, 3)
    gram = 4 * np.pi * (Y.T @ Y) / d.shape[0]
    assert np.abs(gram - np.eye(16)).max() < 0.03


def test_frozen_pipeline_finite_differences():
    """C2(a)-style check in FP64: central differences (h = 1e-6) on grid SDF and SH entries
    through vertex interpolation (t frozen), SH evaluation and compositing (hits frozen),
    against the chained oracle adjoints."""
    sc = OracleScene(small_sphere_spec(resolution=32, image=32, sh_degree=1,
                                       level_sets=(0.06, 0.0, -0.06)))
    t = sc.tri
    rec = t.record
    rng = np.random.default_rng(5)
    P = sc.B * sc.H * sc.W
    g = rng.standard_normal((P, 4))
    uvw = (sc.hits.u.astype(np.float64), sc.hits.v.astype(np.float64), sc.hits.w.astype(np.float64))
    k = sc.spec.sharpness

    def loss(grid_sdf, grid_sh):
        tt = rec.t.astype(np.float64)
        tsdf = grid_sdf[rec.i0] + tt * (grid_sdf[rec.i1] - grid_sdf[rec.i0])
        tsh = grid_sh[:, rec.i0] + tt[None, :, None] * (grid_sh[:, rec.i1] - grid_sh[:, rec.i0])
        col, _ = orr.evaluate_spherical_harmonics(t.vertices, tsh, sc.E, 1)
        rgb, T = orr.composite_hits(sc.hits, col, tsdf, t.faces, k, P, dtype=np.float64, uvw=uvw)
        return float((rgb * g[:, :3]).sum() + (T * g[:, 3]).sum()), col, tsdf, tsh, rgb, T

    gs0 = sc.sdf.astype(np.float64)
    gh0 = sc.sh.astype(np.float64)
    _, col, tsdf, tsh, rgb, T = loss(gs0, gh0)
    out = np.empty((sc.B, 4, sc.H, sc.W))
    out[:, :3] = rgb.reshape(sc.B, sc.H, sc.W, 3).transpose(0, 3, 1, 2)
    out[:, 3] = T.reshape(sc.B, sc.H, sc.W)
    gimg = np.empty_like(out)
    gimg[:, :3] = g[:, :3].reshape(sc.B, sc.H, sc.W, 3).transpose(0, 3, 1, 2)
    gimg[:, 3] = g[:, 3].reshape(sc.B, sc.H, sc.W)
    cg, sg, _ = orr.fused_backward(sc.hits, col, tsdf, t.faces, k, out, gimg, uvw=uvw)
    shg = orr.evaluate_spherical_harmonics_backward(t.vertices, tsh, sc.E, 1, cg)
    _, grid_sdf_grad, grid_sh_grad = ox.create_triangle_vertices_backward(
        rec, gs0.shape[0], gh0.shape[0], grad_d=sg, grad_sh=shg)
    h = 1e-6
    touched = np.unique(np.concatenate([rec.i0, rec.i1]))
    for _ in range(25):
        i = rng.choice(touched)
        a, b = gs0.copy(), gs0.copy()
        a[i] += h
        b[i] -= h
        fd = (loss(a, gh0)[0] - loss(b, gh0)[0]) / (2 * h)
        assert abs(fd - grid_sdf_grad[i]) <= 1e-3 * max(abs(fd), 1e-3), (i, fd, grid_sdf_grad[i])
    for _ in range(25):
        i = rng.choice(touched)
        c, ch = rng.integers(0, 4), rng.integers(0, 3)
        a, b = gh0.copy(), gh0.copy()
        a[c, i, ch] += h
        b[c, i, ch] -= h
        fd = (loss(gs0, a)[0] - loss(gs0, b)[0]) / (2 * h)
        assert abs(fd - grid_sh_grad[c, i, ch]) <= 1e-3 * max(abs(fd), 1e-3)
