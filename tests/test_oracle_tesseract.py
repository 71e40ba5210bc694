"""Pins the tesseract oracle (oracle/tesseract.py) and checks the product's host-side
tesseract logic (meshtryoshka_amd/tesseract.py) against it on the CPU.

Pins: hand-computed corner tables for the centre and an arm; the welded vertex count of a
fully occupied tesseract, which has a closed form only if every arm meets the centre and
its neighbouring arms face to face ((c+1)^3 + a ((c+1)^3 - (c-1)^3) for centre c and arms
a); hand-computed initializer offsets, shell level sets and uncontracted points, and the
contraction's continuity at the centre's boundary."""

import math

import numpy as np
import pytest
import torch

from meshtryoshka_amd import tesseract as ts
from meshtryoshka_amd.triangle_extraction.create_voxels import Voxels
from oracle import tesseract as ot
from oracle.extraction import extract_voxels, pack_occupancy
from oracle.subdivision import index_vertices


def _full_flat(c, a, e=0.5):
    return np.concatenate([pack_occupancy(np.ones(shape, bool)).reshape(-1)
                           for shape, _ in ot.component_table(c, a, e)])


def test_corner_tables_known_answers():
    comps = ot.component_table(32, 64, 0.5)
    assert [s for s, _ in comps] == [(32, 32, 32), (64, 32, 32), (64, 32, 32), (32, 64, 32),
                                     (32, 64, 32), (32, 32, 64), (32, 32, 64)]
    centre = comps[0][1]
    assert np.array_equal(centre[0, 0, 0], [0.25, 0.25, 0.25])
    assert np.array_equal(centre[1, 0, 1], [0.75, 0.25, 0.75])  # (i, j, k) -> (x, y, z) = (k, j, i)
    # The arm along grid axis 0 (z) on the low side: its i = 0 face is the unit cube's
    # z = 0 face, its i = 1 face the centre's z = 0.25 face.
    arm = comps[1][1]
    assert np.array_equal(arm[0, 1, 0], [0, 1, 0]) and np.array_equal(arm[0, 1, 1], [1, 1, 0])
    assert np.array_equal(arm[1, 0, 1], [0.75, 0.25, 0.25])
    # The arm along axis 2 (x) on the high side: its k = 1 face is the cube's x = 1 face.
    arm = comps[6][1]
    assert np.array_equal(arm[1, 1, 1], [1, 1, 1]) and np.array_equal(arm[1, 0, 0], [0.75, 0.25, 0.75])
    assert np.array_equal(ot.trilinear(np.float32([[0.5, 0.5, 0.0]]), comps[1][1]), [[0.5, 0.5, 0.0]])
    assert np.array_equal(ot.trilinear(np.float32([[0.5, 0.5, 1.0]]), comps[1][1]), [[0.5, 0.5, 0.25]])


@pytest.mark.parametrize("c,a", [(32, 32), (64, 32)])
def test_full_tesseract_welds_to_the_closed_form(c, a):
    vox, bounds = ot.extract_tesseract(_full_flat(c, a), c, a, 0.5)
    assert bounds[-1] == (c + 1) ** 3 + 6 * (c + 1) ** 2 * (a + 1)
    used = np.unique(np.concatenate([vox.lower_corners.ravel(), vox.upper_corners.ravel()]))
    used = used[used >= 0]
    assert used.size == (c + 1) ** 3 + a * ((c + 1) ** 3 - (c - 1) ** 3)
    assert vox.grid_shape == (c + 2 * a,) * 3
    assert np.array_equal(np.sort(vox.indices), np.arange(c**3 + 6 * a * c * c))


def test_single_grid_is_plain_extraction():
    rng = np.random.default_rng(3)
    occ = rng.random((32, 32, 32)) < 0.2
    flat = pack_occupancy(occ).reshape(-1)
    vox, bounds = ot.extract_tesseract(flat, 32, 0, 1.0)
    plain = extract_voxels(pack_occupancy(occ))
    assert np.array_equal(vox.vertices, plain.vertices)  # the unit corners map is the identity
    assert np.array_equal(vox.neighbors, plain.neighbors) and np.array_equal(vox.indices, plain.indices)
    assert np.array_equal(vox.lower_corners, plain.lower_corners)  # grid vertices are distinct


def test_offsets_shells_and_contraction_known_answers():
    v = np.float32([[0.5, 0.5, 0.5], [0.5, 0.5, 1.0], [0.0, 0.5, 0.5]])
    assert np.allclose(ot.sphere_offset(v, 0.25), [-0.25, 0.25, 0.25])
    assert np.allclose(ot.level_sets_from_shells([0.5, 0.9], 400.0), [0.0, math.log(9) / 400])
    e, r0, k = 0.5, 0.6, 1.5
    got = ot.uncontract(np.array([[0.6, 0.5, 0.5], [1.0, 0.5, 0.5], [0.5, 0.0, 0.5]]), r0, k, e)
    far = r0 * math.exp(2 * k)
    assert np.allclose(got, [[0.24, 0, 0], [far, 0, 0], [0, -far, 0]])
    # Continuous across the centre's boundary (Chebyshev radius e).
    inside, outside = ot.uncontract(np.array([[0.75 - 1e-9, 0.5, 0.5], [0.75 + 1e-9, 0.5, 0.5]]), r0, k, e)
    assert abs(inside[0] - outside[0]) < 1e-6 and abs(inside[0] - r0) < 1e-6


# ------------------------------------------------------------------------------------
# The product's host logic against the pinned oracle
# ------------------------------------------------------------------------------------

def test_schedules():
    sched = {0: 32, 2401: 64, 4801: 128}
    assert [ts.value_at_step(sched, s) for s in (0, 2400, 2401, 9999)] == [32, 32, 64, 128]
    assert ts.value_at_step({"0": 7}, 5) == 7 and ts.value_at_step(9, 123) == 9
    with pytest.raises(ValueError):
        ts.value_at_step({10: 1}, 5)


@pytest.mark.parametrize("disable", [False, True])
def test_components_match_the_oracle(disable):
    layout = ts.TesseractLayout({0: 32, 100: 64}, {0: 32}, 0.5, disable_tesseract=disable)
    got = layout.components(150)
    want = ot.component_table(64, 32, 0.5, disable)
    assert [c.shape for c in got] == [s for s, _ in want]
    rng = np.random.default_rng(0)
    pts = rng.random((1000, 3)).astype(np.float32)
    dyadic = (rng.integers(0, 65, (1000, 3)) / 64).astype(np.float32)
    for comp, (_, table) in zip(got, want):
        assert np.array_equal(comp.corners.numpy(), table.astype(np.float32))
        assert np.array_equal(comp.map(torch.from_numpy(dyadic)).numpy(),
                              ot.trilinear(dyadic, table).astype(np.float32))
        assert np.allclose(comp.map(torch.from_numpy(pts)).numpy(), ot.trilinear(pts, table), atol=2e-7)
    assert layout.extent(150) == 128


def _to_torch(v):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    return Voxels(t(v.vertices), t(v.neighbors), t(v.lower_corners), t(v.upper_corners),
                  t(v.indices), v.grid_shape)


def test_concatenation_matches_the_oracle():
    """The product's concatenation (torch, on the CPU) of oracle-extracted components, welded
    with the oracle's index_vertices, equals the oracle's tesseract extraction."""
    c, a = 32, 32
    layout = ts.TesseractLayout(c, a, 0.5)
    rng = np.random.default_rng(1)
    flat = np.concatenate([pack_occupancy(rng.random(s) < 0.15).reshape(-1)
                           for s, _ in ot.component_table(c, a, 0.5)])
    parts, cells = [], []
    for comp, grid in zip(layout.components(0), ts.split_occupancy(torch.from_numpy(flat), layout, 0)):
        v = _to_torch(extract_voxels(grid.numpy()))
        parts.append(v._replace(vertices=comp.map(v.vertices)))
        cells.append(comp.shape[0] * comp.shape[1] * comp.shape[2])
    got = ts.concatenate_voxels(parts, cells)
    first = torch.from_numpy(index_vertices(got.vertices.numpy(), 1e-6)).long()
    want, _ = ot.extract_tesseract(flat, c, a, 0.5)
    assert np.array_equal(got.vertices.numpy(), want.vertices)
    assert np.array_equal(got.neighbors.numpy(), want.neighbors)
    assert np.array_equal(got.indices.numpy(), want.indices)
    lower = got.lower_corners.long()
    assert bool((lower < 0).any())  # subvoxel columns with missing corners are exercised
    welded = torch.where(lower >= 0, first[lower.clamp(min=0)], lower)
    assert np.array_equal(welded.numpy(), want.lower_corners)
    assert np.array_equal(first[got.upper_corners.long()].numpy(), want.upper_corners)
    with pytest.raises(ValueError, match="words"):
        ts.split_occupancy(torch.from_numpy(flat[:-1]), layout, 0)


def test_sdf_parameterisation():
    v = torch.tensor([[0.5, 0.5, 0.5], [0.5, 0.5, 1.0], [0.1, 0.2, 0.3]])
    raw = torch.tensor([0.01, -0.02, 0.03])
    sphere = {"name": "sphere", "radius": 0.25}
    want = raw.double() + torch.from_numpy(ot.sphere_offset(v.numpy(), 0.25))
    got = ts.sdf_from_raw(raw, v, sphere, 5)
    assert torch.allclose(got.double(), want, atol=1e-7)
    assert torch.allclose(ts.raw_from_sdf(got, v, sphere, 5), raw, atol=1e-7)
    pc = {"name": "point_cloud", "interpolate_steps": 200, "num_frozen_steps": 100}
    assert torch.equal(ts.sdf_from_raw(raw, v, pc, 300), raw)
    with pytest.raises(ValueError, match="point cloud"):
        ts.sdf_from_raw(raw, v, pc, 299)
    assert torch.equal(ts.sdf_from_raw(raw, v, {"name": "none"}, 0), raw)
    with pytest.raises(ValueError, match="unknown initializer"):
        ts.sdf_from_raw(raw, v, {"name": "cube"}, 0)
    assert torch.allclose(ts.level_sets_from_shells([0.9, 0.5, 0.1], 400.0).double(),
                          torch.from_numpy(ot.level_sets_from_shells([0.9, 0.5, 0.1], 400.0)))


def test_uncontract_matches_the_oracle():
    rng = np.random.default_rng(2)
    pts = rng.random((2000, 3)).astype(np.float32)
    pts[:10] = [0.5, 0.5, 0.5]
    pts[10:20, 0] = 1.0
    for e, r0, k in ((0.5, 0.6, 1.5), (0.25, 1.0, 0.5)):
        got = ts.Contraction(r0, k, e).uncontract(torch.from_numpy(pts)).double().numpy()
        want = ot.uncontract(pts, r0, k, e)
        assert np.allclose(got, want, rtol=2e-6, atol=1e-6)
    flat = ts.Contraction(0.6, 1.5, 0.5, disable_exponential=True).uncontract(torch.tensor([[1.0, 0.5, 0.0]]))
    assert torch.allclose(flat, torch.tensor([[1.0, 0.0, -1.0]]) * math.exp(3.0))
