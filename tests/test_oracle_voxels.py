"""Pins the oracle's rank formulation of the voxel structure against a literal
per-thread transcription of compute_vertex_occupancy / count_occupancy / create_voxels
(including their last-bit special cases), plus structural invariants."""

import numpy as np
import pytest

from oracle import extraction as ox


def _packed(shape, density, seed):
    rng = np.random.default_rng(seed)
    return ox.pack_occupancy(rng.random(shape) < density)


CASES = [((3, 2, 32), 0.4, 0), ((2, 3, 64), 0.3, 1), ((4, 4, 32), 1.0, 2), ((3, 3, 32), 0.0, 3),
         ((2, 2, 96), 0.6, 4), ((5, 1, 32), 0.5, 5)]


@pytest.mark.parametrize("shape,density,seed", CASES)
def test_rank_formulation_matches_bitserial(shape, density, seed):
    packed = _packed(shape, density, seed)
    a = ox.extract_voxels(packed)
    b = ox.extract_voxels_bitserial(packed)
    assert a.grid_shape == b.grid_shape
    np.testing.assert_array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
    for name in ("neighbors", "lower_corners", "upper_corners", "indices"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)


def test_word_edge_cells_match_bitserial():
    occ = np.zeros((3, 3, 64), dtype=bool)
    for (i, j, k) in [(0, 0, 31), (0, 0, 32), (2, 2, 63), (1, 0, 31), (2, 1, 0)]:
        occ[i, j, k] = True
    packed = ox.pack_occupancy(occ)
    a, b = ox.extract_voxels(packed), ox.extract_voxels_bitserial(packed)
    for name in ("neighbors", "lower_corners", "upper_corners", "indices"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)


def test_structure_invariants():
    packed = _packed((6, 5, 64), 0.3, 9)
    v = ox.extract_voxels(packed)
    I, J, K = v.grid_shape
    n_vox = v.num_voxels
    n_col = v.lower_corners.shape[1]
    assert n_col == v.vertices.shape[0]
    # Column origins are a permutation of the vertices.
    assert np.array_equal(np.sort(v.lower_corners[0]), np.arange(n_col))
    # Voxel linear indices decode to occupied cells, in increasing order.
    occ = ox.unpack_occupancy(packed)
    idx = v.indices.astype(np.int64)
    i, rem = np.divmod(idx, J * K)
    j, k = np.divmod(rem, K)
    assert occ[i, j, k].all() and occ.sum() == n_vox and (np.diff(idx) > 0).all()
    # Upper corners of a voxel sit at the expected offsets from its origin.
    pos = np.rint(v.vertices * np.array([K, J, I])).astype(np.int64)
    o = pos[v.lower_corners[0, :n_vox]]
    for c, off in zip(range(4), [(1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]):
        np.testing.assert_array_equal(pos[v.upper_corners[c]] - o, np.broadcast_to(off, o.shape))
    # Neighbor columns have their origin at the neighbor offset.
    offs = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
    for n, off in enumerate(offs):
        np.testing.assert_array_equal(pos[v.lower_corners[0, v.neighbors[n]]] - o,
                                      np.broadcast_to(off, o.shape))
    # Valid lower corners point one step along their axis.
    for a in range(3):
        ok = v.lower_corners[a + 1] >= 0
        d = pos[v.lower_corners[a + 1, ok]] - pos[v.lower_corners[0, ok]]
        e = np.zeros(3, np.int64)
        e[a] = 1
        np.testing.assert_array_equal(d, np.broadcast_to(e, d.shape))
    assert (v.lower_corners[1:, :n_vox] >= 0).all()


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0)
    occ = rng.random((3, 4, 64)) < 0.5
    assert np.array_equal(ox.unpack_occupancy(ox.pack_occupancy(occ)), occ)
