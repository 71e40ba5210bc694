"""Pins the marching-cubes tables semantically (no reference vectors exist for them)."""

import numpy as np

from oracle import tables


def test_table_shape_and_prefix_invariants():
    start, edges, owner = tables.load()
    assert start[0] == 0 and start[256] == 2460 and edges.size == 2460
    counts = np.diff(start)
    assert (counts >= 0).all() and (counts % 3 == 0).all() and counts.max() <= 15
    assert counts[0] == 0 and counts[255] == 0  # all-outside / all-inside cubes emit nothing
    assert ((edges >= 0) & (edges < 12)).all()
    assert set(owner[:, 0].tolist()) <= {-1, 0, 1, 2, 3, 4, 5}
    assert set(owner[:, 1].tolist()) <= {0, 1, 2}


def test_edges_connect_adjacent_corners():
    for a, b in tables.edge_corners():
        ca, cb = np.array(tables.CORNERS[a]), np.array(tables.CORNERS[b])
        assert np.abs(ca - cb).sum() == 1
    assert len(set(tables.edge_corners())) == 12


def test_every_case_lists_exactly_its_crossing_edges():
    start, edges, _ = tables.load()
    ec = tables.edge_corners()
    for code in range(256):
        listed = edges[start[code] : start[code + 1]]
        crossing = {e for e, (a, b) in enumerate(ec) if ((code >> a) & 1) != ((code >> b) & 1)}
        assert set(listed.tolist()) == crossing, code


def test_single_corner_cases_emit_one_triangle():
    start, _, _ = tables.load()
    counts = np.diff(start)
    for c in range(8):
        assert counts[1 << c] == 3 and counts[255 ^ (1 << c)] == 3
