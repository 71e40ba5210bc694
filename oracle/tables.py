"""Marching-cubes tables, parsed from the same data header the HIP kernels include.

The data follows triangle_extraction/constants.slang:1-158. Sharing one data file
keeps the table single-sourced; it is pinned independently by
``tests/test_oracle_tables.py`` (every case lists exactly the edges whose end
corners straddle the iso value, under the edge geometry of MT_MC_EDGE_OWNER).
"""

from __future__ import annotations

import re
from functools import lru_cache
from pathlib import Path

import numpy as np

HEADER = Path(__file__).resolve().parent.parent / "meshtryoshka_amd" / "csrc" / "mc_tables.h"


def _parse(name: str, text: str) -> np.ndarray:
    m = re.search(name + r"\[[^\]]*\](?:\[[^\]]*\])?\s*=\s*\{(.*?)\};", text, re.S)
    if m is None:
        raise RuntimeError(f"table {name} not found in {HEADER}")
    return np.array([int(x) for x in re.findall(r"-?\d+", m.group(1))], dtype=np.int64)


@lru_cache(maxsize=1)
def load():
    text = HEADER.read_text()
    case_start = _parse("MT_MC_CASE_START", text)
    case_edges = _parse("MT_MC_CASE_EDGES", text)
    edge_owner = _parse("MT_MC_EDGE_OWNER", text).reshape(12, 2)
    assert case_start.shape == (257,) and case_edges.shape == (2460,)
    return case_start, case_edges, edge_owner


# Cube corner order used by count_triangle_faces.slang:29-38, as (x, y, z) offsets
# with x = k (bit axis), y = j, z = i.
CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))

# Neighbor slot -> (x, y, z) offset, create_voxels.slang:134-179.
NEIGHBOR_OFFSETS = {
    -1: (0, 0, 0),
    0: (1, 0, 0),
    1: (0, 1, 0),
    2: (0, 0, 1),
    3: (1, 1, 0),
    4: (0, 1, 1),
    5: (1, 0, 1),
    6: (1, 1, 1),
}
AXES = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def edge_corners() -> list[tuple[int, int]]:
    """Cube edge id -> (corner a, corner b) implied by MT_MC_EDGE_OWNER."""
    _, _, owner = load()
    out = []
    for e in range(12):
        o = NEIGHBOR_OFFSETS[int(owner[e, 0])]
        d = AXES[int(owner[e, 1])]
        a = CORNERS.index(o)
        b = CORNERS.index(tuple(o[i] + d[i] for i in range(3)))
        out.append((a, b))
    return out
