"""The C-ABI library loads without a GPU, exports every entry point the header declares,
and the Python ops refuse CPU tensors (no CPU fallback exists)."""

import re
from pathlib import Path

import pytest
import torch

from meshtryoshka_amd import _native

HEADER = Path(__file__).resolve().parent.parent / "include" / "meshtryoshka_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    names = declared_functions()
    assert len(names) >= 29
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.SIGNATURES), "ctypes table and header disagree"


def test_version_and_workspace_queries():
    assert _native.query("mt_version") >= 100
    assert _native.query("mt_scan_workspace_bytes", 10**6) >= 256
    n = 10**6
    assert _native.query("mt_sort_workspace_bytes", n) >= 12 * n


def test_error_path_reports_message():
    # A negative length is rejected on the host before any launch.
    with pytest.raises(RuntimeError, match="negative"):
        _native.call("mt_exclusive_scan_i32", None, -1, None, 0, None)


def test_ops_refuse_cpu_tensors():
    from meshtryoshka_amd import triangle_extraction as te

    with pytest.raises(RuntimeError, match="GPU only"):
        te.compute_exclusive_cumsum(torch.zeros(8, dtype=torch.int32))


def test_aliases_expose_reference_package_names():
    import meshtryoshka_amd

    meshtryoshka_amd.install_aliases()
    import triangle_extraction
    import triangle_rasterization

    for name in ("tessellate_voxels", "extract_voxels", "CreateTriangleVertices", "Voxels",
                 "Triangles", "compute_exclusive_cumsum", "pack_occupancy"):
        assert hasattr(triangle_extraction, name)
    for name in ("render", "render_fused", "render_explicit", "rasterize", "TileGrid",
                 "IntersectAndComposite", "EvaluateSphericalHarmonics", "Interpolate", "Composite"):
        assert hasattr(triangle_rasterization, name)
