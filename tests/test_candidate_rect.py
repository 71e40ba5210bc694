"""The compositing kernels test a staged triangle only at the pixels of its candidate
rectangle (meshtryoshka_amd/csrc/geometry.h), and the key kernels of small, well-shaped
triangles only at the pixel centres near their bounding box (centre_range). Either is
exact only if every pixel the reference's barycentric test accepts
(intersect_and_composite.slang:65-145) lies inside. tests/native/rect_check.cpp checks
both by brute force over every pixel of random tiles, for random triangles of every kind
(sub-pixel, huge, slivers, collinear, repeated corners, far away), compiled for the host
with the kernels' float rules (-ffp-contract=off, IEEE division). Without its shape
certificate the bounding-box route would miss hits of slivers (hundreds per run): the
certificate is what the sweep exercises."""

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def rect_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("rect") / "rect_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    str(HERE / "native" / "rect_check.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_every_hit_lies_in_its_candidate_rectangle(rect_check, seed):
    res = subprocess.run([str(rect_check), "150000", str(seed)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    fields = dict(kv.split("=") for kv in res.stdout.split("\n")[-2].split())
    assert int(fields["outside"]) == 0
    assert int(fields["centre_outside"]) == 0
    assert int(fields["box_outside"]) == 0 and int(fields["box_none"]) > 5000
    assert int(fields["hits"]) > 100000  # the sample exercises plenty of accepted pixels
    # Both answers of the bounding-box route occur often, and so does its fallback.
    assert min(int(fields["centre_none"]), int(fields["centre_few"]),
               int(fields["centre_unknown"])) > 5000
    # Most triangles get a rectangle much smaller than the tile or region.
    assert int(fields["area"]) < 0.5 * int(fields["pixels"])
