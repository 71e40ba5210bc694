"""Committed golden vectors (tests/golden/make_golden.py): the oracle must reproduce them
bit for bit on CPU, and the HIP path must reproduce them on the GPU."""

import numpy as np
import pytest
import torch

from tests.golden import make_golden
from tests.helpers import to_dev

GOLD = np.load(make_golden.PATH)


def test_oracle_reproduces_golden_vectors():
    out = make_golden.compute()
    for name in GOLD.files:
        a, b = out[name], GOLD[name]
        if a.dtype.kind == "f":
            assert np.array_equal(np.asarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64),
                                  np.asarray(b).view(np.uint32 if b.dtype == np.float32 else np.uint64)), name
        else:
            assert np.array_equal(a, b), name


@pytest.mark.gpu
def test_hip_path_reproduces_golden_vectors():
    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import triangle_rasterization as tr

    g = GOLD
    vox = te.extract_voxels(to_dev(g["occupancy"]))
    tri = te.tessellate_voxels(to_dev(g["sdf"]), to_dev(g["grid_vertices"]), to_dev(g["sh"]),
                               vox.neighbors, vox.lower_corners, vox.upper_corners, vox.indices,
                               to_dev(g["level_sets"]))
    np.testing.assert_array_equal(tri.faces.cpu().numpy(), g["faces"])
    np.testing.assert_array_equal(tri.face_boundaries.cpu().numpy(), g["face_boundaries"])
    assert np.array_equal(tri.vertices.detach().cpu().numpy().view(np.uint32),
                          g["triangle_vertices"].view(np.uint32))
    img = tr.render(to_dev(g["extrinsics"]), to_dev(g["intrinsics"]), (40, 40), tri.vertices.detach(),
                    tri.faces, tri.face_boundaries, tri.signed_distances, tri.spherical_harmonics, 1,
                    torch.tensor(float(g["sharpness"]), device="cuda"), torch.ones(3, device="cuda"))
    want = np.clip(g["image"][:, :3] + g["image"][:, 3:4], 0, 1)
    assert np.abs(img.detach().cpu().numpy() - want).max() < 1e-4
