"""The scene checkpoint layout (the reference trainer's "model" entries plus a "scene" entry):
a written file reads back with the same tensors through weights_only loading, the SH bands
are split and joined as the reference's tesseract keeps them, and missing config is an
error."""

import pytest
import torch

from meshtryoshka_amd import serving


def _scene(nsh):
    gen = torch.Generator().manual_seed(0)
    occ = torch.randint(-2**31, 2**31 - 1, (4, 5, 2), generator=gen, dtype=torch.int64).to(torch.int32)
    sdf = torch.randn(100, generator=gen)
    sh = torch.randn(nsh, 100, 3, generator=gen)
    return occ, sdf, sh


@pytest.mark.parametrize("nsh,deg", [(16, 3), (1, 0)])
def test_round_trip(tmp_path, nsh, deg):
    occ, sdf, sh = _scene(nsh)
    path = tmp_path / "scene.ckpt"
    serving.save_scene_checkpoint(path, occ, sdf, sh, torch.tensor([0.1, 0.0, -0.1]), deg, 250.0,
                                  torch.tensor([1.0, 0.5, 0.25]), step=7)
    raw = torch.load(path, weights_only=True)
    assert raw["step"] == 7 and raw["model"]["tesseract.sh_rgb"].shape == (1, 100, 3)
    assert raw["model"]["tesseract.occupancy"].shape == (40,)
    c = serving.read_scene_checkpoint(path)
    assert torch.equal(c["occupancy"], occ) and torch.equal(c["signed_distances"], sdf)
    assert torch.equal(c["spherical_harmonics"], sh)
    assert c["sh_degree"] == deg and c["sharpness"] == 250.0
    assert torch.allclose(c["level_sets"], torch.tensor([0.1, 0.0, -0.1]))
    assert torch.equal(c["background"], torch.tensor([1.0, 0.5, 0.25]))


def test_reference_layout_needs_config(tmp_path):
    occ, sdf, sh = _scene(4)
    model = {"tesseract.occupancy": occ.reshape(-1), "tesseract.signed_distances_raw": sdf,
             "tesseract.sh_rgb": sh[:1], "tesseract.sh_rest": sh[1:]}
    path = tmp_path / "ref.ckpt"
    torch.save({"model": model, "step": 3}, path)  # no "scene" entry, as the reference writes
    with pytest.raises(ValueError):
        serving.read_scene_checkpoint(path)
    with pytest.raises(ValueError):
        serving.read_scene_checkpoint(path, grid_shape=(4, 5, 64))  # no sharpness
    c = serving.read_scene_checkpoint(path, grid_shape=(4, 5, 64), sharpness=100.0)
    assert c["sh_degree"] == 1 and torch.equal(c["spherical_harmonics"], sh)
    with pytest.raises(ValueError):  # two components' worth of words
        serving.read_scene_checkpoint(path, grid_shape=(4, 5, 32), sharpness=100.0)
