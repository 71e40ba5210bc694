"""Writes tests/golden/tesseract_c32_a32.ckpt: a trained-scene checkpoint in the reference
trainer's layout (trainer.py:111-121: "model", "optimizer", "step", "benchmarks"; the model
entries of tesseract.py:161-214 and model_meshtryoshka.py:92-96), with no extra entries.

The tesseract is a 32^3 centre (extent 0.5) with six 32-cell arms. Its occupancy marks the
cells of every component whose mapped centre lies within 0.018 of a sphere of radius 0.33
about (0.5, 0.5, 0.5), so the surface crosses the centre and all six arms. The raw SDF is
-0.08 plus noise: with the default sphere initializer (radius 0.25) the rendered field is
|xyz - 0.5| - 0.33 plus noise. SH degree 0, step 5000.
Run: python tests/golden/make_tesseract_fixture.py
"""

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import tesseract as ot  # noqa: E402
from oracle.extraction import extract_voxels, pack_occupancy  # noqa: E402

PATH = os.path.join(HERE, "tesseract_c32_a32.ckpt")
CENTER, ARM, EXTENT, STEP = 32, 32, 0.5, 5000


def build():
    rng = np.random.default_rng(7)
    words, nverts = [], 0
    for shape, table in ot.component_table(CENTER, ARM, EXTENT):
        i, j, k = shape
        ci, cj, ck = np.meshgrid(np.arange(i), np.arange(j), np.arange(k), indexing="ij")
        unit = np.stack([(ck + 0.5) / k, (cj + 0.5) / j, (ci + 0.5) / i], axis=-1).reshape(-1, 3)
        r = np.linalg.norm(ot.trilinear(unit, table) - 0.5, axis=1)
        occ = (np.abs(r - 0.33) < 0.018).reshape(i, j, k)
        packed = pack_occupancy(occ)
        nverts += extract_voxels(packed).vertices.shape[0]
        words.append(packed.reshape(-1))
    flat = np.concatenate(words)
    raw = (-0.08 + 0.002 * rng.standard_normal(nverts)).astype(np.float32)
    sh_rgb = (0.2 + 0.6 * rng.random((1, nverts, 3))).astype(np.float32)
    model = {
        "background": torch.tensor([1.0, 1.0, 1.0]),
        "step": torch.tensor([STEP], dtype=torch.int32),
        "tesseract.step": torch.tensor(STEP, dtype=torch.int32),
        "tesseract.occupancy": torch.from_numpy(flat),
        "tesseract.signed_distances_raw": torch.from_numpy(raw),
        "tesseract.sh_rgb": torch.from_numpy(sh_rgb),
        "tesseract.sh_rest": torch.empty([], dtype=torch.float32),
        "tessellator.step": torch.tensor([STEP], dtype=torch.int32),
    }
    return {"model": model, "optimizer": {"state": {}, "param_groups": []}, "step": STEP,
            "benchmarks": {}}


if __name__ == "__main__":
    content = build()
    torch.save(content, PATH)
    m = content["model"]
    print(f"wrote {PATH}: {m['tesseract.occupancy'].numel()} words, "
          f"{m['tesseract.signed_distances_raw'].numel()} vertices, "
          f"{os.path.getsize(PATH)} bytes")
