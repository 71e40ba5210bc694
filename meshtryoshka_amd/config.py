"""Process-wide switches of the MI355X path.

* ``reference_gradient_quirks`` (env ``MESHTRYOSHKA_REFERENCE_QUIRKS=1``): the fused
  backward scatters each delayed sample gradient with the *next* triangle's barycentrics
  (intersect_and_composite.slang:452, 497-513) and the explicit composite backward writes
  delayed SDF gradients to layer s - 1 / S - 1 (composite.slang:197, 213), exactly as the
  reference does. Off by default: the kernels then compute the true adjoint.
* ``composite_chunk``: triangles staged in LDS per round by the tile kernels
  (0 = one per tile pixel, the reference's coupling). Set by the autotuner.
"""

import os

reference_gradient_quirks: bool = os.environ.get("MESHTRYOSHKA_REFERENCE_QUIRKS", "0") == "1"
composite_chunk: int = int(os.environ.get("MESHTRYOSHKA_COMPOSITE_CHUNK", "0"))


def set_reference_gradient_quirks(enabled: bool) -> None:
    global reference_gradient_quirks
    reference_gradient_quirks = bool(enabled)


def set_composite_chunk(chunk: int) -> None:
    global composite_chunk
    if not 0 <= int(chunk) <= 512:
        raise ValueError("composite chunk must be in [0, 512]")
    composite_chunk = int(chunk)
