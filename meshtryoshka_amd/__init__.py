"""MI355X (gfx950) implementation of Meshtryoshka's extract -> rasterize -> composite path.

Subpackages ``triangle_extraction`` and ``triangle_rasterization`` keep the reference's
Python API; every op runs in the hand-written HIP library ``_lib/libmeshtryoshka_hip.so``
(C ABI in ``include/meshtryoshka_hip.h``). ``install_aliases()`` registers the two
subpackages under their reference top-level names so that unmodified callers
(``from triangle_extraction import tessellate_voxels``) resolve to this implementation.
"""

import sys

__version__ = "0.1.0"


def install_aliases() -> None:
    from . import triangle_extraction, triangle_rasterization

    sys.modules.setdefault("triangle_extraction", triangle_extraction)
    sys.modules.setdefault("triangle_rasterization", triangle_rasterization)
