"""Image tiling shared by the rasterizer ops. ``TileGrid`` keeps the reference's fields and
properties (triangle_rasterization/common.py:11-60), since callers construct it by keyword
and read ``grid_shape``, ``num_tiles_per_image`` and ``tile_numel``."""

from dataclasses import dataclass
from functools import wraps

from torch.profiler import record_function as _profiler_range


def ceildiv(numerator, denominator):
    """Integer division rounded towards positive infinity (for positive divisors)."""
    return -(-numerator // denominator)


@dataclass(frozen=True)
class TileGrid:
    """How a batch of images is cut into tiles.

    ``tile_shape`` and ``image_shape`` are (height, width) in pixels; a partial tile at the
    bottom or right edge counts as a whole tile. ``num_cameras`` images are rendered, each
    ``num_shells`` times (one layer per level-set shell)."""

    tile_shape: tuple[int, int]
    image_shape: tuple[int, int]
    num_cameras: int
    num_shells: int

    @property
    def grid_shape(self) -> tuple[int, int]:
        """Tiles per image as (rows, columns)."""
        (ih, iw), (th, tw) = self.image_shape, self.tile_shape
        return (ceildiv(ih, th), ceildiv(iw, tw))

    @property
    def num_tiles_per_image(self) -> int:
        rows, cols = self.grid_shape
        return rows * cols

    @property
    def tile_numel(self) -> int:
        """Pixels in one (whole) tile."""
        return self.tile_shape[0] * self.tile_shape[1]


def record_function(tag: str):
    """Decorator that runs the function inside a profiler range named ``tag`` (visible in
    torch and rocprof traces); the wrapper carries the function's name, docstring and
    signature."""

    def attach(fn):
        @wraps(fn)
        def ranged(*args, **kwargs):
            with _profiler_range(tag):
                return fn(*args, **kwargs)

        return ranged

    return attach
