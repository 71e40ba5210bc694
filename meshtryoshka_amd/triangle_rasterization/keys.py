"""64-bit sort-key layouts (keys.py:4-22)."""

from typing import Literal

from .._native import KeyLayout


def create_key_defines(*sequence: tuple[Literal["camera", "shell", "tile", "depth"], int]) -> tuple[tuple[str, int], ...]:
    """MSB-first (field, bits) sequence -> ((NAME_BIT_OFFSET, o), (NAME_BIT_LENGTH, l), ...)."""
    offset = 0
    defines = []
    for key, length in reversed(sequence):
        if key == "depth" and length != 32:
            raise ValueError("Depth must have 32 bits.")
        upper = key.upper()
        defines.append((f"{upper}_BIT_OFFSET", offset))
        defines.append((f"{upper}_BIT_LENGTH", length))
        offset += length
    assert offset == 64
    return tuple(defines)


def to_layout(key_defines: tuple[tuple[str, int], ...]) -> KeyLayout:
    d = dict(key_defines)
    return KeyLayout(
        d["CAMERA_BIT_OFFSET"], d["CAMERA_BIT_LENGTH"], d["SHELL_BIT_OFFSET"], d["SHELL_BIT_LENGTH"],
        d["TILE_BIT_OFFSET"], d["TILE_BIT_LENGTH"], d["DEPTH_BIT_OFFSET"],
    )
