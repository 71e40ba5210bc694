"""ctypes binding of ``libmeshtryoshka_hip.so`` (declared in include/meshtryoshka_hip.h).

The library is loaded lazily on first use, after torch (so that the HIP runtime torch
already loaded is the one the library binds to, and torch streams are valid handles).
There is no fallback: if the library is missing, or a tensor is not on a HIP device,
calls raise.
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libmeshtryoshka_hip.so"

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
c_p = ctypes.c_void_p
c_sz = ctypes.c_size_t


class KeyLayout(ctypes.Structure):
    _fields_ = [
        ("camera_offset", c_i32),
        ("camera_bits", c_i32),
        ("shell_offset", c_i32),
        ("shell_bits", c_i32),
        ("tile_offset", c_i32),
        ("tile_bits", c_i32),
        ("depth_offset", c_i32),
    ]


class CompositeArgs(ctypes.Structure):
    _fields_ = [
        ("xy", c_p),
        ("colors", c_p),
        ("sdf", c_p),
        ("num_vertices", c_i64),
        ("faces", c_p),
        ("sharpness", c_p),
        ("sorted_keys", c_p),
        ("sorted_values", c_p),
        ("tile_ranges", c_p),
        ("num_cameras", c_i32),
        ("num_shells", c_i32),
        ("grid_rows", c_i32),
        ("grid_cols", c_i32),
        ("tile_h", c_i32),
        ("tile_w", c_i32),
        ("height", c_i32),
        ("width", c_i32),
        ("layout", KeyLayout),
        ("chunk", c_i32),
        ("stats", c_p),
        # Camera table (cameras of different resolutions); all None / 0 when uniform.
        ("cam_dims", c_p),
        ("tile_base", c_p),
        ("pixel_base", c_p),
        ("total_tiles", c_i64),
        ("vertex_touched", c_p),
        # 0: colours f32[B, V, 3]; s > 0: f32[V, s, 3] (the batched step's layout).
        ("colors_stride", c_i32),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "mt_get_last_error_string": (ctypes.c_char_p, []),
    "mt_version": (c_i32, []),
    "mt_scan_workspace_bytes": (c_sz, [c_i64]),
    "mt_exclusive_scan_i32": (c_i32, [c_p, c_i64, c_p, c_sz, c_p]),
    "mt_exclusive_scan_i32_noalloc": (c_i32, [c_p, c_i64, c_p]),
    "mt_inclusive_scan_i32_to_i64": (c_i32, [c_p, c_p, c_i64, c_p, c_sz, c_p]),
    "mt_compute_vertex_occupancy": (c_i32, [c_p, c_i32, c_i32, c_i32, c_p, c_p, c_p]),
    "mt_count_occupancy": (c_i32, [c_p, c_i32, c_i32, c_i32, c_p, c_p]),
    "mt_create_voxels": (
        c_i32,
        [c_p, c_p, c_p, c_p, c_i32, c_i32, c_i32, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p],
    ),
    "mt_count_triangle_vertices": (c_i32, [c_p, c_p, c_i64, c_p, c_i32, c_p, c_p]),
    "mt_count_triangle_vertices_axes": (c_i32, [c_p, c_p, c_i64, c_p, c_i32, c_p, c_p, c_p]),
    "mt_create_triangle_vertices_forward": (
        c_i32,
        [c_p, c_p, c_p, c_i64, c_i32, c_p, c_i64, c_p, c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p,
         c_p],
    ),
    "mt_create_triangle_vertices_backward_workspace_bytes": (c_sz, [c_i64]),
    "mt_create_triangle_vertices_backward": (
        c_i32,
        [c_p, c_i64, c_i32, c_p, c_i64, c_p, c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
         c_sz, c_p],
    ),
    "mt_create_triangle_vertices_colors_forward": (
        c_i32,
        [c_p, c_p, c_p, c_i64, c_i32, c_i32, c_p, c_i64, c_p, c_i32, c_p, c_i64, c_p, c_i32, c_p,
         c_p, c_p, c_p, c_p, c_i32, c_i32, c_p, c_p],
    ),
    "mt_sh_row_floats": (c_i64, [c_i32]),
    "mt_pack_sh_rows": (c_i32, [c_p, c_i64, c_i32, c_i32, c_p, c_p]),
    "mt_create_triangle_vertices_colors_backward_workspace_bytes": (c_sz, [c_i64, c_i64, c_i32]),
    "mt_create_triangle_vertices_colors_backward": (
        c_i32,
        [c_p, c_i64, c_i32, c_i32, c_p, c_i64, c_p, c_i32, c_p, c_i64, c_p, c_p, c_p, c_i32, c_p,
         c_p, c_p, c_p, c_p, c_sz, c_i32, c_p, c_i32, c_i32, c_p, c_i32, c_p],
    ),
    "mt_count_triangle_faces": (c_i32, [c_p, c_p, c_i64, c_p, c_i64, c_p, c_i32, c_p, c_p, c_p]),
    "mt_write_occupancy": (c_i32, [c_p, c_i64, c_p, c_i64, c_i32, c_i32, c_p]),
    "mt_dilate_occupancy": (c_i32, [c_p, c_i32, c_i32, c_i32, c_i32, c_p, c_p]),
    "mt_upscale_occupancy": (c_i32, [c_p, c_i32, c_i32, c_i32, c_p, c_i32, c_i32, c_i32, c_p]),
    "mt_upscale_parameters": (
        c_i32, [c_p, c_p, c_i32, c_i32, c_i32, c_p, c_p, c_i64, c_i32, c_p, c_p, c_i32, c_i32,
                c_i32, c_p, c_p, c_i64, c_p]),
    "mt_quantized_max_abs": (c_i32, [c_p, c_i64, c_f64, c_p, c_p]),
    "mt_index_vertices_workspace_bytes": (c_sz, [c_i64]),
    "mt_index_vertices": (c_i32, [c_p, c_i64, c_f64, c_i64, c_p, c_p, c_sz, c_p]),
    "mt_index_keys": (c_i32, [c_p, c_i64, c_p, c_p, c_sz, c_p]),
    "mt_compute_sdf_regularizers_forward": (
        c_i32, [c_p, c_p, c_p, c_i64, c_p, c_i64, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p]),
    "mt_compute_sdf_regularizers_backward": (
        c_i32, [c_p, c_p, c_p, c_i64, c_p, c_i64, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p]),
    "mt_create_triangle_faces": (
        c_i32,
        [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p, c_i32, c_p, c_p, c_p, c_i64, c_i64, c_p, c_p],
    ),
    "mt_create_triangle_faces_i64": (
        c_i32,
        [c_p, c_p, c_i64, c_i64, c_p, c_p, c_p, c_i32, c_p, c_p, c_p, c_i64, c_i64, c_p, c_p],
    ),
    "mt_create_triangle_faces_axes_i64": (
        c_i32, [c_p, c_i64, c_i64, c_p, c_p, c_p, c_i32, c_p, c_p, c_i64, c_i64, c_p, c_p]
    ),
    "mt_create_triangle_faces_words_i64": (
        c_i32, [c_p, c_i64, c_i64, c_p, c_p, c_i32, c_p, c_p, c_i64, c_i64, c_p, c_p]
    ),
    "mt_project_vertices": (c_i32, [c_p, c_i64, c_p, c_p, c_i32, c_p, c_p, c_p]),
    "mt_project_vertices_live": (c_i32, [c_p, c_i64, c_p, c_p, c_p, c_i32, c_p, c_p, c_p]),
    "mt_bound_triangles": (
        c_i32,
        [c_p, c_p, c_i64, c_p, c_i64, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_f32,
         c_p, c_p, c_p, c_p, c_p],
    ),
    "mt_generate_keys": (
        c_i32,
        [c_p, c_i64, c_p, c_i64, c_p, c_i32, c_p, c_p, c_p, c_i32, c_i32, KeyLayout, c_p, c_p,
         c_i64, c_p, c_p],
    ),
    "mt_key_block_counts": (
        c_i32,
        [c_p, c_p, c_i64, c_p, c_i64, c_p, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_f32,
         c_p, c_i32, c_p, c_p],
    ),
    "mt_generate_keys_from_geometry": (
        c_i32,
        [c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32,
         c_i32, c_f32, c_p, c_i32, c_p, c_i32, KeyLayout, c_p, c_p, c_i64, c_p, c_p],
    ),
    "mt_generate_keys_single_pass_workspace_bytes": (c_sz, [c_i64, c_i32]),
    "mt_generate_keys_single_pass": (
        c_i32,
        [c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32,
         c_i32, c_f32, c_p, c_i32, c_i32, KeyLayout, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_sz,
         c_p],
    ),
    "mt_sort_workspace_bytes": (c_sz, [c_i64]),
    "mt_sort_keys": (c_i32, [c_p, c_p, c_i64, c_p, c_p, c_p, c_sz, c_p, c_p]),
    "mt_sort_keys_high": (c_i32, [c_p, c_p, c_i64, c_p, c_p, c_p, c_sz, c_p, c_i32, c_p]),
    "mt_sort_key_ranges_workspace_bytes": (c_sz, [c_i64]),
    "mt_sort_key_ranges": (c_i32, [c_p, c_p, c_i64, c_p, c_i64, c_i32, c_p, c_p, c_p, c_sz, c_p]),
    "mt_delineate_tiles": (c_i32, [c_p, c_i64, KeyLayout, c_i32, c_i32, c_i32, c_i32, c_p, c_p, c_p,
                                   c_p]),
    "mt_camera_positions": (c_i32, [c_p, c_i32, c_p, c_p]),
    "mt_evaluate_sh_forward": (c_i32, [c_p, c_i64, c_p, c_i32, c_i32, c_p, c_i32, c_p, c_p]),
    "mt_evaluate_sh_backward": (c_i32, [c_p, c_i64, c_p, c_i32, c_i32, c_p, c_i32, c_p, c_p, c_p]),
    "mt_intersect_and_composite_forward": (c_i32, [ctypes.POINTER(CompositeArgs), c_p, c_p]),
    "mt_intersect_and_composite_backward": (
        c_i32,
        [ctypes.POINTER(CompositeArgs), c_p, c_p, c_p, c_p, c_p, c_i32, c_p],
    ),
    "mt_image_l2_loss": (c_i32, [c_p, c_p, c_p, c_i32, c_i32, c_i32, c_p, c_i64, c_i64, c_p, c_p,
                                 c_p]),
    "mt_intersect": (
        c_i32,
        [c_p, c_p, c_p, c_i64, c_p, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_p,
         c_p, c_p],
    ),
    "mt_interpolate_forward": (
        c_i32,
        [c_p, c_p, c_p, c_p, c_p, c_i64, c_i32, c_i32, c_i32, c_i32, c_p, c_p, c_p],
    ),
    "mt_interpolate_backward": (
        c_i32,
        [c_p, c_p, c_p, c_i64, c_i32, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p],
    ),
    "mt_composite_forward": (c_i32, [c_p, c_p, c_p, c_i64, c_i32, c_p, c_p, c_p, c_p]),
    "mt_composite_backward": (
        c_i32,
        [c_p, c_p, c_p, c_i64, c_i32, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i32, c_p],
    ),
}

_lock = threading.Lock()
_lib = None


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load (once) and return the library with every C-ABI signature declared."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"{p} is missing: build it with `python -m meshtryoshka_amd.build` "
                "(this package has no CPU fallback)"
            )
        lib = ctypes.CDLL(str(p), mode=os.RTLD_NOW | os.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:  # a stale build; calling it raises in call()
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    lib = load_library()
    fn = getattr(lib, name, None)
    if fn is None:
        raise RuntimeError(f"{name} is missing from {LIB_PATH}: rebuild the library")
    rc = fn(*args)
    if rc != 0:
        msg = lib.mt_get_last_error_string().decode(errors="replace")
        raise RuntimeError(f"{name} failed (code {rc}): {msg}")


def query(name: str, *args):
    return getattr(load_library(), name)(*args)


def require_device(*tensors) -> None:
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError(
                "meshtryoshka_amd runs on the GPU only: got a tensor on "
                f"{t.device}; move inputs to a HIP device"
            )


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=device)
