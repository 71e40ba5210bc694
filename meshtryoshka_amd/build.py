"""Ahead-of-time build of the gfx950 HIP library ``_lib/libmeshtryoshka_hip.so``.

Each ``csrc/*.hip`` file is compiled by hipcc into an object (in parallel), then the
objects are linked into one shared library. Flags:

* ``--offload-arch=gfx950``: MI355X only, no other targets.
* ``-ffp-contract=off``: no fused multiply-add, so float expressions round exactly as
  written (keys embed depth bits; hit decisions must match the CPU oracle bit for bit).
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and square root.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
* ``-munsafe-fp-atomics``: hardware float atomics (all targets are device-local,
  coarse-grained allocations).

Usage: ``python -m meshtryoshka_amd.build`` (or ``build()`` from Python).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
LIBDIR = PACKAGE / "_lib"
LIBNAME = "libmeshtryoshka_hip.so"
INCLUDE = PACKAGE.parent / "include"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    cmd = [HIPCC, *COMMON_FLAGS, f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objects = [objdir / (s.stem + ".o") for s in sources]
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(lambda so: _compile(so[0], so[1], verbose), zip(sources, objects)))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objects), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
