"""CPU oracle for the extract -> rasterize -> composite path.

TEST INFRASTRUCTURE ONLY. Nothing in the product (the ``meshtryoshka_amd``
package and its HIP library) imports, calls or links anything in this
directory. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

The oracle is a NumPy restatement of the reference semantics, written from the
reference's Slang/CUDA sources read as text (each function cites the file:line
it follows). It has two modes:

* FP32-exact (default): IEEE float32 elementwise arithmetic in the reference's
  operation order, no fused multiply-add. Integer outputs (cube codes, counts,
  cumsums, vertex/face lists, keys, tile ranges, hit lists) are bit-exact
  targets for the HIP kernels.
* FP64-continuous (``oracle.differentiable``): given the FP32 run's discrete
  decisions (crossings, hit lists), recompute the continuous quantities in
  float64 with torch autograd. This is the target for the image and gradient
  tolerances.

How each function is pinned (closed forms, invariants, brute force, known answers,
autograd) is listed in ``oracle/PINNING.md``; every function there has at least one
independent check in ``tests/``.
"""

from . import extraction, rasterization, tables  # noqa: F401
