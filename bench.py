#!/usr/bin/env python
"""Training-step benchmark of the extract -> rasterize -> composite path on one MI355X.

Default workload (sphere-field training step, 4 cameras): a sparse 1024^3 virtual grid
whose occupied cells lie within 3 cells of the zero set of min(64 seeded spheres, ground
plane y = -0.6); corner SDF = clean SDF + 0.01 N(0, 1); SH degree 3; 5 shells spaced 1/512
apart; 4 cameras at 1237x822 on the radius-2 upper hemisphere; sharpness 50. One step =
tessellate_voxels (5 level sets) + fused render of the 4 cameras + mean squared error
against a seeded U(0, 1) target + backward to the grid SDF and SH. Inputs are resident in
HBM before timing starts.

``--workload sphere-field-32cam-mixed``: 256 spheres, 7 shells, 32 cameras whose
resolutions are drawn from {320x240, 640x480, 1237x822, 1920x1080, 2473x1643}, rendered in
camera groups inside one captured step (a 512^3 virtual grid: ~16M occupied cells).

Prints one JSON line (metric, value = steps/s, roofline of the dominant kernel measured
with HIP events over the instrumented steps, CPU baseline of the oracle on a bounded
sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


# Preset scenes: option values that replace the defaults.
WORKLOADS = {
    "sphere-field-4cam": {},
    "sphere-field-32cam-mixed": {
        "spheres": 256, "resolution": 1024, "shells": 7, "cameras": 32, "camera_group": 4,
        "resolutions": ["320x240", "640x480", "1237x822", "1920x1080", "2473x1643"],
        "steps": 3, "warmup": 1,
    },
}


def parse_args(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--resolution", type=int, default=1024)
    p.add_argument("--spheres", type=int, default=64)
    p.add_argument("--cameras", type=int, default=4)
    p.add_argument("--height", type=int, default=822)
    p.add_argument("--width", type=int, default=1237)
    p.add_argument("--shells", type=int, default=5)
    p.add_argument("--resolutions", type=str, nargs="+", default=None, metavar="WxH",
                   help="draw each camera's resolution uniformly from these (mixed resolutions)")
    p.add_argument("--camera-group", type=int, default=0,
                   help="cameras rendered together inside the batched step (0 = all)")
    p.add_argument("--workload", choices=sorted(WORKLOADS), default="sphere-field-4cam",
                   help="preset scene sizes (explicit options given on the command line win)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-pair-budget", type=int, default=1_500_000_000,
                   help="(pixel, key) pairs the CPU baseline composites on its tile sample")
    p.add_argument("--cpu-slab", type=float, default=1 / 64,
                   help="fraction of the grid's z layers the CPU baseline runs on")
    p.add_argument("--breakdown", action="store_true", help="print the per-kernel table to stderr")
    p.add_argument("--no-autotune", action="store_true",
                   help="skip the autotuner and use --tile / --chunk")
    p.add_argument("--tile", type=int, nargs=2, default=(16, 16), metavar=("H", "W"))
    p.add_argument("--chunk", type=int, default=256)
    p.add_argument("--only-batched", action="store_true",
                   help="time only the batched step with --tile / --chunk (no plain path, no "
                        "autotuner, no CPU baseline): the run profiled per kernel")
    p.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    p.add_argument("--no-c3", action="store_true",
                   help="skip the 32-camera mixed-resolution leg that follows the default workload")
    p.add_argument("--c3-steps", type=int, default=3)
    p.add_argument("--only-serving", action="store_true",
                   help="run only the serving sweep (prints its JSON object)")
    p.add_argument("--no-mode-compare", action="store_true",
                   help="skip the fused-vs-explicit peak memory and gradient comparison")
    p.add_argument("--serve-spheres", type=int, nargs="+", default=[8, 16, 32, 64, 96, 128, 192, 256],
                   help="sphere count of each resident scene")
    p.add_argument("--serve-resolution", type=int, default=1024,
                   help="virtual grid of the served scenes (the training generator's 1024^3)")
    p.add_argument("--serve-colours", choices=("grid", "vertex"), default="grid",
                   help="resident scenes keep the grid SH and colour each batch from it "
                        "(grid), or keep every triangle vertex's SH (vertex)")
    p.add_argument("--serve-rates", type=float, nargs="+",
                   default=[50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0])
    p.add_argument("--serve-seconds", type=float, default=3.0)
    p.add_argument("--serve-budget-gib", type=int, default=96)
    p.add_argument("--serve-max-batch", type=int, default=16)
    p.add_argument("--serve-deadline", type=float, default=1.0,
                   help="requests not admitted within this many seconds of arrival are shed")
    p.add_argument("--calibrate-counters", action="store_true",
                   help="run only three elementwise kernels that each read exactly 1 GiB and "
                        "write exactly 1 GiB (the scale for rocprofv3's FETCH_SIZE / WRITE_SIZE)")
    p.add_argument("--probe", action="store_true",
                   help="build the scene, print its sizes and memory, and exit")
    args = p.parse_args(argv)
    given = {a.split("=")[0] for a in (sys.argv[1:] if argv is None else argv) if a.startswith("--")}
    for key, value in WORKLOADS[args.workload].items():
        if "--" + key.replace("_", "-") not in given:
            setattr(args, key, value)
    return args


def parse_resolutions(items):
    """'WxH' strings -> (H, W) tuples."""
    out = []
    for it in items:
        w, h = it.lower().split("x")
        out.append((int(h), int(w)))
    return out


# ------------------------------------------------------------------------------------
# Scene
# ------------------------------------------------------------------------------------


class Scene:
    pass


def build_scene(args, device):
    import torch

    from meshtryoshka_amd import batched
    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import workloads

    res = parse_resolutions(args.resolutions) if args.resolutions else None
    spec = workloads.sphere_field(seed=args.seed, num_spheres=args.spheres,
                                  resolution=args.resolution, num_cameras=args.cameras,
                                  num_shells=args.shells, image_shape=(args.height, args.width),
                                  device=device, resolutions=res)
    sc = Scene()
    sc.spec = spec
    sc.voxels = te.extract_voxels(spec.occupancy.to(device))
    world, sdf, sh = workloads.finish_fields(spec, sc.voxels.vertices)
    sc.world = world.to(device)
    sc.sdf = sdf.to(device).requires_grad_(True)
    sc.sh = sh.to(device).requires_grad_(True)
    sc.cams = spec.cameras.to(device)
    b = sc.cams.extrinsics.shape[0]
    sc.shapes = batched.camera_shapes(sc.cams.image_shape, b)
    sc.uniform = len(set(sc.shapes)) == 1
    sc.num_pixels = sum(h * w for h, w in sc.shapes)
    if sc.uniform:
        h, w = sc.shapes[0]
        sc.target = workloads.target_image(spec, (b, 3, h, w)).to(device)
    else:  # flat concatenation of the per-camera [3, H_b, W_b] planes
        sc.target = workloads.target_image(spec, (3 * sc.num_pixels,)).to(device)
    sc.level_sets = spec.level_sets.to(device)
    sc.sharpness = torch.tensor(spec.sharpness, device=device)
    sc.background = torch.tensor(spec.background, device=device)
    return sc


# ------------------------------------------------------------------------------------
# Plain per-camera path: the reference's public API, one render call per camera
# (host size syncs inside every call, no graph), loss averaged over cameras.
# ------------------------------------------------------------------------------------


def plain_per_camera_step(sc, mode="fused"):
    import torch

    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import triangle_rasterization as tr

    v = sc.voxels
    sc.sdf.grad = None
    sc.sh.grad = None
    tri = te.tessellate_voxels(sc.sdf, sc.world, sc.sh, v.neighbors, v.lower_corners,
                               v.upper_corners, v.indices, sc.level_sets)
    b = sc.cams.extrinsics.shape[0]
    loss = 0.0
    p = 0
    for c in range(b):
        h, w = sc.shapes[c]
        img = tr.render(sc.cams.extrinsics[c : c + 1], sc.cams.intrinsics[c : c + 1],
                        (h, w), tri.vertices.detach(), tri.faces, tri.face_boundaries,
                        tri.signed_distances, tri.spherical_harmonics, sc.spec.sh_degree,
                        sc.sharpness, sc.background, tile_shape=sc.spec.tile_shape,
                        near_plane=sc.spec.near_plane, mode=mode)
        if sc.uniform:
            tgt = sc.target[c : c + 1]
        else:
            tgt = sc.target[3 * p : 3 * (p + h * w)].view(1, 3, h, w)
        p += h * w
        # One mean over every camera's pixels.
        loss = loss + ((img - tgt) ** 2).sum() / (3 * sc.num_pixels)
    loss.backward()
    return loss


def compare_render_modes(sc, log):
    """Plain per-camera step in fused mode (backward recomputes the hits) and in explicit
    mode (stores 12 B per (camera, shell, pixel) of hit data): peak device memory of each
    step, its duration, and how far the explicit gradients are from the fused ones."""
    import torch

    out = {}
    grads = {}
    for mode in ("fused", "explicit"):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        t0 = time.perf_counter()
        loss = plain_per_camera_step(sc, mode=mode)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        peak = torch.cuda.max_memory_allocated() - base
        grads[mode] = (sc.sdf.grad.clone(), sc.sh.grad.clone())
        out[mode] = {"peak_step_gib": round(peak / 2**30, 3), "ms_per_step": round(1000 * dt, 1),
                     "loss": float(loss)}
        log(f"[modes] {mode}: peak {peak / 2**30:.2f} GiB above the resident scene, {1000 * dt:.1f} ms")
    (fs, fh), (es, eh) = grads["fused"], grads["explicit"]
    out["grad_sdf_rel_diff"] = float((es - fs).abs().max() / fs.abs().max())
    out["grad_sh_rel_diff"] = float((eh - fh).abs().max() / fh.abs().max())
    out["loss_abs_diff"] = abs(out["fused"]["loss"] - out["explicit"]["loss"])
    log(f"[modes] explicit vs fused gradients: sdf {out['grad_sdf_rel_diff']:.2e}, "
        f"sh {out['grad_sh_rel_diff']:.2e} (max-abs / max)")
    sc.sdf.grad = None
    sc.sh.grad = None
    return out


# ------------------------------------------------------------------------------------
# Batched path: autotuned, graph-captured step; an eager instrumented run of the same
# kernels gives the per-kernel durations (HIP events on the launch stream).
# ------------------------------------------------------------------------------------


class LaunchTimer:
    def __init__(self):
        import torch

        self.torch = torch
        self.marks = []  # (name, event)
        self.totals = {}
        self.counts = {}

    def start(self):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        self.marks.append(("start", ev))

    def hook(self, name):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        self.marks.append((name, ev))

    def collect(self):
        prev = None
        for name, ev in self.marks:
            if name != "start" and prev is not None:
                ms = prev.elapsed_time(ev)
                self.totals[name] = self.totals.get(name, 0.0) + ms
                self.counts[name] = self.counts.get(name, 0) + 1
            prev = ev
        self.marks = []

    def mean_ms(self, name):
        """Average duration of one launch."""
        return self.totals[name] / self.counts[name]

    def step_ms(self, name, steps):
        """Duration per step (all launches of the stage in one step)."""
        return self.totals[name] / steps


def make_batched(sc, tile_shape, chunk, camera_group=None, margin=1.15):
    from meshtryoshka_amd import batched

    cap = batched.measure_capacities(sc.voxels, sc.level_sets, sc.cams.extrinsics, sc.cams.intrinsics,
                                     sc.shapes, sc.world, sc.sdf.detach(), sc.sh.detach(),
                                     tile_shape=tile_shape, near_plane=sc.spec.near_plane,
                                     camera_group=camera_group).scaled(margin)
    return batched.BatchedTrainingStep(sc.voxels, sc.level_sets, sc.cams.extrinsics,
                                       sc.cams.intrinsics, sc.shapes, sc.sh.shape[0],
                                       sc.spec.sh_degree, sc.spec.sharpness, sc.spec.background,
                                       cap, tile_shape=tile_shape, chunk=chunk,
                                       near_plane=sc.spec.near_plane, camera_group=camera_group)


# ------------------------------------------------------------------------------------
# Roofline accounting (algorithmic bytes / flops per launch)
# ------------------------------------------------------------------------------------


def algorithmic_bytes(z):
    """Unique bytes each kernel stage must move per step at minimum (perfect on-chip reuse).

    ``G`` camera groups run one after another: stages that read per-vertex or per-face
    data once per group count it G times; ``P`` is the pixel count and ``TT`` the tile
    count summed over all cameras; ``K`` is the key count summed over all groups."""
    nb = z["NSH"] * 12
    n_lc, n_vox, vt, f, k, b, s = z["N_lc"], z["N_vox"], z["V_t"], z["F"], z["K"], z["B"], z["S"]
    g, px, tt = z.get("G", 1), z["P"], z["TT"]
    blocks = b * -(-f // 256)
    return {
        "mt_count_triangle_vertices": n_lc * 16 + n_lc * 4 + 4 * s * n_lc,
        "mt_exclusive_scan_i32": 8 * (s * n_lc + 1),  # vertex counts, in place
        "mt_create_triangle_vertices_forward": n_lc * 16 + n_lc * (12 + 4 + nb) + 4 * s * n_lc
        + vt * (12 + 4 + nb + 1),
        "mt_count_triangle_faces": n_vox * 32 + n_lc * 4 + s * n_vox * (4 + 1),
        "mt_create_triangle_faces": n_vox * (28 + 4) + s * n_vox * (1 + 8) + 4 * s * n_lc + vt + f * 16,
        # Batched faces: corners from the columns' crossing-axis bits, no voxel indices.
        "mt_count_triangle_vertices_axes": n_lc * 16 + n_lc * 4 + 5 * s * n_lc,
        "mt_create_triangle_faces_axes_i64": n_vox * 28 + s * n_vox * (1 + 8) + 5 * s * n_lc + f * 12,
        # Faces from one (first vertex << 3 | axes) word per (level set, column).
        "mt_create_triangle_faces_words_i64": n_vox * 28 + s * n_vox * (1 + 8) + 4 * s * n_lc + f * 12,
        "mt_project_vertices": g * vt * 12 + b * vt * 12,
        "mt_project_vertices_live": g * vt * 12 + b * vt * 12,  # live vertices only
        "mt_bound_triangles": f * 12 + b * vt * 12 + b * f * 20,
        "mt_inclusive_scan_i32_to_i64": 12 * s * n_vox,  # face counts -> int64 offsets
        "mt_key_block_counts": g * f * 12 + b * vt * 12 + 4 * blocks,
        "mt_generate_keys_from_geometry": g * f * 12 + b * vt * 12 + k * 12 + 8 * blocks,
        # One pass: faces once per group, projected vertices once, keys written once.
        "mt_generate_keys_single_pass": g * f * 12 + b * vt * 12 + k * 12,
        "mt_sort_keys": 8 * k + 24 * k * z.get("sort_passes", 8),
        "mt_delineate_tiles": 8 * k + 16 * tt,
        "mt_evaluate_sh_forward": vt * (12 + nb) + b * vt * 12,
        "mt_intersect_and_composite_forward": 96 * k + 16 * px + 16 * tt,
        "mt_image_l2_loss": 16 * px + 12 * px + 16 * px,
        # Gradients go only to the vertices of accepted hits (the flagged vertices, counted
        # in the last group): per camera colour (12 B) and SDF (4 B), read and written.
        "mt_intersect_and_composite_backward": 96 * k + 32 * px + 16 * tt
        + 2 * z.get("V_touched", vt) * (12 * b + 4 * g),
        "mt_evaluate_sh_backward": vt * (12 + nb) + b * vt * 12 + vt * nb,
        "mt_create_triangle_vertices_backward": n_lc * 16 + vt * (12 + 4 + nb) + 2 * n_lc * (4 + nb),
        # Fused vertex + colour kernels of the batched step (grid-vertex count ~ N_lc).
        # grid_sh read once, its SH rows written once (the row holds the nb used bytes).
        "mt_pack_sh_rows": 2 * n_lc * nb,
        # The first group also writes the vertices; later groups write colours only.
        "mt_create_triangle_vertices_colors_forward": g * (n_lc * (16 + 4 + 12 + nb) + 4 * s * n_lc)
        + vt * 17 + b * vt * 12 + 4 * s * n_lc,  # + the column words (first group)
        # Columns, offsets and flags once per group; the flagged vertices' position, SDF and
        # per-camera colour / colour-gradient reads plus the clearing writes; the grid
        # gradients written once.
        "mt_create_triangle_vertices_colors_backward": g * (n_lc * (16 + 4) + 4 * s * n_lc + vt)
        + z.get("V_touched", vt) * (12 + 8 + 36 * b) + n_lc * (4 + nb),
        # Per group the vertex flags and tile ranges (the first group's adjoint writes the
        # grid gradients whole, so they are not cleared).
        "zero_fill": g * vt + 16 * tt,
    }


def sort_passes(keys):
    """Digit passes the radix sort executes: 8-bit digits over the key bits that vary."""
    import torch

    if keys.numel() == 0:
        return 0
    u = keys.view(torch.int64)
    varying = 0
    for b in range(64):
        bit = (u >> b) & 1
        if int(bit.max()) != int(bit.min()):
            varying += 1
    return (varying + 7) // 8


def workload_description(args, sc):
    if sc.uniform:
        h, w = sc.shapes[0]
        cams = f"{args.cameras} cameras {w}x{h}"
    else:
        counts = {}
        for h, w in sc.shapes:
            counts[f"{w}x{h}"] = counts.get(f"{w}x{h}", 0) + 1
        cams = f"{args.cameras} cameras of mixed resolution (" + ", ".join(
            f"{n}x {k}" for k, n in sorted(counts.items())) + ")"
    return (f"sphere-field training step: {args.resolution}^3 virtual grid, {args.spheres} spheres "
            f"+ ground plane, band 3 cells, SH deg 3, {args.shells} shells, {cams}")


def workload_key(args, best):
    """Names the committed profile summary of this workload and tile choice."""
    th, tw = best.tile_shape
    return f"{args.workload}_tile{th}x{tw}_chunk{best.chunk}"


def profile_traffic(entry_point, key):
    """HBM bytes per launch of ``entry_point`` from the committed rocprofv3 --pmc summary
    (profiles/batched_step_<key>.md, FETCH_SIZE and WRITE_SIZE scaled by the calibration
    kernel), or None when no summary of this workload exists."""
    path = os.path.join(ROOT, "profiles", f"batched_step_{key}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            # Rows are named by the entry point, possibly followed by a note in parentheses.
            if cells and (cells[0] == entry_point or cells[0].startswith(entry_point + " (")) \
                    and len(cells) >= 6:
                try:
                    return int(float(cells[5]) * 2**30)  # "HBM GiB per launch" column
                except ValueError:
                    return None
    return None


# Device kernel behind each entry point that can be the step's longest, for the counter lookup.
DOMINANT_KERNELS = {
    "mt_generate_keys_single_pass": "mt::keys_single_pass_kernel<true>",
    "mt_create_triangle_vertices_colors_forward": "mt::create_vertices_colors_kernel<3, true>",
    "mt_intersect_and_composite_backward": "mt::fused_backward_kernel",
    "mt_intersect_and_composite_forward": "mt::fused_forward_kernel",
}
# 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction.
VALU_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4


def profile_counter(kernel, counter, key):
    """A counter of ``kernel`` (largest dispatch) from the instruction-counter table of the
    committed profile of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"batched_step_{key}.md")
    if not kernel or not os.path.exists(path):
        return None
    header = None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells and cells[0] == "kernel" and counter in cells:
                header = cells
            elif header and cells and cells[0] == kernel and len(cells) == len(header):
                try:
                    return float(cells[header.index(counter)])
                except ValueError:
                    return None
    return None


# ------------------------------------------------------------------------------------
# SDF regularizers (the other per-step differentiable kernel of the reference's training
# step): forward and adjoint on the scene's voxels, HIP events on the launch stream
# ------------------------------------------------------------------------------------


def regularizer_benchmark(sc, cpu_sample: int = 0):
    """ms per call of mt_compute_sdf_regularizers_{forward,backward} over all voxels, their
    algorithmic GB/s (per voxel: 7 neighbour + 4 lower + 4 upper + 3 subvoxel corner indices
    read, 3 losses written / 3 upstream gradients read; per grid vertex: its SDF read once,
    and for the adjoint its gradient read and written once), and the oracle's single-thread
    time on a sample of cpu_sample voxels scaled to all of them (0: not timed)."""
    import torch

    from meshtryoshka_amd import _native

    v = sc.voxels
    n = int(v.upper_corners.shape[1])
    ncol = int(v.lower_corners.shape[1])
    ng = int(sc.sdf.shape[0])
    i, j, kp = sc.spec.occupancy.shape
    grid = (int(i), int(j), 32 * int(kp))
    sdf = sc.sdf.detach().contiguous()
    dev = sdf.device
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3)]
    gen = torch.Generator("cpu").manual_seed(3)
    ups = [torch.rand(n, generator=gen).to(dev) for _ in range(3)]
    grad = torch.zeros_like(sdf)
    common = (sdf.data_ptr(), v.neighbors.data_ptr(), v.lower_corners.data_ptr(), ncol,
              v.upper_corners.data_ptr(), n, grid[2], grid[1], grid[0])
    st = torch.cuda.current_stream(dev).cuda_stream

    def fwd():
        _native.call("mt_compute_sdf_regularizers_forward", *common,
                     *(o.data_ptr() for o in outs), st)

    def bwd():
        _native.call("mt_compute_sdf_regularizers_backward", *common,
                     *(u.data_ptr() for u in ups), grad.data_ptr(), st)

    def timed(fn, reps=10):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_f, ms_b = timed(fwd), timed(bwd)
    per_voxel = 4 * (7 + 4 + 4 + 3) + 12
    bytes_f = n * per_voxel + 4 * ng
    bytes_b = n * per_voxel + 12 * ng
    res = {"voxels": n, "grid_vertices": ng,
           "forward_ms": round(ms_f, 4), "backward_ms": round(ms_b, 4),
           "forward_GB/s": round(bytes_f / (ms_f * 1e-3) / 1e9, 1),
           "backward_GB/s": round(bytes_b / (ms_b * 1e-3) / 1e9, 1),
           "algorithmic_bytes": {"forward": int(bytes_f), "backward": int(bytes_b)}}
    res["forward_hbm_frac"] = round(res["forward_GB/s"] / HBM_PEAK_GBS, 4)
    res["backward_hbm_frac"] = round(res["backward_GB/s"] / HBM_PEAK_GBS, 4)
    if cpu_sample > 0:
        import numpy as np
        from threadpoolctl import threadpool_limits

        from oracle import regularizers as orz

        m = min(cpu_sample, n)
        nb = v.neighbors[:, :m].cpu().numpy()
        lo = v.lower_corners.cpu().numpy()
        up = v.upper_corners[:, :m].cpu().numpy()
        s_np = sdf.cpu().numpy()
        g_np = [u[:m].cpu().numpy() for u in ups]
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            orz.compute_sdf_regularizers(s_np, nb, lo, up, grid)
            orz.compute_sdf_regularizers_backward(s_np, nb, lo, up, grid, *g_np)
            sec = time.perf_counter() - t0
        res["cpu_oracle"] = {"seconds_per_fwd_bwd": round(sec * n / m, 3), "cores": 1,
                             "sample": f"NumPy oracle forward + adjoint on the first {m} of {n} "
                             f"voxels ({sec:.1f} s), scaled by voxels"}
    return res


# ------------------------------------------------------------------------------------
# Coarse-to-fine subdivision (the reference's schedule step, tesseract.py:436-481): mark the
# cells that hold triangles, dilate by 4, upscale the parameters to twice the resolution
# ------------------------------------------------------------------------------------


def subdivision_benchmark(device, resolution: int = 512, spheres: int = 64):
    """ms per call of write_occupancy / dilate_occupancy / upscale_occupancy /
    upscale_parameters on a sphere-field scene at ``resolution``^3 upscaled to twice that,
    with their algorithmic GB/s."""
    import torch

    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import workloads

    spec = workloads.sphere_field(seed=11, num_spheres=spheres, resolution=resolution,
                                  num_cameras=1, image_shape=(64, 64), device=device)
    occ = spec.occupancy.to(device)
    vox = te.extract_voxels(occ)
    _, sdf, sh = workloads.finish_fields(spec, vox.vertices)
    sdf, sh = sdf.to(device), sh.to(device)
    with torch.no_grad():
        tri = te.tessellate_voxels(sdf, vox.vertices, sh, vox.neighbors, vox.lower_corners,
                                   vox.upper_corners, vox.indices, spec.level_sets.to(device))
    i, j, kp = occ.shape
    target = (2 * i, 2 * j, 64 * kp)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            out = fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps, out

    faces_voxels = tri.voxel_indices
    ms_w, marked = timed(lambda: te.write_occupancy(faces_voxels, occ.shape, 0, 1))
    ms_d, dilated = timed(lambda: te.dilate_occupancy(marked, 4))
    ms_u, _ = timed(lambda: te.upscale_occupancy(occ & dilated, target))
    ms_p, up = timed(lambda: te.upscale_parameters(sdf, sh, occ, dilated, target), reps=3)
    words, words2 = occ.numel(), 8 * occ.numel()
    nsh = int(sh.shape[0])
    v_new = int(up.signed_distances.shape[0])
    bytes_ = {"write": 4 * faces_voxels.numel() + 4 * words, "dilate": 8 * words,
              "upscale_occupancy": 4 * words + 4 * words2,
              # vertex grids, old and new parameters; the vertex-occupancy and offset
              # passes of both grids are inside the call
              "upscale_parameters": 8 * words + 8 * words2 + 4 * (1 + 3 * nsh) * (
                  int(sdf.shape[0]) + v_new)}
    ms = {"write": ms_w, "dilate": ms_d, "upscale_occupancy": ms_u, "upscale_parameters": ms_p}
    return {"grid": [int(i), int(j), 32 * int(kp)], "target": list(target),
            "cells_marked": int(faces_voxels.numel()), "new_vertices": v_new,
            "ops": {k: {"ms": round(v, 4), "GB/s": round(bytes_[k] / (v * 1e-3) / 1e9, 1)}
                    for k, v in ms.items()}}


# ------------------------------------------------------------------------------------
# CPU baseline: the oracle on a bounded sample of the same step
# ------------------------------------------------------------------------------------


def cpu_threads() -> int:
    """Host threads for the CPU baseline: every core the process may use (OMP_NUM_THREADS
    caps it where set, as on a shared host)."""
    n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def c0_forward(device, log, runs: int = 5):
    """The dense 64^3 forward (S = 3, one 128 x 128 camera, 16 x 16 tiles, SH degree 0):
    extraction, tessellation, keys, sort, tile runs, SH and the fused composite, timed in
    full with the PyTorch-CPU transcription on the host threads and with the plain GPU path
    (1 warm-up, median of ``runs``); the two images are compared."""
    import statistics

    import torch

    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import triangle_rasterization as tr
    from meshtryoshka_amd import workloads
    from oracle import cpu_torch as ct

    spec = workloads.dense_sphere(seed=0, resolution=64, image_size=128)
    packed = spec.occupancy.cpu()
    vox0 = ct.extract_voxels(packed)
    world, sdf, sh = workloads.finish_fields(spec, vox0.vertices)
    world, sdf, sh = world.float(), sdf.float(), sh.float()
    ls = spec.level_sets.float()
    E, K = spec.cameras.extrinsics.float(), spec.cameras.intrinsics.float()
    shape = spec.cameras.image_shape

    def cpu_run():
        vox = ct.extract_voxels(packed)
        tess = ct.tessellate(sdf, world, sh, vox, ls)
        xy, depth = ct.project(tess.vertices, E, K)
        sk, sv, ranges = ct.front_end(xy, depth, tess.faces, tess.face_boundaries, spec.tile_shape,
                                      shape, spec.near_plane)
        col, _ = ct.sh_forward(tess.vertices, tess.sh, E, spec.sh_degree)
        res = ct.fused_composite(xy, tess.faces, tess.sdf, col, sk, sv, ranges, spec.tile_shape,
                                 shape, spec.sharpness)
        bg = torch.tensor(spec.background)[None, :, None, None]
        return (res.image[:, :3] + res.image[:, 3:4] * bg).clamp(0, 1), tess, sk

    dev = [t.to(device) for t in (world, sdf, sh, ls, E, K)]
    occ_d = spec.occupancy.to(device)
    sharp = torch.tensor(spec.sharpness, device=device)
    bg_d = torch.tensor(spec.background, device=device)

    def gpu_run():
        with torch.no_grad():
            vox = te.extract_voxels(occ_d)
            tri = te.tessellate_voxels(dev[1], dev[0], dev[2], vox.neighbors, vox.lower_corners,
                                       vox.upper_corners, vox.indices, dev[3])
            return tr.render(dev[4], dev[5], shape, tri.vertices, tri.faces, tri.face_boundaries,
                             tri.signed_distances, tri.spherical_harmonics, spec.sh_degree, sharp,
                             bg_d, tile_shape=spec.tile_shape, near_plane=spec.near_plane)

    def median_s(fn, sync=None):
        fn()
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            out = fn()
            if sync:
                sync()
            times.append(time.perf_counter() - t0)
        return statistics.median(times), out

    cpu_s, (cpu_img, tess, sk) = median_s(cpu_run)
    log(f"[bench] C0 forward on the CPU ({torch.get_num_threads()} threads): {1000 * cpu_s:.1f} ms")
    gpu_s, gpu_img = median_s(gpu_run, torch.cuda.synchronize)
    log(f"[bench] C0 forward on the GPU (plain path): {1000 * gpu_s:.2f} ms")
    return {"cpu_ms": round(1000 * cpu_s, 2), "gpu_ms": round(1000 * gpu_s, 3),
            "speedup": round(cpu_s / gpu_s, 1),
            "image_max_abs_cpu_vs_gpu": float((cpu_img - gpu_img.cpu()).abs().max()),
            "faces": int(tess.faces.shape[0]), "triangle_vertices": int(tess.vertices.shape[0]),
            "keys": int(sk.numel()),
            "workload": "dense 64^3 grid, S=3, one 128x128 camera, 16x16 tiles, SH degree 0"}


def cpu_baseline(sc, args, live, log):
    """The PyTorch-CPU transcription of the path (oracle/cpu_torch.py: FP32, vectorised,
    on the host's intra-op thread pool) timed on the host cores.

    * C0 forward in full (c0_forward), beside the plain GPU path.
    * The measured workload's training step on a bounded sample, each stage scaled to the
      full step by its own unit: a slab of the grid's z layers (extraction, tessellation
      and the vertex adjoint, scaled by voxels); projection..tile runs, SH and the SH
      adjoint on the slab's faces for every camera (scaled by faces); the fused composite
      forward + backward of the L2 loss on a random sample of the slab's non-empty tiles
      (scaled by the full step's unculled key count: the transcription, like the
      reference, composites every key of every face's tile box)."""
    import numpy as np
    import torch

    from oracle import cpu_torch as ct

    threads = cpu_threads()
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        c0 = c0_forward(sc.world.device, log)
        packed = sc.spec.occupancy.cpu()
        res = packed.shape[0]
        layers = max(2, int(round(res * args.cpu_slab)))
        i0 = (res - layers) // 2
        t0 = time.perf_counter()
        vox = ct.extract_voxels(packed[i0 : i0 + layers].contiguous(), i0, res)
        t_ext = time.perf_counter() - t0
        gen = torch.Generator("cpu").manual_seed(args.seed + 1)
        world, sdf, sh = (x.float().contiguous() for x in sc.spec.field_fn(vox.vertices, gen))
        ls = sc.level_sets.cpu()
        t0 = time.perf_counter()
        tess = ct.tessellate(sdf, world, sh, vox, ls)
        t_tess = time.perf_counter() - t0
        log(f"[bench] CPU sample: slab of {layers}/{res} layers, {vox.upper.shape[1]} voxels, "
            f"{tess.faces.shape[0]} faces; extraction {t_ext:.1f} s, tessellation {t_tess:.1f} s")
        E, K = sc.cams.extrinsics.cpu(), sc.cams.intrinsics.cpu()
        h, w = sc.cams.image_shape
        deg = sc.spec.sh_degree
        t0 = time.perf_counter()
        xy, depth = ct.project(tess.vertices, E, K)
        sk, sv, ranges = ct.front_end(xy, depth, tess.faces, tess.face_boundaries, sc.spec.tile_shape,
                                      (h, w), sc.spec.near_plane)
        col, pre = ct.sh_forward(tess.vertices, tess.sh, E, deg)
        t_front = time.perf_counter() - t0
        log(f"[bench] CPU sample: projection..tile runs and SH {t_front:.1f} s, {sk.numel()} keys")
        cnt = ranges[..., 1] - ranges[..., 0]
        nonempty = torch.nonzero(cnt > 0)
        rng = np.random.default_rng(0)
        pick = nonempty[torch.from_numpy(rng.permutation(nonempty.shape[0]))]
        # Tiles up to a budget of (pixel, key) pairs.
        th, tw = sc.spec.tile_shape
        budget = args.cpu_pair_budget
        take = int(torch.searchsorted(torch.cumsum(cnt[pick[:, 0], pick[:, 1]] * th * tw, 0),
                                      budget)) + 1
        pick = pick[: max(1, min(take, pick.shape[0]))]
        keys_sampled = int(cnt[pick[:, 0], pick[:, 1]].sum())
        t0 = time.perf_counter()
        comp = ct.fused_composite(xy, tess.faces, tess.sdf, col, sk, sv, ranges, sc.spec.tile_shape,
                                  (h, w), sc.spec.sharpness, sc.background.cpu(), sc.target.cpu(),
                                  tiles=pick)
        t_comp = time.perf_counter() - t0
        log(f"[bench] CPU sample: composite fwd+bwd on {pick.shape[0]} tiles ({keys_sampled} keys, "
            f"{comp.pairs} pixel-key pairs) {t_comp:.1f} s")
        t0 = time.perf_counter()
        g_sh = ct.sh_backward(tess.vertices, tess.sh, E, deg, pre, comp.colors_grad)
        t_shb = time.perf_counter() - t0
        t0 = time.perf_counter()
        ct.vertex_backward(tess, world.shape[0], comp.sdf_grad, g_sh)
        t_vb = time.perf_counter() - t0
    finally:
        torch.set_num_threads(prev)
    vox_ratio = sc.voxels.num_voxels / max(vox.upper.shape[1], 1)
    face_ratio = live["F"] / max(tess.faces.shape[0], 1)
    k_full = int(sk.numel() * face_ratio)
    t_comp_full = t_comp * k_full / max(keys_sampled, 1)
    step_s = (t_ext + t_tess + t_vb) * vox_ratio + (t_front + t_shb) * face_ratio + t_comp_full
    sample_s = t_ext + t_tess + t_front + t_comp + t_shb + t_vb
    return {
        "value": round(1.0 / step_s, 6),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"PyTorch-CPU transcription (FP32, vectorised, {threads} threads) on a slab of "
                   f"{layers}/{res} z layers of the same scene ({vox.upper.shape[1]} voxels, "
                   f"{tess.faces.shape[0]} faces, {sk.numel()} unculled keys for {E.shape[0]} "
                   f"cameras; {sample_s:.1f} s of CPU work): extraction + tessellation "
                   f"{t_ext + t_tess:.1f} s and vertex adjoint {t_vb:.1f} s scaled x{vox_ratio:.1f} "
                   f"by voxels; projection..tile runs + SH {t_front:.1f} s and SH adjoint "
                   f"{t_shb:.1f} s scaled x{face_ratio:.1f} by faces; fused composite fwd+bwd on "
                   f"{pick.shape[0]} random non-empty tiles ({keys_sampled} keys, {t_comp:.1f} s) "
                   f"scaled to the step's {k_full} unculled keys"),
        "seconds_per_step": round(step_s, 1),
        "c0_forward": c0,
    }


# ------------------------------------------------------------------------------------
# Serving: device-resident scenes, Poisson request stream, continuous batching
# ------------------------------------------------------------------------------------

SERVE_SHAPES = ((256, 256), (512, 512), (768, 1024), (1080, 1920))


def serving_benchmark(args, device, log):
    """Resident scenes of the sphere-field generator (one per entry of args.serve_spheres);
    requests = (scene, random hemisphere pose, resolution from SERVE_SHAPES), forward only.
    For each offered rate a Poisson stream runs for args.serve_seconds and then drains;
    requests not admitted within args.serve_deadline seconds are shed. Reports completed
    req/s, p50/p99 latency of the completed requests and the number shed."""
    import random

    import torch

    from meshtryoshka_amd import serving
    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import workloads

    spheres = list(args.serve_spheres)
    n = len(spheres)
    scenes, cells = {}, []
    cal_gen = torch.Generator("cpu").manual_seed(args.seed + 5)
    t0 = time.time()
    for sid, ns in enumerate(spheres):
        spec = workloads.sphere_field(seed=100 + sid, num_spheres=ns,
                                      resolution=args.serve_resolution, num_cameras=4,
                                      image_shape=(512, 512), device=device)
        vox = te.extract_voxels(spec.occupancy.to(device))
        world, sdf, sh = workloads.finish_fields(spec, vox.vertices)
        del spec.occupancy
        sc = serving.tessellate_scene(vox, world.to(device), sdf.to(device), sh.to(device),
                                      spec.level_sets.to(device), spec.sh_degree, spec.sharpness,
                                      spec.background, colours=args.serve_colours)
        cal = {}
        for shape in SERVE_SHAPES:
            poses = workloads.hemisphere_cameras(cal_gen, 4, shape)
            cal[shape] = (torch.stack([p[0] for p in poses]).to(device),
                          torch.stack([p[1] for p in poses]).to(device))
        serving.calibrate_keys(sc, cal)
        cells.append(int(vox.num_voxels))
        scenes[sid] = sc
        del vox, world, sdf, sh
        log(f"[serve] scene {sid} ({ns} spheres): {cells[-1]} cells, {sc.faces.shape[0]} faces, "
            f"{sc.resident_bytes() / 2**30:.1f} GiB resident")
    torch.cuda.synchronize()
    resident = torch.cuda.memory_allocated(device)
    log(f"[serve] {n} scenes ({min(cells)}..{max(cells)} cells) resident, "
        f"{resident / 2**30:.1f} GiB, built in {time.time() - t0:.1f} s")
    poses = {}
    gen = torch.Generator("cpu").manual_seed(args.seed + 7)
    for shape in SERVE_SHAPES:
        poses[shape] = workloads.hemisphere_cameras(gen, 64, shape)
    pick = random.Random(args.seed)

    def make_request(rid, arrival):
        shape = SERVE_SHAPES[pick.randrange(len(SERVE_SHAPES))]
        e, k, _ = poses[shape][pick.randrange(64)]
        return serving.Request(arrival=arrival, rid=rid, scene_id=pick.randrange(n),
                               extrinsics=e, intrinsics=k, image_shape=shape)

    # Warm-up: one request per (scene, shape) so first-use costs stay out of the sweep.
    server = serving.RenderServer(scenes, hbm_budget_bytes=args.serve_budget_gib << 30,
                                  max_batch=args.serve_max_batch, deadline_s=args.serve_deadline)
    for sid in scenes:
        for shape in SERVE_SHAPES:
            e, k, _ = poses[shape][0]
            server.submit(serving.Request(time.perf_counter(), 0, sid, e, k, shape))
    server.drain()
    torch.cuda.synchronize()
    sweep = []
    for rate in args.serve_rates:
        server.reset_stats()
        st = serving.poisson_load(server, make_request, rate, args.serve_seconds, seed=int(rate))
        log(f"[serve] offered {rate} req/s: {st}")
        sweep.append(st)
    return {
        "scenes": n, "spheres": spheres, "cells": cells, "resolution": args.serve_resolution,
        "colours": args.serve_colours,
        "shapes": [list(x) for x in SERVE_SHAPES], "hbm_budget_gib": args.serve_budget_gib,
        "max_batch": args.serve_max_batch, "deadline_s": args.serve_deadline,
        "resident_gib": round(resident / 2**30, 2),
        "sweep": sweep,
    }


# ------------------------------------------------------------------------------------
# The 32-camera mixed-resolution configuration (1024^3, 256 spheres, 7 shells), timed after
# the headline workload with its autotuned tile and key-chunk.
# ------------------------------------------------------------------------------------


def c3_leg(args, device, log, tile_shape, chunk):
    import copy

    import torch

    a = copy.copy(args)
    for key, value in WORKLOADS["sphere-field-32cam-mixed"].items():
        setattr(a, key, value)
    log("[c3] building scene (1024^3, 256 spheres, 7 shells, 32 cameras of mixed resolution)")
    t0 = time.time()
    sc = build_scene(a, device)
    torch.cuda.synchronize()
    log(f"[c3] scene ready in {time.time() - t0:.1f} s: {sc.voxels.num_voxels} voxels")
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    step = make_batched(sc, tile_shape, chunk, camera_group=a.camera_group)
    inputs = (sc.world, sc.sdf.detach(), sc.sh.detach(), sc.target)
    step.capture(*inputs)
    for _ in range(a.warmup):
        step.replay()
    torch.cuda.synchronize()
    log("[c3] captured; timing")
    t0 = time.perf_counter()
    for _ in range(args.c3_steps):
        step.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if step.overflowed():
        raise RuntimeError("C3 batched step overflowed its capacities")
    live = step.live_sizes()
    peak = torch.cuda.max_memory_allocated() - base
    ms = 1000 * elapsed / args.c3_steps
    log(f"[c3] batched graph step: {ms:.1f} ms/step, loss {float(step.loss):.6f}, {live}, "
        f"peak {peak / 2**30:.1f} GiB above the scene")
    timer = LaunchTimer()
    timer.start()
    step.run(*inputs, hook=timer.hook)
    torch.cuda.synchronize()
    timer.collect()
    table = {k: round(timer.step_ms(k, 1), 3) for k in sorted(timer.totals, key=lambda k: -timer.totals[k])
             if k.startswith("mt_")}
    out = {
        "workload": workload_description(a, sc),
        "steps": args.c3_steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 2), "steps_per_s": round(1000 / ms, 4),
        "tile_shape": list(tile_shape), "chunk": chunk, "camera_group": a.camera_group,
        "camera_groups": len(step.groups),
        "sizes": {"N_vox": int(sc.voxels.num_voxels), "V_t": live["V_t"], "F": live["F"],
                  "F_x3_over_2^31": round(3 * live["F"] / 2**31, 3), "K": live["K"],
                  "K_max_group": live["K_max_group"], "pixels": sc.num_pixels},
        "peak_gib_above_scene": round(peak / 2**30, 2),
        "kernels_ms": table,
        "plain_per_camera": {
            "status": "not run: does not fit in HBM",
            "reason": "the reference-structured step keeps every triangle vertex's SH "
                      f"({16 * 12 * live['V_t'] / 2**30:.0f} GiB) and each render call's SH "
                      "gradient of the same size plus its accumulation; the batched step "
                      "never stores per-vertex SH. Images of one camera of every resolution "
                      "are tested identical to the plain render (tests/test_gpu_config_c3_c4.py)",
        },
    }
    del step, sc, inputs
    import gc

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by camera, with one all-reduce of the grid gradients)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from meshtryoshka_amd import autotune

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.calibrate_counters:
        src = torch.ones(1 << 28, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
            torch.cuda.synchronize()
        log("[bench] calibration: 3 x (1 GiB read + 1 GiB written)")
        return
    if args.only_serving:
        print(json.dumps({"serving": serving_benchmark(args, device, log)}), flush=True)
        return
    log("[bench] building scene")
    t0 = time.time()
    sc = build_scene(args, device)
    torch.cuda.synchronize()
    log(f"[bench] scene ready in {time.time() - t0:.1f} s: {sc.voxels.num_voxels} voxels, "
        f"{sc.voxels.lower_corners.shape[1]} columns")
    inputs = (sc.world, sc.sdf.detach(), sc.sh.detach(), sc.target)
    b = sc.cams.extrinsics.shape[0]
    group = args.camera_group or None
    if args.probe:
        from meshtryoshka_amd import batched

        for ts in ((16, 16), (8, 8)):
            cap = batched.measure_capacities(sc.voxels, sc.level_sets, sc.cams.extrinsics,
                                             sc.cams.intrinsics, sc.shapes, sc.world, sc.sdf.detach(),
                                             sc.sh.detach(), tile_shape=ts, camera_group=group)
            log(f"[probe] tiles {ts}: {cap}")
        log(f"[probe] voxels {sc.voxels.num_voxels} columns {sc.voxels.lower_corners.shape[1]} "
            f"peak memory {torch.cuda.max_memory_allocated() / 2**30:.1f} GiB")
        return

    # 1. Plain per-camera API path.
    plain_s = None
    if not args.only_batched:
        for _ in range(max(1, args.warmup // 2)):
            plain_per_camera_step(sc)
        torch.cuda.synchronize()
        plain_steps = max(2, args.steps // 3)
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            plain_loss = plain_per_camera_step(sc).detach()
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t0) / plain_steps
        log(f"[bench] plain per-camera path: {1000 * plain_s:.2f} ms/step (loss {float(plain_loss):.6f})")
    modes = None
    if not args.only_batched and sc.uniform and not args.no_mode_compare:
        modes = compare_render_modes(sc, log)

    # 2. Autotune the batched step (tile shape x chunk).
    if args.no_autotune or args.only_batched:
        best = autotune.Choice(tuple(args.tile), args.chunk, float("nan"))
        tuned = []
    else:
        best, tuned = autotune.autotune(sc.voxels, sc.level_sets, sc.cams.extrinsics, sc.cams.intrinsics,
                                        sc.shapes, sc.sh.shape[0], sc.spec.sh_degree, sc.spec.sharpness,
                                        sc.spec.background, inputs, near_plane=sc.spec.near_plane, log=log,
                                        camera_group=group)
    log(f"[bench] autotune choice: tile {best.tile_shape} chunk {best.chunk}")
    step = make_batched(sc, best.tile_shape, best.chunk, camera_group=group)

    # 3. Graph-captured batched step: the timed region.
    step.capture(*inputs)
    for _ in range(args.warmup):
        step.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if step.overflowed():
        raise RuntimeError("batched step overflowed its capacities")
    ms_per_step = 1000 * elapsed / args.steps
    live = step.live_sizes()
    log(f"[bench] batched graph step: {ms_per_step:.2f} ms/step, loss {float(step.loss):.6f}, {live}")

    if args.only_batched:
        # Reference transfer for HBM counters (FETCH_SIZE / WRITE_SIZE): one elementwise
        # kernel that reads exactly 1 GiB and writes exactly 1 GiB (torch's MulFunctor
        # kernel, the last one of the run), so counter readings can be scaled to bytes.
        src = torch.ones(1 << 28, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy: 1 GiB read + 1 GiB written")

    # 4. Per-kernel durations: the same kernels, eager, bracketed by HIP events.
    timer = LaunchTimer()
    stats = torch.zeros(2, dtype=torch.int64, device=device)
    inst_steps = 2
    for _ in range(inst_steps):
        step_args = step._composite_args
        def with_stats(g=None, _orig=step_args):
            a = _orig(g)
            a.stats = stats.data_ptr()
            return a
        step._composite_args = with_stats
        timer.start()
        step.run(*inputs, hook=timer.hook)
        step._composite_args = step_args
    torch.cuda.synchronize()
    timer.collect()
    sizes = {
        "N_lc": int(sc.voxels.lower_corners.shape[1]), "N_vox": int(sc.voxels.upper_corners.shape[1]),
        "V_t": live["V_t"], "F": live["F"], "K": live["K"], "B": b, "S": int(sc.level_sets.shape[0]),
        "G": len(step.groups), "P": sc.num_pixels, "TT": step.total_tiles, "NSH": int(sc.sh.shape[0]),
        "tile_shape": list(best.tile_shape), "chunk": best.chunk,
    }
    if sc.uniform:
        sizes["H"], sizes["W"] = sc.shapes[0]
        sizes["T"] = step.T
    sizes["sort_passes"] = sort_passes(step.sorted_keys[: live["K_last_group"]])
    # Vertices that received a gradient (last group; the colour adjoint's work unit).
    sizes["V_touched"] = int(step.touched[: live["V_t"]].to(torch.int64).sum())
    alg = algorithmic_bytes(sizes)
    table = {}
    for name in timer.totals:
        ms = timer.step_ms(name, inst_steps)
        entry = {"ms": round(ms, 4), "launches": timer.counts[name] // inst_steps}
        if alg.get(name):
            entry["GB/s"] = round(alg[name] / (ms * 1e-3) / 1e9, 1)
            entry["hbm_frac"] = round(entry["GB/s"] / HBM_PEAK_GBS, 4)
        table[name] = entry
    s = stats.tolist()
    tests = s[0] / (2 * inst_steps)
    hits = s[1] / (2 * inst_steps)
    sizes["composite_tests_per_step"] = tests
    sizes["composite_hits_per_step"] = hits
    if args.breakdown:
        for name, e in sorted(table.items(), key=lambda kv: -kv[1]["ms"]):
            log(f"  {name:40s} {e}")
    # Roofline of the dominant stage: algorithmic bytes of one launch over the average
    # launch duration (HIP events on the launch stream).
    dom = max((k for k in table if k.startswith("mt_")), key=lambda k: table[k]["ms"])
    launches = table[dom]["launches"]
    ms_launch = timer.mean_ms(dom)
    bytes_launch = alg.get(dom, 0) / max(launches, 1)
    achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profile_traffic(dom, workload_key(args, best)),
            "ms_per_launch": round(ms_launch, 4), "launches_per_step": launches,
            "algorithmic_bytes_per_launch": int(bytes_launch)}
    instr = profile_counter(DOMINANT_KERNELS.get(dom, ""), "SQ_INSTS_VALU", workload_key(args, best))
    if instr:
        # VALU instruction issue: a wave64 VALU instruction takes 4 cycles of a 16-lane SIMD.
        issue = instr / (ms_launch * 1e-3) / 1e9
        roof["valu"] = {"achieved": round(issue, 1), "peak": VALU_ISSUE_PEAK_G, "unit": "G wave-instr/s",
                        "frac": round(issue / VALU_ISSUE_PEAK_G, 4), "instructions_per_launch": int(instr),
                        "source": "SQ_INSTS_VALU of the committed counter profile of this workload"}
        roof["limiter"] = "VALU instruction issue and per-block latency (HBM fraction shown for the contract)"
    if dom.startswith("mt_intersect_and_composite"):
        fl = (58 * tests + (25 if dom.endswith("forward") else 60) * hits) / max(launches, 1)
        roof["valu"] = {"achieved": round(fl / (ms_launch * 1e-3) / 1e12, 3), "peak": FP32_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "note": "FP32 vector ops: 58 per (pixel, triangle) "
                        "test + 25 (fwd) / 60 (bwd) per hit"}

    log("[bench] SDF regularizers")
    reg = regularizer_benchmark(sc, 0 if (args.only_batched or args.no_cpu_baseline) else 2_000_000)
    log(f"[bench] regularizers: {reg}")

    cpu = None
    if not (args.no_cpu_baseline or args.only_batched or not sc.uniform):
        log(f"[bench] CPU baseline (PyTorch-CPU transcription, {cpu_threads()} threads)")
        cpu = cpu_baseline(sc, args, live, log)
        log(f"[bench] CPU baseline: {cpu['seconds_per_step']} s/step")

    workload = workload_description(args, sc)
    sub = None
    if not args.only_batched:
        log("[bench] subdivision")
        sub = subdivision_benchmark(device)
        log(f"[bench] subdivision: {sub}")
    serve = None
    c3 = None
    run_c3 = not (args.only_batched or args.no_c3) and args.workload == "sphere-field-4cam"
    run_serve = not (args.no_serving or args.only_batched)
    if run_c3 or run_serve:
        # Free the training step's arenas and the scene before the next legs.
        import gc

        del step, step_args, with_stats, inputs
        sc = None
        timer = None
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        log(f"[bench] device memory after freeing the training step: "
            f"{torch.cuda.memory_allocated(device) / 2**30:.1f} GiB")
    if run_c3:
        c3 = c3_leg(args, device, log, best.tile_shape, best.chunk)
    if run_serve:
        log("[bench] serving sweep")
        serve = serving_benchmark(args, device, log)

    out = {
        "metric": "train steps/s fwd+bwd (batched graph-captured step; plain per-camera path beside it)",
        "value": round(args.steps / elapsed, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "name": args.workload,
            "camera_group": args.camera_group or args.cameras,
            "sizes": sizes,
        },
        "plain_per_camera": None if plain_s is None else {
            "steps_per_s": round(1.0 / plain_s, 4), "ms_per_step": round(1000 * plain_s, 3)},
        "speedup_vs_plain": None if plain_s is None else round(plain_s / (elapsed / args.steps), 3),
        "autotune": {"tile_shape": list(best.tile_shape), "chunk": best.chunk,
                     "candidates": [[list(c.tile_shape), c.chunk, round(c.ms, 3)] for c in tuned]},
        "roofline": roof,
        "kernels": table,
        "regularizers": reg,
        "subdivision": sub,
        "render_modes": modes,
        "c3": c3,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
