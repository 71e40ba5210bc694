#!/usr/bin/env python
"""Training-step benchmark of the extract -> rasterize -> composite path on one MI355X.

Workload (sphere-field training step): a sparse 1024^3 virtual grid whose occupied cells
lie within 3 cells of the zero set of min(64 seeded spheres, ground plane y = -0.6);
corner SDF = clean SDF + 0.01 N(0, 1); SH degree 3; 5 shells spaced 1/512 apart;
4 cameras at 1237x822 on the radius-2 upper hemisphere; 16x16 tiles; sharpness 50.
One step = tessellate_voxels (5 level sets) + fused render of the 4 cameras + mean
squared error against a seeded U(0, 1) target + backward to the grid SDF and SH.
Inputs are resident in HBM before timing starts.

Prints one JSON line (metric, value = steps/s, roofline of the dominant kernel measured
with HIP events over the timed steps, CPU baseline of the oracle on a bounded sample).
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


def parse_args():
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
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-tiles", type=int, default=24)
    p.add_argument("--breakdown", action="store_true", help="print the per-kernel table to stderr")
    return p.parse_args()


# ------------------------------------------------------------------------------------
# Scene
# ------------------------------------------------------------------------------------


class Scene:
    pass


def build_scene(args, device):
    import torch

    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import workloads

    spec = workloads.sphere_field(seed=args.seed, num_spheres=args.spheres,
                                  resolution=args.resolution, num_cameras=args.cameras,
                                  image_shape=(args.height, args.width), device=device)
    sc = Scene()
    sc.spec = spec
    sc.voxels = te.extract_voxels(spec.occupancy.to(device))
    world, sdf, sh = workloads.finish_fields(spec, sc.voxels.vertices)
    sc.world = world.to(device)
    sc.sdf = sdf.to(device).requires_grad_(True)
    sc.sh = sh.to(device).requires_grad_(True)
    sc.cams = spec.cameras.to(device)
    b = sc.cams.extrinsics.shape[0]
    h, w = sc.cams.image_shape
    sc.target = workloads.target_image(spec, (b, 3, h, w)).to(device)
    sc.level_sets = spec.level_sets.to(device)
    sc.sharpness = torch.tensor(spec.sharpness, device=device)
    sc.background = torch.tensor(spec.background, device=device)
    return sc


# ------------------------------------------------------------------------------------
# Plain step (the reference's calling pattern, per-op calls with host size syncs),
# instrumented with HIP events around each kernel stage.
# ------------------------------------------------------------------------------------


class StageTimer:
    """Pairs of torch.cuda events on the current stream (the stream every op launches on)."""

    def __init__(self):
        import torch

        self.torch = torch
        self.pending = []  # (name, start, end)
        self.totals = {}
        self.counts = {}

    def event(self):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def add(self, name, start, end):
        self.pending.append((name, start, end))

    def collect(self):
        for name, s, e in self.pending:
            self.totals[name] = self.totals.get(name, 0.0) + s.elapsed_time(e)
            self.counts[name] = self.counts.get(name, 0) + 1
        self.pending = []

    def mean_ms(self, name):
        return self.totals[name] / self.counts[name]


def plain_step(sc, timer, stats=None):
    import torch

    from meshtryoshka_amd import triangle_extraction as te
    from meshtryoshka_amd import triangle_rasterization as tr
    from meshtryoshka_amd.triangle_rasterization.common import TileGrid

    v = sc.voxels
    sc.sdf.grad = None
    sc.sh.grad = None
    ev = timer.event
    # --- tessellate_voxels (triangle_extraction/__init__.py:75-157)
    e0 = ev()
    with torch.no_grad():
        vertex_offsets = te.count_triangle_vertices(sc.sdf, v.lower_corners, sc.level_sets)
        e1 = ev()
        te.compute_exclusive_cumsum(vertex_offsets)
        e2 = ev()
    tv, td, tsh, ttype = te.create_triangle_vertices(sc.world, sc.sdf, sc.sh, v.lower_corners,
                                                     sc.level_sets, vertex_offsets)
    e3 = ev()
    with torch.no_grad():
        face_offsets, codes = te.count_triangle_faces(sc.sdf, v.lower_corners, v.upper_corners,
                                                      sc.level_sets)
        e4 = ev()
        te.compute_exclusive_cumsum(face_offsets)
        e5 = ev()
        faces, _ = te.create_triangle_faces(sc.sdf, v.neighbors, v.indices, vertex_offsets, ttype,
                                            codes, sc.level_sets, face_offsets)
        e6 = ev()
    s_count = sc.level_sets.shape[0]
    face_boundaries = face_offsets[1:].view(s_count, -1)[:, -1].clone() // 3
    timer.add("count_triangle_vertices", e0, e1)
    timer.add("exclusive_scan(vertices)", e1, e2)
    timer.add("create_triangle_vertices", e2, e3)
    timer.add("count_triangle_faces", e3, e4)
    timer.add("exclusive_scan(faces)", e4, e5)
    timer.add("create_triangle_faces", e5, e6)
    # --- render_fused (triangle_rasterization/__init__.py:167-287)
    b = sc.cams.extrinsics.shape[0]
    h, w = sc.cams.image_shape
    grid = TileGrid(sc.spec.tile_shape, (h, w), b, s_count)
    kd = tr.create_key_defines(("camera", 6), ("tile", 20), ("depth", 32), ("shell", 6))
    f0 = ev()
    xy, depth = tr.project_vertices(tv.detach(), sc.cams.extrinsics, sc.cams.intrinsics)
    f1 = ev()
    bounds = tr.bound_triangles(xy, depth, faces, grid, sc.spec.near_plane)
    f2 = ev()
    pk = tr.generate_keys(depth, faces, face_boundaries, bounds.tile_minima, bounds.tile_maxima,
                          bounds.num_tiles_overlapped, grid, kd)
    f3 = ev()
    spk = tr.sort_keys(pk, grid)
    f4 = ev()
    ranges = tr.delineate_tiles(spk, grid, kd, False)
    f5 = ev()
    colors = tr.evaluate_spherical_harmonics(tv.detach(), tsh, sc.cams.extrinsics, sc.spec.sh_degree)
    f6 = ev()
    t = grid.num_tiles_per_image
    img = _composite_with_stats(xy, colors, td, faces, sc.sharpness, spk, ranges.view(b, t, 2),
                                s_count, grid, kd, stats)
    f7 = ev()
    timer.add("project_vertices", f0, f1)
    timer.add("bound_triangles", f1, f2)
    timer.add("generate_keys(+scan)", f2, f3)
    timer.add("sort_keys", f3, f4)
    timer.add("delineate_tiles", f4, f5)
    timer.add("evaluate_sh_forward", f5, f6)
    timer.add("intersect_and_composite_forward", f6, f7)
    rgb = img[:, :3] + img[:, 3:4] * sc.background[:, None, None]
    rgb = rgb.clip(min=0, max=1)
    loss = ((rgb - sc.target) ** 2).mean()
    # Backward stage boundaries from gradient hooks (each fires on the current stream).
    marks = {}
    img.register_hook(lambda g: marks.__setitem__("c0", ev()))
    colors.register_hook(lambda g: marks.__setitem__("c1", ev()))
    tsh.register_hook(lambda g: marks.__setitem__("s1", ev()))
    sc.sdf.register_hook(lambda g: marks.__setitem__("v1", ev()))
    loss.backward()
    timer.add("intersect_and_composite_backward", marks["c0"], marks["c1"])
    timer.add("evaluate_sh_backward", marks["c1"], marks["s1"])
    if "v1" in marks:
        timer.add("create_triangle_vertices_backward", marks["s1"], marks["v1"])
    sizes = {
        "N_lc": int(v.lower_corners.shape[1]), "N_vox": int(v.upper_corners.shape[1]),
        "V_t": int(tv.shape[0]), "F": int(faces.shape[0]), "K": int(spk.keys.shape[0]),
        "B": b, "S": s_count, "H": h, "W": w, "T": t, "NSH": int(sc.sh.shape[0]),
    }
    sc.last_sorted_keys = spk.keys
    return loss, sizes


def _composite_with_stats(xy, colors, sdf, faces, sharpness, spk, ranges, s_count, grid, kd, stats):
    """intersect_and_composite through the public autograd Function; when `stats` is given,
    the kernels also count their (pixel, triangle) tests and hits into it."""
    from meshtryoshka_amd.triangle_rasterization import intersect_and_composite as iac

    if stats is None:
        return iac.intersect_and_composite(xy, colors, sdf, faces, sharpness, spk.keys,
                                           spk.triangle_indices, ranges, s_count, grid, kd)
    orig = iac._args

    def with_stats(*a, **k):
        args = orig(*a, **k)
        args.stats = stats.data_ptr()
        return args

    iac._args = with_stats
    try:
        return iac.intersect_and_composite(xy, colors, sdf, faces, sharpness, spk.keys,
                                           spk.triangle_indices, ranges, s_count, grid, kd)
    finally:
        iac._args = orig


# ------------------------------------------------------------------------------------
# Roofline accounting (algorithmic bytes / flops per launch)
# ------------------------------------------------------------------------------------


def algorithmic_bytes(z):
    """Unique bytes each kernel stage must move at minimum (assumes perfect on-chip reuse)."""
    nb = z["NSH"] * 12
    n_lc, n_vox, vt, f, k, b, s = z["N_lc"], z["N_vox"], z["V_t"], z["F"], z["K"], z["B"], z["S"]
    hw = z["H"] * z["W"]
    return {
        "count_triangle_vertices": n_lc * 16 + n_lc * 4 + 4 * s * n_lc,
        "exclusive_scan(vertices)": 8 * (s * n_lc + 1),
        "create_triangle_vertices": n_lc * 16 + n_lc * (12 + 4 + nb) + 4 * s * n_lc + vt * (12 + 4 + nb + 1),
        "count_triangle_faces": n_vox * 32 + n_lc * 4 + s * n_vox * (4 + 1),
        "exclusive_scan(faces)": 8 * (s * n_vox + 1),
        "create_triangle_faces": n_vox * (28 + 4) + s * n_vox * (1 + 8) + 4 * s * n_lc + vt + f * 16,
        "project_vertices": vt * 12 + b * vt * 12,
        "bound_triangles": f * 12 + b * vt * 12 + b * f * 20,
        "generate_keys(+scan)": b * f * (4 + 8 + 8) + f * 12 + b * vt * 4 + k * 12,
        "sort_keys": None,  # filled from the number of active passes
        "delineate_tiles": 8 * k + 16 * b * z["T"],
        "evaluate_sh_forward": vt * (12 + nb) + b * vt * 12,
        "intersect_and_composite_forward": 96 * k + 16 * b * hw + 16 * b * z["T"],
        "intersect_and_composite_backward": 96 * k + 32 * b * hw + 16 * b * z["T"] + 2 * (12 * b * vt + 4 * vt),
        "evaluate_sh_backward": vt * (12 + nb) + b * vt * 12 + vt * nb,
        "create_triangle_vertices_backward": n_lc * 16 + vt * (12 + 4 + nb) + 2 * n_lc * (4 + nb),
    }


def sort_passes(keys):
    """Digit passes the radix sort executes: 8-bit digits that are not constant."""
    passes = 0
    for p in range(8):
        d = (keys >> (8 * p)) & 0xFF
        if keys.numel() and int(d.min()) != int(d.max()):
            passes += 1
    return passes


# ------------------------------------------------------------------------------------
# CPU baseline: the oracle on a bounded sample of the same step
# ------------------------------------------------------------------------------------


def cpu_baseline(sc, args):
    import numpy as np

    from oracle import extraction as ox
    from oracle import rasterization as orr

    v = sc.voxels
    t0 = time.perf_counter()
    sdf = sc.sdf.detach().cpu().numpy()
    world = sc.world.cpu().numpy()
    sh = sc.sh.detach().cpu().numpy()
    lower = v.lower_corners.cpu().numpy()
    upper = v.upper_corners.cpu().numpy()
    nbr = v.neighbors.cpu().numpy()
    idx = v.indices.cpu().numpy()
    ls = sc.level_sets.cpu().numpy()
    t_copy = time.perf_counter() - t0
    t0 = time.perf_counter()
    tri = ox.tessellate_voxels(sdf, world, sh, nbr, lower, upper, idx, ls)
    t_tess = time.perf_counter() - t0
    E = sc.cams.extrinsics.cpu().numpy()
    K = sc.cams.intrinsics.cpu().numpy()
    h, w = sc.cams.image_shape
    b = E.shape[0]
    th, tw = sc.spec.tile_shape
    rows, cols = -(-h // th), -(-w // tw)
    t0 = time.perf_counter()
    xy, depth = orr.project_vertices(tri.vertices, E, K)
    tmin, tmax, cnt = orr.bound_triangles(xy, depth, tri.faces, (th, tw), (rows, cols), sc.spec.near_plane)
    lay = orr.fused_layout()
    keys, vals, _ = orr.generate_keys(depth, tri.faces, tri.face_boundaries, tmin, tmax, cnt, cols, lay)
    sk, sv = orr.sort_keys(keys, vals)
    ranges = orr.delineate_tiles(sk, lay, b, ls.shape[0], rows * cols, False)
    col, _ = orr.evaluate_spherical_harmonics(tri.vertices, tri.spherical_harmonics, E, sc.spec.sh_degree)
    col = col.astype(np.float32)
    t_front = time.perf_counter() - t0
    # Composite forward + backward on a seeded sample of non-empty tiles, scaled by keys.
    rng = np.random.default_rng(0)
    nonempty = np.argwhere(ranges[:, 0, :, 1] > ranges[:, 0, :, 0])
    pick = nonempty[rng.choice(len(nonempty), size=min(args.cpu_sample_tiles, len(nonempty)), replace=False)]
    sub = np.zeros_like(ranges[:, 0])
    sub[pick[:, 0], pick[:, 1]] = ranges[pick[:, 0], 0, pick[:, 1]]
    keys_sampled = int((sub[..., 1] - sub[..., 0]).sum())
    keys_total = int((ranges[:, 0, :, 1] - ranges[:, 0, :, 0]).sum())
    t0 = time.perf_counter()
    hits = orr.fused_hits(xy, tri.faces, sk, sv, sub, lay, ls.shape[0], (th, tw), (h, w))
    img = orr.fused_forward_image(hits, col, tri.signed_distances, tri.faces, sc.spec.sharpness, b, (h, w))
    g = np.ones_like(img)
    orr.fused_backward(hits, col.astype(np.float64), tri.signed_distances.astype(np.float64), tri.faces,
                       sc.spec.sharpness, img, g)
    t_comp_sample = time.perf_counter() - t0
    t_comp = t_comp_sample * keys_total / max(keys_sampled, 1)
    step_s = t_tess + t_front + t_comp
    return {
        "value": 1.0 / step_s,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"NumPy oracle, FP32: full tessellation ({t_tess:.1f} s) and front end "
                   f"(projection..SH, {t_front:.1f} s) of the step, plus fused composite fwd+bwd "
                   f"on {len(pick)} random non-empty tiles ({keys_sampled} of {keys_total} keys, "
                   f"{t_comp_sample:.1f} s) scaled by key count to {t_comp:.0f} s; "
                   f"excludes {t_copy:.1f} s of device-to-host copies"),
    }


# ------------------------------------------------------------------------------------


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by camera, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    print("[bench] building scene", file=sys.stderr, flush=True)
    t0 = time.time()
    sc = build_scene(args, device)
    torch.cuda.synchronize()
    print(f"[bench] scene ready in {time.time() - t0:.1f} s: {sc.voxels.num_voxels} voxels, "
          f"{sc.voxels.lower_corners.shape[1]} columns", file=sys.stderr, flush=True)

    timer = StageTimer()
    for i in range(args.warmup):
        loss, sizes = plain_step(sc, timer)
        torch.cuda.synchronize()
        print(f"[bench] warmup {i} loss {loss.item():.6f}", file=sys.stderr, flush=True)
    timer = StageTimer()
    stats = torch.zeros(2, dtype=torch.int64, device=device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, sizes = plain_step(sc, timer, stats)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timer.collect()
    ms_per_step = 1000 * elapsed / args.steps
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s ({ms_per_step:.2f} ms/step); sizes {sizes}",
          file=sys.stderr, flush=True)

    # Per-kernel table.
    alg = algorithmic_bytes(sizes)
    passes = sort_passes(sc.last_sorted_keys)
    sizes["sort_passes"] = passes
    table = {}
    for name in timer.totals:
        ms = timer.mean_ms(name)
        entry = {"ms": round(ms, 4)}
        b_alg = alg.get(name)
        if name == "sort_keys":
            # Histogram read + (read + write) of 12-byte pairs per active pass.
            b_alg = 8 * sizes["K"] + 24 * sizes["K"] * passes
        if b_alg:
            entry["GB/s"] = round(b_alg / (ms * 1e-3) / 1e9, 1)
            entry["hbm_frac"] = round(entry["GB/s"] / HBM_PEAK_GBS, 4)
        table[name] = entry
    s = stats.tolist()
    n_steps = args.steps
    tests_per_launch = s[0] / (2 * n_steps)  # forward and backward both count
    hits_per_launch = s[1] / (2 * n_steps)
    dom = max(table, key=lambda k: table[k]["ms"])
    flops_fwd = 58 * tests_per_launch + 25 * hits_per_launch
    if args.breakdown:
        for name, e in sorted(table.items(), key=lambda kv: -kv[1]["ms"]):
            print(f"  {name:36s} {e}", file=sys.stderr)
        print(f"  composite tests/launch {tests_per_launch:.3e} hits/launch {hits_per_launch:.3e}",
              file=sys.stderr)
    if dom.startswith("intersect_and_composite"):
        ms = timer.mean_ms(dom)
        fl = flops_fwd if dom.endswith("forward") else 58 * tests_per_launch + 60 * hits_per_launch
        ach = fl / (ms * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": dom, "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "ms": round(ms, 4), "flops_per_launch": fl,
                "note": "FP32 vector work: 58 flops per (pixel, triangle) test + 25 (fwd) / 60 (bwd) per hit"}
    else:
        e = table[dom]
        roof = {"bound": "hbm", "kernel": dom, "achieved": e.get("GB/s"), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": e.get("hbm_frac"), "traffic": None, "ms": e["ms"]}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle on a bounded sample)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sc, args)

    out = {
        "metric": "train steps/s fwd+bwd (sphere-field step, plain per-op path)",
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
            "workload": (f"sphere-field training step: {args.resolution}^3 virtual grid, "
                         f"{args.spheres} spheres + ground plane, band 3 cells, SH deg 3, 5 shells, "
                         f"{args.cameras} cameras {args.width}x{args.height}, tiles 16x16"),
            "path": "plain",
            "sizes": sizes,
        },
        "roofline": roof,
        "kernels": table,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
