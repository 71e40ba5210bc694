"""Seeded synthetic scenes for the extract -> rasterize -> composite path.

Every scene draws from one ``torch.Generator("cpu")`` seeded with ``seed`` in a fixed
order (scene primitives, then corner SDF noise, then SH coefficients, then cameras,
then the target image), so the same seed gives the same tensors on any device.

* ``dense_sphere``: a dense 64^3 cell grid on [-1, 1]^3 (every cell occupied; other
  resolutions are padded along k to whole 32-cell words), corner
  SDF ``|p| - 0.5 + 0.02 N(0, 1)``, SH degree 0 colour ``U(0, 1)``, three level sets
  {+0.02, 0, -0.02}, one 128x128 pinhole camera (60 deg FOV) at distance 2.5 on +z.
* ``sphere_field``: a sparse grid at ``resolution^3`` virtual cells; cells within
  ``band`` cells of the zero set of ``min(spheres, ground plane y = -0.6)`` are occupied;
  corner SDF = clean SDF + ``noise`` N(0, 1); SH degree 3 with DC ``U(0, 1)`` and higher
  bands ``N(0, 0.05)``; shells spaced ``1/512`` apart around 0; cameras on the upper
  hemisphere of radius 2 looking at the origin (+y up, 60 deg vertical FOV). With
  ``resolutions``, each camera's (H, W) is drawn uniformly from that list (after the
  SH coefficients, before the camera poses), and ``Cameras.image_shape`` is the per-camera
  list.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

from .triangle_extraction.common import pack_occupancy


@dataclass
class Cameras:
    extrinsics: torch.Tensor  # f32[B, 4, 4] OpenCV world-to-camera
    intrinsics: torch.Tensor  # f32[B, 3, 3] pixels
    image_shape: tuple  # (H, W) shared by every camera, or a list of per-camera (H, W)

    def to(self, device):
        return Cameras(self.extrinsics.to(device), self.intrinsics.to(device), self.image_shape)


@dataclass
class SceneSpec:
    occupancy: torch.Tensor  # packed int32[I, J, K/32]
    level_sets: torch.Tensor  # f32[S] descending
    sh_degree: int
    sharpness: float
    background: tuple[float, float, float]
    tile_shape: tuple[int, int]
    near_plane: float
    seed: int
    field_fn: object = None  # (unit_vertices[V, 3], generator) -> (world vertices, sdf, sh)
    cameras: Cameras | None = None
    extra: dict = field(default_factory=dict)


def look_at_w2c(position: torch.Tensor, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> torch.Tensor:
    """OpenCV world-to-camera (+x right, +y down, +z forward) looking at ``target``."""
    pos = position.double()
    fwd = torch.tensor(target, dtype=torch.float64) - pos
    fwd = fwd / fwd.norm()
    upv = torch.tensor(up, dtype=torch.float64)
    right = torch.linalg.cross(fwd, upv)
    if right.norm() < 1e-6:
        right = torch.linalg.cross(fwd, torch.tensor((0.0, 0.0, 1.0), dtype=torch.float64))
    right = right / right.norm()
    down = torch.linalg.cross(fwd, right)
    rot = torch.stack([right, down, fwd])
    w2c = torch.eye(4, dtype=torch.float64)
    w2c[:3, :3] = rot
    w2c[:3, 3] = -rot @ pos
    return w2c.float()


def pinhole(image_shape, fov_y_deg=60.0) -> torch.Tensor:
    h, w = image_shape
    f = (h / 2) / math.tan(math.radians(fov_y_deg) / 2)
    return torch.tensor([[f, 0.0, w / 2], [0.0, f, h / 2], [0.0, 0.0, 1.0]], dtype=torch.float32)


# ----------------------------------------------------------------------------------------
# Dense sphere (every cell occupied)
# ----------------------------------------------------------------------------------------


def dense_sphere(seed: int = 0, resolution: int = 64, image_size: int = 128) -> SceneSpec:
    # Packed rows hold whole 32-cell words: a resolution that is not a multiple of 32 is
    # padded with unoccupied cells along k (x), and x is rescaled so that the occupied
    # cells still span [-1, 1] (the scale is exactly 1 without padding).
    kpad = -(-resolution // 32) * 32
    occ = torch.zeros((resolution, resolution, kpad), dtype=torch.bool)
    occ[:, :, :resolution] = True
    scale = torch.tensor([kpad / resolution, 1.0, 1.0], dtype=torch.float32)

    def field_fn(unit_vertices: torch.Tensor, gen: torch.Generator):
        p = 2 * (unit_vertices * scale.to(unit_vertices.device)) - 1
        noise = torch.randn(p.shape[0], generator=gen, dtype=torch.float32).to(p.device)
        sdf = p.norm(dim=-1) - 0.5 + 0.02 * noise
        sh = torch.rand((1, p.shape[0], 3), generator=gen, dtype=torch.float32).to(p.device)
        return p.contiguous(), sdf.float().contiguous(), sh.contiguous()

    e = torch.tensor(
        [[1.0, 0.0, 0.0, 0.0], [0.0, -1.0, 0.0, 0.0], [0.0, 0.0, -1.0, 2.5], [0.0, 0.0, 0.0, 1.0]]
    )
    k = pinhole((image_size, image_size), 60.0)
    cams = Cameras(e[None].clone(), k[None].clone(), (image_size, image_size))
    return SceneSpec(
        occupancy=pack_occupancy(occ),
        level_sets=torch.tensor([0.02, 0.0, -0.02], dtype=torch.float32),
        sh_degree=0,
        sharpness=50.0,
        background=(1.0, 1.0, 1.0),
        tile_shape=(16, 16),
        near_plane=0.2,
        seed=seed,
        field_fn=field_fn,
        cameras=cams,
    )


# ----------------------------------------------------------------------------------------
# Sparse sphere field
# ----------------------------------------------------------------------------------------


def _clean_sdf(p: torch.Tensor, centres: torch.Tensor, radii: torch.Tensor, chunk: int = 1 << 20):
    out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
    c = centres.to(p.device)
    r = radii.to(p.device)
    for s in range(0, p.shape[0], chunk):
        q = p[s : s + chunk]
        d = (q[:, None, :] - c[None]).norm(dim=-1) - r[None]
        out[s : s + chunk] = torch.minimum(d.min(dim=1).values, q[:, 1] + 0.6)
    return out


def sphere_occupancy(resolution: int, centres, radii, band_cells: float, device="cpu",
                     packed: bool = True):
    """Occupancy of cells whose centre lies within band_cells cells of the zero set.

    Evaluated coarse-to-fine: the SDF is 1-Lipschitz, so a coarse block whose centre
    value exceeds the band plus the block's half-diagonal holds no occupied cell.
    Returns the packed int32[n, n, n / 32] grid (or bool[n, n, n] if packed=False)."""
    n = resolution
    h = 2.0 / n
    band = band_cells * h
    block = 8
    nb = n // block
    hb = 2.0 / nb
    idx = torch.arange(nb, device=device, dtype=torch.float32)
    ctr = -1 + (idx + 0.5) * hb
    gz, gy, gx = torch.meshgrid(ctr, ctr, ctr, indexing="ij")  # (i -> z, j -> y, k -> x)
    pc = torch.stack([gx, gy, gz], dim=-1).reshape(-1, 3)
    sc = _clean_sdf(pc, centres, radii)
    keep = (sc.abs() <= band + math.sqrt(3) * hb / 2 + 1e-6).nonzero().squeeze(1)
    words = torch.zeros((n, n, n // 32), dtype=torch.int64, device=device)
    occ = None if packed else torch.zeros((n, n, n), dtype=torch.bool, device=device)
    if keep.numel() > 0:
        bi = keep // (nb * nb)
        bj = (keep // nb) % nb
        bk = keep % nb
        local = torch.arange(block, device=device)
        li, lj, lk = torch.meshgrid(local, local, local, indexing="ij")
        li, lj, lk = li.reshape(-1), lj.reshape(-1), lk.reshape(-1)
        step = 4096
        for s in range(0, keep.numel(), step):
            ci = (bi[s : s + step, None] * block + li[None]).reshape(-1)
            cj = (bj[s : s + step, None] * block + lj[None]).reshape(-1)
            ck = (bk[s : s + step, None] * block + lk[None]).reshape(-1)
            p = torch.stack([-1 + (ck + 0.5) * h, -1 + (cj + 0.5) * h, -1 + (ci + 0.5) * h], dim=-1)
            d = _clean_sdf(p.float(), centres, radii)
            m = d.abs() <= band
            ci, cj, ck = ci[m], cj[m], ck[m]
            if occ is not None:
                occ[ci, cj, ck] = True
            # Each cell is visited once, so summing distinct bits is a bitwise or.
            words.index_put_((ci, cj, ck // 32), torch.ones_like(ck) << (ck % 32), accumulate=True)
    if occ is not None:
        return occ
    # Reinterpret the low 32 bits as int32 (two's complement) without a uint32 dtype.
    return torch.where(words >= 2**31, words - 2**32, words).to(torch.int32)


def hemisphere_cameras(gen: torch.Generator, num_cameras: int, image_shape, radius=2.0, fov_y=60.0,
                       shapes=None) -> list:
    out = []
    for c in range(num_cameras):
        d = torch.randn(3, generator=gen, dtype=torch.float64)
        d = d / d.norm()
        d[1] = d[1].abs()
        pos = radius * d
        shape = image_shape if shapes is None else shapes[c]
        out.append((look_at_w2c(pos), pinhole(shape, fov_y), shape))
    return out


def sphere_field(
    seed: int = 0,
    num_spheres: int = 64,
    resolution: int = 1024,
    band_cells: float = 3.0,
    noise: float = 0.01,
    sh_degree: int = 3,
    num_shells: int = 5,
    num_cameras: int = 4,
    image_shape=(822, 1237),
    sharpness: float = 50.0,
    device="cpu",
    resolutions=None,
) -> SceneSpec:
    gen = torch.Generator("cpu").manual_seed(seed)
    centres = (torch.rand((num_spheres, 3), generator=gen) * 1.6 - 0.8).float()
    radii = (torch.rand((num_spheres,), generator=gen) * 0.25 + 0.05).float()
    packed = sphere_occupancy(resolution, centres, radii, band_cells, device=device)
    nsh = (sh_degree + 1) ** 2

    def field_fn(unit_vertices: torch.Tensor, gen2: torch.Generator):
        p = (2 * unit_vertices - 1).float()
        n = torch.randn(p.shape[0], generator=gen2, dtype=torch.float32).to(p.device)
        sdf = _clean_sdf(p, centres, radii) + noise * n
        sh = torch.empty((nsh, p.shape[0], 3), dtype=torch.float32)
        sh[0] = torch.rand((p.shape[0], 3), generator=gen2)
        if nsh > 1:
            sh[1:] = torch.randn((nsh - 1, p.shape[0], 3), generator=gen2) * 0.05
        return p.contiguous(), sdf.contiguous(), sh.to(p.device).contiguous()

    half = (num_shells - 1) / 2
    level_sets = torch.tensor([(half - s) / 512.0 for s in range(num_shells)], dtype=torch.float32)
    spec = SceneSpec(
        occupancy=packed,
        level_sets=level_sets,
        sh_degree=sh_degree,
        sharpness=sharpness,
        background=(1.0, 1.0, 1.0),
        tile_shape=(16, 16),
        near_plane=0.2,
        seed=seed,
        field_fn=field_fn,
        extra={"generator": gen, "num_cameras": num_cameras, "image_shape": image_shape,
               "centres": centres, "radii": radii,
               "resolutions": None if resolutions is None else [tuple(r) for r in resolutions]},
    )
    return spec


def finish_fields(spec: SceneSpec, unit_vertices: torch.Tensor):
    """Draw the per-corner fields in the scene's seeded order, after the occupancy.

    The clean SDF is evaluated on ``unit_vertices``' device; random draws always come from
    the scene's CPU generator, so results do not depend on the device."""
    gen = spec.extra.get("generator")
    if gen is None:
        gen = torch.Generator("cpu").manual_seed(spec.seed)
        spec.extra["generator"] = gen
    world, sdf, sh = spec.field_fn(unit_vertices, gen)
    if spec.cameras is None:
        nc = spec.extra["num_cameras"]
        choices = spec.extra.get("resolutions")
        shapes = None
        if choices:
            pick = torch.randint(len(choices), (nc,), generator=gen).tolist()
            shapes = [choices[i] for i in pick]
        cams = hemisphere_cameras(gen, nc, spec.extra["image_shape"], shapes=shapes)
        spec.cameras = Cameras(
            torch.stack([c[0] for c in cams]), torch.stack([c[1] for c in cams]),
            cams[0][2] if shapes is None else [c[2] for c in cams],
        )
    return world, sdf, sh


def target_image(spec: SceneSpec, shape) -> torch.Tensor:
    gen = spec.extra["generator"]
    return torch.rand(shape, generator=gen, dtype=torch.float32)
