"""Workload definitions: scenes, cameras and per-frame parameters.

``small_plain`` is the 64^3 / 256x256 workload the reference's own CPU path
can run; ``hd_di`` is the 1920x1080 ReSTIR DI + 3-bounce path workload on a
256^3 grid in 64 chunks.  ``reference_mode=True`` keeps the reference's constants
(defines.glsl:34-62, taa.glsl:13, and camera.h:12 where fov = 45 is passed
to glm::perspectiveRH_ZO as radians); the default uses the workload's own
values (60 degree FOV, 32 RIS candidates, 5 neighbours at 30 px).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Tuple

import numpy as np

__all__ = ["Workload", "small_plain", "hd_di", "palette_pt", "frame_params", "scene_desc", "batched_views", "vox_view"]


def _lib():
    from . import lib
    return lib()


@dataclass
class Workload:
    name: str
    grid: Tuple[int, int, int]
    chunk: Tuple[int, int, int]
    occupied: int
    emissive: int
    le: Tuple[float, float]
    seed: int
    width: int
    height: int
    frames: int
    fov_radians: float
    camera: Callable[[int], Tuple[tuple, tuple, tuple]]
    features: int
    taa: bool
    restir: object = None
    albedo: Tuple[float, float] = (0.2, 0.8)
    time: float = 1.0          # Status.time as main.cpp:1289 sets it
    afternoon: int = 1         # Status.is_afternoon, main.cpp:1290
    palette: Tuple[int, int] = (0, 0)  # (entries, emissive entries); (0, 0): one material per voxel
    extra: dict = field(default_factory=dict)


def scene_desc(w: Workload):
    from . import SceneDesc, CT_VOXEL_EXTENT
    d = SceneDesc()
    d.grid[:] = list(w.grid)
    d.chunk[:] = list(w.chunk)
    d.occupied_count = w.occupied
    d.emissive_count = w.emissive
    d.le_min, d.le_max = w.le
    d.albedo_min, d.albedo_max = w.albedo
    d.seed = w.seed
    d.palette_size, d.palette_emissive = w.palette
    d.voxel_extent = CT_VOXEL_EXTENT
    return d


def _static_camera(pos, fwd, up):
    return lambda f: (pos, fwd, up)


def _orbit_camera(radius, height, deg_per_frame, start_deg=0.0, center=(0.0, 0.0, 0.0)):
    def cam(f):
        a = math.radians(start_deg + deg_per_frame * f)
        rel = (radius * math.sin(a), height, radius * math.cos(a))
        n = math.sqrt(rel[0] ** 2 + rel[1] ** 2 + rel[2] ** 2)
        fwd = (-rel[0] / n, -rel[1] / n, -rel[2] / n)
        pos = tuple(float(c) + r for c, r in zip(center, rel))
        return pos, fwd, (0.0, 1.0, 0.0)
    return cam


def _indirect(rp, indirect, max_depth):
    rp.indirect = 1 if indirect else 0
    rp.max_depth = max_depth


def small_plain(reference_mode=False, width=256, height=256, frames=4, taa=False, indirect=None, max_depth=1):
    """64^3 voxels, 12 % fill (31,457 voxels), 8 emissive voxels of radiance 10.

    Camera at (0, 0, 2.5) looking down -z (camera.h:8-10), 1 spp, frames 0..3.
    The workload's own mode adds one diffuse bounce (ReSTIR PT with
    max_depth 1); reference mode keeps INDIRECT_ILLUMINATION_ON = 0.
    """
    from . import CT_RIS_CUBE_LIGHT_BIT, restir_params_reference
    rp = restir_params_reference()
    if not reference_mode:
        rp.min_ris_cube_samples = rp.max_ris_cube_samples = 32
        rp.min_neighbors = rp.max_neighbors = 5
        rp.min_neighbor_radius = rp.max_neighbor_radius = 30.0
    _indirect(rp, (not reference_mode) if indirect is None else indirect, max_depth)
    return Workload(
        name="small_plain" + ("_reference" if reference_mode else ""),
        grid=(64, 64, 64), chunk=(64, 64, 64), occupied=round(0.12 * 64 ** 3), emissive=8,
        le=(10.0, 10.0), seed=0xC0, width=width, height=height, frames=frames,
        fov_radians=45.0 if reference_mode else math.radians(60.0),
        camera=_static_camera((0.0, 0.0, 2.5), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),
        features=CT_RIS_CUBE_LIGHT_BIT, taa=taa, restir=rp)


def hd_di(frames=16, width=1920, height=1080, grid=256, chunk=64, indirect=True, max_depth=3):
    """256^3 voxels in 64 chunks of 64^3, 25 % fill (4,194,304 voxels), 2 %
    emissive (83,886 lights, radiance U[1, 20]); ReSTIR DI plus a 3-bounce
    diffuse path (ReSTIR PT, max_depth 3 = MAX_DEPTH) at 1920x1080 with a
    1 degree/frame camera orbit and TAA at 0.1 blend."""
    from . import CT_RIS_CUBE_LIGHT_BIT, restir_params_reference
    rp = restir_params_reference()
    rp.taa_alpha = 0.1
    _indirect(rp, indirect, max_depth)
    occ = round(0.25 * grid ** 3)
    extent = grid * 0.03125
    return Workload(
        name=f"hd_di_{grid}", grid=(grid, grid, grid), chunk=(chunk, chunk, chunk), occupied=occ,
        emissive=round(0.02 * occ), le=(1.0, 20.0), seed=0xC1, width=width, height=height, frames=frames,
        fov_radians=math.radians(60.0), camera=_orbit_camera(1.5 * extent, 0.375 * extent, 1.0),
        features=CT_RIS_CUBE_LIGHT_BIT, taa=True, restir=rp)


def palette_pt(frames=16, width=1920, height=1080, grid=512, chunk=64, max_depth=4):
    """A sparse 512^3 palette scene: 12 % of the cells occupied (16,106,127
    voxels, u32 voxel indices) in 64^3 chunks, one uint8 entry of a
    256-entry palette per voxel (albedo U[0.2, 0.8] per channel; 1 % of the
    entries, 3, emit radiance U[1, 20] per channel, so every voxel of them is
    a cube light); ReSTIR DI plus ReSTIR PT with the hybrid shift, 4
    bounces, path M-cap 20 (MAX_INFLUENCE_FROM_THE_PAST_THRESHOLD_PT) and 3
    spatial neighbours (NEIGHBOR_COUNT), at 1920x1080 with a 1 degree/frame
    camera orbit and TAA at 0.1 blend."""
    from . import CT_RIS_CUBE_LIGHT_BIT, CT_SHIFT_HYBRID, restir_params_reference
    rp = restir_params_reference()
    rp.taa_alpha = 0.1
    _indirect(rp, True, max_depth)
    rp.pt_shift_mapping = CT_SHIFT_HYBRID
    rp.pt_m_cap = 20.0
    rp.pt_neighbor_count = 3
    occ = round(0.12 * grid ** 3)
    extent = grid * 0.03125
    return Workload(
        name=f"palette_pt_{grid}", grid=(grid, grid, grid), chunk=(chunk, chunk, chunk), occupied=occ,
        emissive=0, le=(1.0, 20.0), seed=0xC2, width=width, height=height, frames=frames,
        fov_radians=math.radians(60.0), camera=_orbit_camera(1.5 * extent, 0.375 * extent, 1.0),
        features=CT_RIS_CUBE_LIGHT_BIT, taa=True, restir=rp, palette=(256, 3))


def vox_view(scene, width=1920, height=1080, frames=16, indirect=True, max_depth=3, name="vox"):
    """Frames of a loaded .vox scene: the camera orbits the scene's centre at
    1.5x its largest extent, 1 degree/frame, 60 degree field of view;
    ReSTIR DI over its cube lights plus a max_depth-bounce path, TAA at 0.1."""
    from . import CT_RIS_CUBE_LIGHT_BIT, restir_params_reference
    lo, hi = np.asarray(scene.grid_min, np.float64), np.asarray(scene.grid_max, np.float64)
    ext = float((hi - lo).max())
    rp = restir_params_reference()
    rp.taa_alpha = 0.1
    _indirect(rp, indirect, max_depth)
    return Workload(
        name=name, grid=(0, 0, 0), chunk=(0, 0, 0), occupied=scene.info.primitive_count,
        emissive=scene.info.cube_light_count, le=(0.0, 0.0), seed=0, width=width, height=height, frames=frames,
        fov_radians=math.radians(60.0), camera=_orbit_camera(1.5 * ext, 0.375 * ext, 1.0, center=(lo + hi) / 2),
        features=CT_RIS_CUBE_LIGHT_BIT, taa=True, restir=rp)


def frame_params(w: Workload, f: int):
    """FrameParams of frame f: camera of f, previous camera of f-1 (the
    reference copies the matrices into prev_* at the end of draw,
    main.cpp:1660-1666), TAA only once a previous frame exists
    (main.cpp:1483)."""
    from . import FrameParams, camera_matrices, prev_view_proj, CT_TAA_BIT
    aspect = w.width / w.height
    pos, fwd, up = w.camera(f)
    iv, ip = camera_matrices(pos, fwd, up, w.fov_radians, aspect)
    if f > 0:
        ppos, pfwd, pup = w.camera(f - 1)
        piv, pip = camera_matrices(ppos, pfwd, pup, w.fov_radians, aspect)
    else:
        piv, pip = iv, ip
    p = FrameParams()
    p.width, p.height, p.frame_number = w.width, w.height, f
    p.is_active = w.features | (CT_TAA_BIT if (w.taa and f > 0) else 0)
    p.time = w.time
    p.is_afternoon = w.afternoon
    p.has_previous = 1 if f > 0 else 0
    p.inv_view[:] = iv.tolist()
    p.inv_proj[:] = ip.tolist()
    p.prev_view_proj[:] = prev_view_proj(piv, pip).tolist()
    p.restir = w.restir
    return p


def batched_views(n_views=64, seed=0xC3, grids=(64, 128, 256, 384), max_depth=2, size_scale=1.0):
    """Four scenes (64^3 .. 384^3, noise fill U[10 %, 30 %], 2 % emissive)
    and n_views cameras on seeded orbits with resolutions U[128, 1024] x
    U[128, 768]; every view runs ReSTIR DI + a max_depth-bounce path, 1 spp.
    Returns (scene workloads, [(scene index, view workload)])."""
    from . import CT_RIS_CUBE_LIGHT_BIT, restir_params_reference
    rng = np.random.default_rng(seed)
    scenes = []
    for k, g in enumerate(grids):
        fill = float(rng.uniform(0.10, 0.30))
        occ = round(fill * g ** 3)
        rp = restir_params_reference()
        rp.taa_alpha = 0.1
        _indirect(rp, max_depth > 0, max(max_depth, 1))
        scenes.append(Workload(name=f"batch_scene_{g}", grid=(g, g, g), chunk=(min(g, 64),) * 3, occupied=occ,
                               emissive=round(0.02 * occ), le=(1.0, 20.0), seed=seed + k, width=0, height=0,
                               frames=1, fov_radians=math.radians(60.0), camera=None,
                               features=CT_RIS_CUBE_LIGHT_BIT, taa=True, restir=rp))
    views = []
    for v in range(n_views):
        k = int(rng.integers(0, len(grids)))
        w = max(8, int(round(int(rng.integers(128, 1025)) * size_scale)))
        h = max(8, int(round(int(rng.integers(128, 769)) * size_scale)))
        extent = grids[k] * 0.03125
        radius = float(rng.uniform(1.2, 2.0)) * extent
        height = float(rng.uniform(0.1, 0.5)) * extent
        start = float(rng.uniform(0.0, 360.0))
        sw = scenes[k]
        views.append((k, Workload(name=f"view_{v}", grid=sw.grid, chunk=sw.chunk, occupied=sw.occupied,
                                  emissive=sw.emissive, le=sw.le, seed=sw.seed, width=w, height=h, frames=16,
                                  fov_radians=math.radians(60.0),
                                  camera=_orbit_camera(radius, height, 1.0, start_deg=start),
                                  features=sw.features, taa=True, restir=sw.restir)))
    return scenes, views
