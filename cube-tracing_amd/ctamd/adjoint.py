"""Differentiable voxel rendering: a torch.autograd.Function over
ct_adjoint_render / ct_adjoint_backward (path-replay adjoint, O(1) memory
per path), and the inverse-rendering workload on a 128^3 grid; and
frame_adjoint, the gradient of a rendered ReSTIR frame with its final
reservoirs held fixed (ct_frame_adjoint).

Parameters are per voxel: albedo [slots, 3] and scalar emission [slots],
indexed by the accel's primitive slot (``VoxelSlots`` maps between the
scene's voxel order and the slots).  Torch only carries device memory and
streams here; the forward and adjoint kernels are the library's.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import (AdjointDesc, AdjointView, Accel, Scene, SceneDesc, CT_VOXEL_EXTENT, camera_matrices, check, lib)

__all__ = ["AdjointWorkload", "inverse_rendering_c4", "c4_problem", "fit", "orbit_views", "VoxelSlots", "DiffRenderer", "render",
           "build_scene", "frame_adjoint"]


def orbit_views(n, width, height, center, radius, height_offset, fov_radians=math.radians(60.0), start_deg=0.0):
    """n cameras evenly spaced on a circle around `center`, looking at it."""
    views = []
    for k in range(n):
        a = math.radians(start_deg + 360.0 * k / n)
        pos = (center[0] + radius * math.sin(a), center[1] + height_offset, center[2] + radius * math.cos(a))
        f = np.array(center, np.float64) - np.array(pos, np.float64)
        f /= np.linalg.norm(f)
        iv, ip = camera_matrices(pos, tuple(f.tolist()), (0.0, 1.0, 0.0), fov_radians, width / height)
        v = AdjointView()
        v.width, v.height = width, height
        v.inv_view[:] = iv.tolist()
        v.inv_proj[:] = ip.tolist()
        views.append(v)
    return views


class VoxelSlots:
    """Scene voxel order <-> accel primitive slots."""

    def __init__(self, scene: Scene, accel: Accel):
        inst = accel.read_instances()
        first = np.array([i.first_primitive_index for i in inst[: scene.info.instance_count]], np.int64)
        counts = np.array([scene.instances[i].primitive_count for i in range(scene.info.instance_count)], np.int64)
        self.slots = np.concatenate([f + np.arange(c) for f, c in zip(first, counts)])
        self.capacity = int(self.slots.max()) + 1 if self.slots.size else 0
        self.scene_first = np.array([scene.instances[i].first_primitive_index
                                     for i in range(scene.info.instance_count)], np.int64)

    def scatter(self, values, fill=0.0):
        """Scene-order values [n, ...] -> slot-indexed array [capacity, ...]."""
        values = np.asarray(values, np.float32)
        out = np.full((self.capacity,) + values.shape[1:], fill, np.float32)
        out[self.slots] = values
        return out

    def gather(self, slot_values):
        return np.asarray(slot_values)[self.slots]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _on_torch_stream(call):
    """Runs call(stream) in order with torch's current stream: on that stream,
    or, when torch works on the null stream (handle 0, which the library's
    own non-blocking streams do not wait for and which does not wait for
    them), on the library's stream between two device synchronisations."""
    import torch
    h = torch.cuda.current_stream().cuda_stream
    if h:
        return call(C.c_void_p(h))
    torch.cuda.synchronize()
    r = call(None)
    torch.cuda.synchronize()
    return r


class DiffRenderer:
    """Views, sample count, depth and seed of one differentiable render."""

    def __init__(self, accel: Accel, views, spp=16, max_depth=3, seed=0xC4):
        self.accel = accel
        self.views = (AdjointView * len(views))(*views)
        self.n_views = len(views)
        self.desc = AdjointDesc(len(views), spp, max_depth, seed)
        self.n_pixels = sum(v.width * v.height for v in views)

    def forward(self, albedo, emission):
        import torch
        img = torch.empty((self.n_pixels, 3), dtype=torch.float32, device=albedo.device)
        _on_torch_stream(lambda st: check(lib().ct_adjoint_render(
            self.accel.handle, C.byref(self.desc), self.views, _ptr(albedo), _ptr(emission), _ptr(img), st)))
        return img

    def backward(self, albedo, emission, grad_image):
        import torch
        ga = torch.zeros_like(albedo)
        ge = torch.zeros_like(emission)
        _on_torch_stream(lambda st: check(lib().ct_adjoint_backward(
            self.accel.handle, C.byref(self.desc), self.views, _ptr(albedo), _ptr(emission), _ptr(grad_image),
            _ptr(ga), _ptr(ge), st)))
        return ga, ge


def path_trace(accel: Accel, views, materials, spp=16, max_depth=3, seed=0xC4):
    """Plain path tracing with the material callables (ct_path_trace) on the
    current torch device and stream; returns the image [sum of H*W][3].
    `materials`: a list of Material, a ctypes Material array, or None to
    reuse the set the previous call uploaded."""
    import torch
    from . import Material
    vv = (AdjointView * len(views))(*views)
    desc = AdjointDesc(len(views), spp, max_depth, seed)
    if materials is None:  # the set uploaded by an earlier call
        mm, nm = None, 0
    elif isinstance(materials, C.Array):
        mm, nm = materials, len(materials)
    else:
        mm, nm = (Material * max(len(materials), 1))(*materials), len(materials)
    n = sum(v.width * v.height for v in views)
    img = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    _on_torch_stream(lambda st: check(lib().ct_path_trace(accel.handle, C.byref(desc), vv, mm, nm, _ptr(img), st)))
    return img


def frame_adjoint(ctx, accel: Accel, params, grad_image, n_materials):
    """Gradients of sum(grad_image * shading output) of the frame `ctx` last
    rendered with `params`, its final reservoirs held fixed
    (ct_frame_adjoint), with respect to every material's diffuse albedo and
    emission.  grad_image: a float32 device tensor [H*W][4].  Returns
    (grad_diffuse, grad_emission), float32 device tensors [n_materials][3],
    on the current torch stream."""
    import torch
    g = grad_image.contiguous()
    if g.dtype != torch.float32 or g.numel() != 4 * params.width * params.height:
        raise ValueError("grad_image must be float32 [H*W][4]")
    gd = torch.zeros((n_materials, 3), dtype=torch.float32, device=g.device)
    ge = torch.zeros((n_materials, 3), dtype=torch.float32, device=g.device)
    torch.cuda.synchronize()  # the frame ran on the context's own stream
    _on_torch_stream(lambda st: check(lib().ct_frame_adjoint(ctx.handle, accel.handle, C.byref(params), _ptr(g),
                                                             _ptr(gd), _ptr(ge), st)))
    return gd, ge


def _function():
    import torch

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, albedo, emission, renderer):
            albedo = albedo.contiguous()
            emission = emission.contiguous()
            ctx.renderer = renderer
            ctx.save_for_backward(albedo, emission)
            return renderer.forward(albedo, emission)

        @staticmethod
        def backward(ctx, grad_image):
            albedo, emission = ctx.saved_tensors
            ga, ge = ctx.renderer.backward(albedo, emission, grad_image.contiguous())
            return ga, ge, None

    return _Render


_RENDER = None


def render(renderer: DiffRenderer, albedo, emission):
    """Image [pixels, 3] (views one after another) with autograd."""
    global _RENDER
    if _RENDER is None:
        _RENDER = _function()
    return _RENDER.apply(albedo, emission, renderer)


@dataclass
class AdjointWorkload:
    grid: int = 128
    chunk: int = 64
    fill: float = 0.24
    emissive_frac: float = 0.02
    width: int = 512
    height: int = 512
    views: int = 8
    spp: int = 16
    max_depth: int = 3
    target_seed: int = 0xC4
    init_seed: int = 0xC5
    fd_seed: int = 0xC6


def build_scene(w: AdjointWorkload):
    """Geometry and the hidden parameters (scene-order albedo, emission)."""
    d = SceneDesc()
    d.grid[:] = [w.grid] * 3
    d.chunk[:] = [w.chunk] * 3
    d.occupied_count = round(w.fill * w.grid ** 3)
    d.emissive_count = round(w.emissive_frac * d.occupied_count)
    d.le_min, d.le_max = 1.0, 20.0
    d.albedo_min, d.albedo_max = 0.2, 0.8
    d.seed = w.target_seed
    d.palette_size = 0
    d.voxel_extent = CT_VOXEL_EXTENT
    sc = Scene(d)
    mats = sc.materials
    mi = sc.primitives[:, 0]
    alb = np.array([[m.diffuse[0], m.diffuse[1], m.diffuse[2]] for m in mats], np.float32)[mi]
    emi = np.array([m.emission[0] for m in mats], np.float32)[mi]
    return sc, alb, emi


def initial_parameters(w: AdjointWorkload, hidden_albedo, hidden_emission):
    """Albedo U[0.2, 0.8] and emission scaled by U[0.5, 1.5] on the emissive voxels (seed 0xC5)."""
    rng = np.random.default_rng(w.init_seed)
    alb = rng.uniform(0.2, 0.8, hidden_albedo.shape).astype(np.float32)
    emi = (hidden_emission * rng.uniform(0.5, 1.5, hidden_emission.shape)).astype(np.float32)
    return alb, emi


def fit(r, target, albedo, emission, steps, lr_albedo=0.02, lr_emission=0.5, warmup=10, log=None):
    """Adam on (albedo, emission) against `target` with an L2 loss; the
    learning rates ramp linearly over the first `warmup` steps (without the
    ramp, emission's first full-size step overshoots and the loss rises).
    Albedo is clamped to [0, 1] and emission to >= 0 after each step.
    Returns the loss of every step."""
    import torch
    opt = torch.optim.Adam([{"params": [albedo], "lr": lr_albedo}, {"params": [emission], "lr": lr_emission}])
    hist = []
    for it in range(steps):
        s = min(1.0, (it + 1) / warmup) if warmup else 1.0
        opt.param_groups[0]["lr"] = lr_albedo * s
        opt.param_groups[1]["lr"] = lr_emission * s
        opt.zero_grad()
        img = render(r, albedo, emission)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            emission.clamp_(min=0.0)
        hist.append(float(loss.detach()))
        if log and (it % 20 == 0 or it == steps - 1):
            log(f"step {it}: loss {hist[-1]:.6g}")
    return hist


def c4_problem(w: AdjointWorkload | None = None):
    """The inverse-rendering workload set up on the GPU: the renderer over
    its views, the target images of the hidden parameters, and the initial
    parameters as leaf tensors (per voxel slot)."""
    import torch
    w = w or AdjointWorkload()
    sc, alb_t, emi_t = build_scene(w)
    acc = Accel(max_instances=sc.info.instance_count + 8,
                max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                max_cube_lights=max(sc.info.cube_light_count, 1) + 8)
    _load_geometry(sc, acc)
    slots = VoxelSlots(sc, acc)
    extent = w.grid * CT_VOXEL_EXTENT
    views = orbit_views(w.views, w.width, w.height, (0.0, 0.0, 0.0), 1.6 * extent, 0.4 * extent)
    r = DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    target = r.forward(torch.from_numpy(slots.scatter(alb_t)).to(dev), torch.from_numpy(slots.scatter(emi_t)).to(dev))
    a0, e0 = initial_parameters(w, alb_t, emi_t)
    albedo = torch.from_numpy(slots.scatter(a0)).to(dev).requires_grad_(True)
    emission = torch.from_numpy(slots.scatter(e0)).to(dev).requires_grad_(True)
    return {"workload": w, "scene": sc, "accel": acc, "slots": slots, "renderer": r, "target": target,
            "albedo": albedo, "emission": emission, "hidden": (alb_t, emi_t)}


def inverse_rendering_c4(steps=200, w: AdjointWorkload | None = None, lr_albedo=0.02, lr_emission=0.5, warmup=10,
                         log=None):
    """Fits albedo and emission to target images rendered from the hidden
    parameters with an L2 loss and Adam; returns the loss history."""
    P = c4_problem(w)
    return fit(P["renderer"], P["target"], P["albedo"], P["emission"], steps, lr_albedo, lr_emission, warmup, log)


def _load_geometry(sc: Scene, acc: Accel):
    """Voxels into the accel (instances, AABBs, one BUILD task) without a frame context."""
    from . import Task, CT_TASK_BUILD_BLAS_FROM_CPU
    st = acc.staging()
    ni, np_ = sc.info.instance_count, sc.info.primitive_count
    for i in range(ni):
        st.instances[i] = sc.instances[i]
    prims = sc.primitives.copy()
    prims[:, 1] = 0xFFFFFFFF  # no light list: the adjoint reads emission from its parameters
    C.memmove(st.aabbs, sc.aabbs.ctypes.data, 24 * np_)
    C.memmove(st.primitives, prims.ctypes.data, 8 * np_)
    t = Task()
    t.type = CT_TASK_BUILD_BLAS_FROM_CPU
    t.u.blas_build_from_cpu.instance_count = ni
    t.u.blas_build_from_cpu.primitive_count = np_
    for k in range(16):
        t.u.blas_build_from_cpu.transform[k] = 1.0 if k % 5 == 0 else 0.0
    acc.task_queue_add(t)
    acc.update_scene(True)
