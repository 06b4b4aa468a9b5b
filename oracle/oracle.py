"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker.  The boundary struct layouts are the ones of
include/ct_abi.h (taken from the ctamd binding, which mirrors that header).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_ROOT, "cube-tracing_amd"))
import ctamd  # noqa: E402  (struct layouts only)
from ctamd import (Aabb, Instance, Primitive, Material, Light, LightConfig, FrameParams,  # noqa: E402
                   FrameStats, TraceStats, BvhArrays, Reservoir, DiInfo, PathReservoir, AdjointDesc,
                   AdjointView)

LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
f32, u32 = C.c_float, C.c_uint32


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sig = {
            "oc_tea": ([u32, u32], u32),
            "oc_lcg": ([P(u32)], u32),
            "oc_rnd": ([P(u32)], f32),
            "oc_urnd_interval": ([P(u32), u32, u32], u32),
            "oc_compute_ray_origin": ([P(f32), P(f32), P(f32)], None),
            "oc_intersect_box": ([P(f32), f32, P(f32), P(f32), C.c_int, C.c_int, P(f32), P(f32)], C.c_int),
            "oc_cube_like_normal": ([P(f32), P(f32)], None),
            "oc_sky_color": ([f32, C.c_int, P(f32), P(f32)], None),
            "oc_morton3": ([u32, u32, u32], u32),
            "oc_lbvh_build": ([u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p], C.c_int),
            "oc_scene_create": ([u32, P(Instance), C.c_void_p, u32, C.c_void_p, C.c_void_p, u32, P(Material),
                                 u32, P(Light), u32, P(Light), u32, P(Light), P(LightConfig)], C.c_void_p),
            "oc_scene_destroy": ([C.c_void_p], None),
            "oc_scene_read_bvh": ([C.c_void_p, u32, P(BvhArrays)], C.c_int),
            "oc_trace_closest": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, f32, C.c_int, C.c_void_p, C.c_void_p,
                                  P(TraceStats)], None),
            "oc_trace_any": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                              C.c_void_p, P(TraceStats)], None),
            "oc_trace_bounce": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, P(TraceStats)], None),
            "oc_sincos": ([f32, P(f32), P(f32)], None),
            "oc_adjoint_render": ([C.c_void_p, P(AdjointDesc), P(AdjointView), C.c_void_p, C.c_void_p, C.c_void_p],
                                  C.c_int),
            "oc_adjoint_backward": ([C.c_void_p, P(AdjointDesc), P(AdjointView), C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "oc_random_cosine_direction": ([P(u32), P(f32), P(f32)], None),
            "oc_path_trace": ([C.c_void_p, P(AdjointDesc), P(AdjointView), P(Material), u32, C.c_void_p], C.c_int),
            "oc_scatter_material": ([P(Material), P(f32), P(f32), P(u32), P(f32), P(f32), P(C.c_int)], None),
            "oc_refraction": ([P(f32), P(f32), f32, P(f32)], None),
            "oc_reflectance": ([f32, f32], f32),
            "oc_frame_create": ([C.c_void_p, u32, u32], C.c_void_p),
            "oc_frame_set_scene": ([C.c_void_p, C.c_void_p], C.c_int),
            "oc_scene_set_previous": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, u32, C.c_void_p, u32],
                                      C.c_int),
            "oc_frame_destroy": ([C.c_void_p], None),
            "oc_frame_render": ([C.c_void_p, P(FrameParams), C.c_int, u32, u32], C.c_int),
            "oc_frame_read": ([C.c_void_p, u32, C.c_void_p, C.c_uint64], C.c_int),
            "oc_frame_reshade": ([C.c_void_p, P(FrameParams), C.c_void_p, C.c_void_p, u32, C.c_void_p], C.c_int),
            "oc_frame_adjoint": ([C.c_void_p, P(FrameParams), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "oc_frame_get_stats": ([C.c_void_p, P(FrameStats)], C.c_int),
            "oc_frame_set_previous_scene": ([C.c_void_p, C.c_void_p], C.c_int),
            "oc_probe_ris": ([C.c_void_p, P(FrameParams), P(DiInfo), f32, u32, C.c_void_p, C.c_void_p], C.c_int),
            "oc_probe_temporal": ([C.c_void_p, P(FrameParams), P(DiInfo), P(Reservoir), P(DiInfo), P(Reservoir), u32,
                                   P(Reservoir)], C.c_int),
            "oc_probe_pairwise": ([C.c_void_p, P(FrameParams), P(DiInfo), P(Reservoir), u32, C.c_void_p,
                                   C.c_void_p, u32, P(Reservoir)], C.c_int),
            "oc_probe_retrace": ([C.c_void_p, P(FrameParams), P(DiInfo), P(PathReservoir), P(C.c_int), P(f32), P(f32),
                                  P(f32), P(u32), P(f32)], C.c_int),
            "oc_probe_shift": ([C.c_void_p, P(FrameParams), P(DiInfo), P(DiInfo), C.c_int, P(PathReservoir),
                                P(f32), P(f32)], C.c_int),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _f3(v):
    return (f32 * 3)(*[float(x) for x in v])


def tea(v0, v1):
    return lib().oc_tea(v0 & 0xFFFFFFFF, v1 & 0xFFFFFFFF)


def lcg(state):
    s = u32(state & 0xFFFFFFFF)
    r = lib().oc_lcg(C.byref(s))
    return r, s.value


def rnd(state):
    s = u32(state & 0xFFFFFFFF)
    r = lib().oc_rnd(C.byref(s))
    return r, s.value


def urnd_interval(state, mn, mx):
    s = u32(state & 0xFFFFFFFF)
    r = lib().oc_urnd_interval(C.byref(s), mn, mx)
    return r, s.value


def compute_ray_origin(pos, nrm):
    o = (f32 * 3)()
    lib().oc_compute_ray_origin(_f3(pos), _f3(nrm), o)
    return np.array(o[:], np.float32)


def intersect_box(center, radius, origin, direction, can_start=True, oriented=True):
    t = f32()
    n = (f32 * 3)()
    hit = lib().oc_intersect_box(_f3(center), radius, _f3(origin), _f3(direction), int(can_start), int(oriented),
                                 C.byref(t), n)
    return bool(hit), t.value, np.array(n[:], np.float32)


def cube_like_normal(n):
    o = (f32 * 3)()
    lib().oc_cube_like_normal(_f3(n), o)
    return np.array(o[:], np.float32)


def sky_color(time, afternoon, d):
    o = (f32 * 3)()
    lib().oc_sky_color(time, int(afternoon), _f3(d), o)
    return np.array(o[:], np.float32)


def sincos(x):
    sv, cv = f32(), f32()
    lib().oc_sincos(float(x), C.byref(sv), C.byref(cv))
    return sv.value, cv.value


def random_cosine_direction(state, n):
    s = u32(state)
    o = (f32 * 3)()
    lib().oc_random_cosine_direction(C.byref(s), _f3(n), o)
    return np.array(o[:], np.float32), s.value


def morton3(x, y, z):
    return lib().oc_morton3(x, y, z)


def lbvh_build(centers, lo, hi):
    c = np.ascontiguousarray(centers, np.float32)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    n = c.shape[0]
    keys = np.zeros(n, np.uint32)
    ids = np.zeros(n, np.uint32)
    child = np.zeros(max(2 * (n - 1), 1), np.uint32)
    boxes = np.zeros(max(12 * (n - 1), 1), np.float32)
    parent = np.zeros(2 * n - 1, np.uint32)
    depth = lib().oc_lbvh_build(n, c.ctypes.data, lo.ctypes.data, hi.ctypes.data, keys.ctypes.data, ids.ctypes.data,
                                child.ctypes.data, boxes.ctypes.data, parent.ctypes.data)
    return {"keys": keys, "ids": ids, "child": child[: 2 * (n - 1)], "boxes": boxes[: 12 * (n - 1)],
            "parent": parent, "depth": depth}


def light_config_for(n_cube, n_point=0, n_env=0):
    """load_lights (main.cpp:418-438) with brdf_count = 1 when cube lights
    exist (main.cpp:1317-1320); pdfs use the reference's integer division."""
    c = LightConfig()
    c.point_light_count, c.cube_light_count, c.env_map_count = n_point, n_cube, n_env
    c.brdf_count = 1 if n_cube > 0 else 0
    c.light_count = n_point + n_cube + n_env
    ipdf = (lambda k: 0.0 if k == 0 else float(k // c.light_count))
    c.point_light_pdf, c.cube_light_pdf, c.env_map_pdf = ipdf(n_point), ipdf(n_cube), ipdf(n_env)
    c.brdf_pdf = 0.0 if (c.brdf_count == 0 or n_cube == 0) else float(c.brdf_count // n_cube)
    return c


class OScene:
    """Oracle scene: instances, voxels, materials, lights and its own LBVH."""

    def __init__(self, instances, world_to_object, aabbs, primitives, materials, cube_lights,
                 point_lights=(), env_lights=(), config=None):
        self.n_inst = len(instances)
        self._inst = (Instance * self.n_inst)(*instances)
        self._w2o = np.ascontiguousarray(world_to_object, np.float32).reshape(self.n_inst, 16)
        self._aabbs = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
        self._prims = np.ascontiguousarray(primitives, np.uint32).reshape(-1, 2)
        self._mats = (Material * len(materials))(*materials)
        self._cube = (Light * max(len(cube_lights), 1))(*cube_lights)
        self._point = (Light * max(len(point_lights), 1))(*point_lights)
        self._env = (Light * max(len(env_lights), 1))(*env_lights)
        self.config = config if config is not None else light_config_for(len(cube_lights), len(point_lights),
                                                                          len(env_lights))
        self.handle = lib().oc_scene_create(
            self.n_inst, self._inst, self._w2o.ctypes.data, self._aabbs.shape[0], self._aabbs.ctypes.data,
            self._prims.ctypes.data, len(materials), self._mats, len(cube_lights), self._cube, len(point_lights),
            self._point, len(env_lights), self._env, C.byref(self.config))
        if not self.handle:
            raise ValueError("oc_scene_create rejected the scene")

    @classmethod
    def from_scene(cls, scene, world_to_object=None, **kw):
        ni = scene.info.instance_count
        w2o = world_to_object
        if w2o is None:
            w2o = np.tile(np.eye(4, dtype=np.float32).reshape(16), (ni, 1))
        return cls(list(scene.instances), w2o, scene.aabbs, scene.primitives, list(scene.materials),
                   list(scene.lights), **kw)

    def set_previous(self, prev_transforms=None, prev_first=None, remap_prims=None, remap_lights=None):
        """The previous frame's instance transforms [n, 16] and first
        indices [n] (None: the current ones) and the remapping tables."""
        self._prev_t = None if prev_transforms is None else np.ascontiguousarray(prev_transforms, np.float32)
        self._prev_f = None if prev_first is None else np.ascontiguousarray(prev_first, np.uint32)
        self._rp = np.ascontiguousarray(remap_prims if remap_prims is not None else [], np.uint32)
        self._rl = np.ascontiguousarray(remap_lights if remap_lights is not None else [], np.uint32)
        rc = lib().oc_scene_set_previous(self.handle, None if self._prev_t is None else self._prev_t.ctypes.data,
                                         None if self._prev_f is None else self._prev_f.ctypes.data,
                                         self._rp.ctypes.data if self._rp.size else None, self._rp.size,
                                         self._rl.ctypes.data if self._rl.size else None, self._rl.size)
        if rc != 0:
            raise ValueError(f"oc_scene_set_previous: {rc}")

    # ------------------------------------------------------------ probes
    def probe_ris(self, params, di, seeds, confidence=0.0):
        """RIS with its visibility pass at the shading point `di` (a DiInfo),
        once per seed: (n, 9) uint32 reservoir words (Y, type, seed, W_y,
        W_sum, M, F.rgb; floats as their bits)."""
        seeds = np.ascontiguousarray(seeds, np.uint32)
        out = (Reservoir * len(seeds))()
        rc = lib().oc_probe_ris(self.handle, C.byref(params), C.byref(di), confidence, len(seeds),
                                seeds.ctypes.data, C.addressof(out))
        assert rc == 0, rc
        return np.frombuffer(bytes(out), np.uint32).reshape(len(seeds), 9).copy()

    def probe_temporal(self, params, di, res, prev_di, prev_res, seed):
        """Temporal reuse of `res` at `di` with the previous frame's
        reservoir and DI at the predicted pixel."""
        out = Reservoir()
        rc = lib().oc_probe_temporal(self.handle, C.byref(params), C.byref(di), C.byref(res), C.byref(prev_di),
                                     C.byref(prev_res), seed, C.byref(out))
        assert rc == 0, rc
        return out

    def probe_pairwise(self, params, center, res, neighbours, neighbour_res, seed):
        """Pairwise-MIS combination of spatial reuse for `res` at `center`
        with the given neighbour DiInfos and Reservoirs."""
        k = len(neighbours)
        nd = (DiInfo * k)(*neighbours)
        nr = (Reservoir * k)(*neighbour_res)
        out = Reservoir()
        rc = lib().oc_probe_pairwise(self.handle, C.byref(params), C.byref(center), C.byref(res), k,
                                     C.addressof(nd), C.addressof(nr), seed, C.byref(out))
        assert rc == 0, rc
        return out

    def probe_shift(self, params, dst, src, res, normalize_dst_wo=True):
        """Shift of path reservoir `res` from primary vertex `src` to `dst`:
        (jacobian, integrand rgb)."""
        jac = f32(0.0)
        f = (f32 * 3)()
        rc = lib().oc_probe_shift(self.handle, C.byref(params), C.byref(dst), C.byref(src),
                                  1 if normalize_dst_wo else 0, C.byref(res), C.byref(jac), f)
        assert rc == 0, rc
        return float(jac.value), np.array(f[:], np.float32)

    def probe_retrace(self, params, dst, res):
        """The hybrid shift's retrace of `res` from primary vertex `dst`:
        dict(ok, pos, nrm, wo, ids, tp) of the vertex before the
        reconnection vertex."""
        ok = C.c_int(0)
        pos, nrm, wo, tp = (f32 * 3)(), (f32 * 3)(), (f32 * 3)(), (f32 * 3)()
        ids = (u32 * 2)()
        rc = lib().oc_probe_retrace(self.handle, C.byref(params), C.byref(dst), C.byref(res), C.byref(ok), pos, nrm,
                                    wo, ids, tp)
        assert rc == 0, rc
        arr = lambda a: np.array(a[:], np.float32)
        return dict(ok=bool(ok.value), pos=arr(pos), nrm=arr(nrm), wo=arr(wo), ids=tuple(ids[:]), tp=arr(tp))

    def read_bvh(self, instance, n):
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.uint32)
        child = np.zeros(max(2 * (n - 1), 1), np.uint32)
        boxes = np.zeros(max(12 * (n - 1), 1), np.float32)
        parent = np.zeros(2 * n - 1, np.uint32)
        a = BvhArrays(keys.ctypes.data_as(C.POINTER(u32)), ids.ctypes.data_as(C.POINTER(u32)),
                      child.ctypes.data_as(C.POINTER(u32)), boxes.ctypes.data_as(C.POINTER(f32)),
                      parent.ctypes.data_as(C.POINTER(u32)), 0, 0, 0, 0)
        rc = lib().oc_scene_read_bvh(self.handle, instance, C.byref(a))
        assert rc == 0
        return {"keys": keys, "ids": ids, "child": child[: 2 * (n - 1)], "boxes": boxes[: 12 * (n - 1)],
                "parent": parent, "depth": a.depth, "root_ref": a.root_ref, "n": a.n}

    def trace_closest(self, origins, directions, tmax=1e9, brute=False, stats=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        t = np.zeros(n, np.float32)
        ids = np.zeros((n, 2), np.uint32)
        st = TraceStats()
        lib().oc_trace_closest(self.handle, n, o.ctypes.data, d.ctypes.data, tmax, int(brute), t.ctypes.data,
                               ids.ctypes.data, C.byref(st))
        return (t, ids, st) if stats else (t, ids)

    def trace_any(self, origins, directions, distances, targets=None, brute=False, stats=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        dist = np.ascontiguousarray(distances, np.float32)
        n = o.shape[0]
        tg = None if targets is None else np.ascontiguousarray(targets, np.uint32)
        occ = np.zeros(n, np.uint8)
        st = TraceStats()
        lib().oc_trace_any(self.handle, n, o.ctypes.data, d.ctypes.data, dist.ctypes.data,
                           None if tg is None else tg.ctypes.data, int(brute), occ.ctypes.data, C.byref(st))
        return (occ, st) if stats else occ

    def trace_bounce(self, origins, directions, brute=False, stats=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        ids = np.zeros((n, 2), np.uint32)
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        dist = np.zeros(n, np.float32)
        wo = np.zeros((n, 3), np.float32)
        st = TraceStats()
        lib().oc_trace_bounce(self.handle, n, o.ctypes.data, d.ctypes.data, int(brute), ids.ctypes.data,
                              pos.ctypes.data, nrm.ctypes.data, dist.ctypes.data, wo.ctypes.data, C.byref(st))
        out = {"ids": ids, "pos": pos, "nrm": nrm, "dist": dist, "wo": wo}
        if stats:
            out["stats"] = st
        return out

    def adjoint_render(self, desc, views, albedo, emission):
        """The differentiable path tracer: image [sum of H*W][3]."""
        a = np.ascontiguousarray(albedo, np.float32)
        e = np.ascontiguousarray(emission, np.float32)
        n = sum(v.width * v.height for v in views)
        img = np.zeros((n, 3), np.float32)
        vv = (AdjointView * len(views))(*views)
        rc = lib().oc_adjoint_render(self.handle, C.byref(desc), vv, a.ctypes.data, e.ctypes.data, img.ctypes.data)
        assert rc == 0
        return img

    def path_trace(self, desc, views, materials):
        """Plain path tracing with the material callables: image [sum of H*W][3]."""
        n = sum(v.width * v.height for v in views)
        img = np.zeros((n, 3), np.float32)
        vv = (AdjointView * len(views))(*views)
        mm = (Material * max(len(materials), 1))(*materials)
        rc = lib().oc_path_trace(self.handle, C.byref(desc), vv, mm, len(materials), img.ctypes.data)
        assert rc == 0
        return img

    def adjoint_backward(self, desc, views, albedo, emission, grad_image):
        a = np.ascontiguousarray(albedo, np.float32)
        e = np.ascontiguousarray(emission, np.float32)
        g = np.ascontiguousarray(grad_image, np.float32)
        ga = np.zeros_like(a)
        ge = np.zeros_like(e)
        vv = (AdjointView * len(views))(*views)
        rc = lib().oc_adjoint_backward(self.handle, C.byref(desc), vv, a.ctypes.data, e.ctypes.data, g.ctypes.data,
                                       ga.ctypes.data, ge.ctypes.data)
        assert rc == 0
        return ga, ge

    def close(self):
        if getattr(self, "handle", None):
            lib().oc_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OFrame:
    def __init__(self, scene: OScene, w, h):
        self.scene = scene
        self.w, self.h = w, h
        self.handle = lib().oc_frame_create(scene.handle, w, h)

    def set_scene(self, scene: OScene):
        """Render the following frames with `scene`, keeping the history."""
        self.scene = scene
        if lib().oc_frame_set_scene(self.handle, scene.handle) != 0:
            raise ValueError("oc_frame_set_scene rejected the scene")

    def set_previous_scene(self, scene):
        """The scene before the last update (an OScene, or None for the
        current one): temporal reuse traces its target-at-previous query
        there while CT_REMAP_BIT is set."""
        self.prev_scene = scene
        if lib().oc_frame_set_previous_scene(self.handle, None if scene is None else scene.handle) != 0:
            raise ValueError("oc_frame_set_previous_scene failed")

    def render(self, params, threads=0, rows=None):
        r0, r1 = (0, self.h) if rows is None else rows
        rc = lib().oc_frame_render(self.handle, C.byref(params), threads, r0, r1)
        if rc != 0:
            raise RuntimeError(f"oc_frame_render failed: {rc}")

    def read(self, which):
        n = self.w * self.h
        if which in (ctamd.CT_BUF_IMAGE, ctamd.CT_BUF_PREV_IMAGE, ctamd.CT_BUF_TAA_FRAME):
            out = np.zeros((n, 4), np.float32)
        elif which == ctamd.CT_BUF_VELOCITY:
            out = np.zeros((n, 2), np.float32)
        elif which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
            out = (DiInfo * n)()
        elif which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
            out = (PathReservoir * n)()
        elif which == ctamd.CT_BUF_INDIRECT:
            out = np.zeros((n, 3), np.float32)
        else:
            out = (Reservoir * n)()
        nbytes = out.nbytes if isinstance(out, np.ndarray) else C.sizeof(out)
        ptr = out.ctypes.data if isinstance(out, np.ndarray) else C.addressof(out)
        rc = lib().oc_frame_read(self.handle, which, ptr, nbytes)
        assert rc == 0
        return out

    def stats(self):
        s = FrameStats()
        lib().oc_frame_get_stats(self.handle, C.byref(s))
        return s

    def reshade(self, params, diffuse, emission):
        """The last frame's image [W*H][4] with its final reservoirs held
        fixed, for materials' diffuse and emission arrays [materials][3]."""
        d = np.ascontiguousarray(diffuse, np.float32)
        e = np.ascontiguousarray(emission, np.float32)
        out = np.zeros((self.w * self.h, 4), np.float32)
        rc = lib().oc_frame_reshade(self.handle, C.byref(params), d.ctypes.data, e.ctypes.data, d.shape[0],
                                    out.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"oc_frame_reshade failed: {rc}")
        return out

    def adjoint(self, params, grad_image, n_mats):
        """d sum(grad_image * reshade) / d (diffuse, emission) at the scene's
        own materials: two [n_mats][3] arrays."""
        g = np.ascontiguousarray(grad_image, np.float32).reshape(-1, 4)
        gd = np.zeros((n_mats, 3), np.float32)
        ge = np.zeros((n_mats, 3), np.float32)
        rc = lib().oc_frame_adjoint(self.handle, C.byref(params), g.ctypes.data, gd.ctypes.data, ge.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"oc_frame_adjoint failed: {rc}")
        return gd, ge

    def close(self):
        if getattr(self, "handle", None):
            lib().oc_frame_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
