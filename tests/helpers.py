"""Shared test helpers: scenes on both sides, bitwise comparisons."""
import ctypes as C

import numpy as np

import ctamd
import oracle as O


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def struct_array(arr, dtype=np.uint8):
    return np.frombuffer(bytes(arr), dtype)


def gpu_scene(w, max_prims=None):
    sc = ctamd.Scene(ctamd.scene_desc(w))
    n = sc.info.primitive_count
    acc = ctamd.Accel(max_instances=max(sc.info.instance_count, 1) + 8,
                      max_primitives=max_prims or (n + 64 * sc.info.instance_count + 1024),
                      max_cube_lights=max(sc.info.cube_light_count, 1))
    ctx = ctamd.Ctx(w.width, w.height)
    ctamd.load_scene(sc, acc, ctx)
    return sc, acc, ctx


def oracle_scene(sc, acc=None):
    ni = sc.info.instance_count
    w2o = None
    if acc is not None:
        w2o = np.stack([acc.world_to_object(i) for i in range(ni)])
    return O.OScene.from_scene(sc, world_to_object=w2o)


def random_rays(rng, n, lo, hi, inside_frac=0.2):
    """Rays from a box around the scene toward random targets inside it."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    ext = hi - lo
    o = lo - 0.5 * ext + rng.random((n, 3)) * 2.0 * ext
    k = int(n * inside_frac)
    o[:k] = lo + rng.random((k, 3)) * ext
    tgt = lo + rng.random((n, 3)) * ext
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def di_array(arr):
    return np.frombuffer(bytes(arr), np.uint32).reshape(len(arr), 15)


def res_array(arr):
    return np.frombuffer(bytes(arr), np.uint32).reshape(len(arr), 9)


def pres_array(arr):
    return np.frombuffer(bytes(arr), np.uint32).reshape(len(arr), 23)
