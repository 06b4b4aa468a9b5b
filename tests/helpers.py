"""Shared test helpers: scenes on both sides, bitwise comparisons."""
import ctypes as C

import numpy as np

import ctamd
import oracle as O


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def struct_array(arr, dtype=np.uint8):
    return np.frombuffer(bytes(arr), dtype)


def gpu_scene(w, max_prims=None, scene=None):
    sc = scene if scene is not None else ctamd.Scene(ctamd.scene_desc(w))
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


_HIP = None


def hip_runtime():
    """The HIP runtime library this process already runs (the one torch
    and libctamd.so share), for raw device copies in tests."""
    global _HIP
    if _HIP is None:
        ctamd.lib()
        path = None
        with open("/proc/self/maps") as fh:
            for line in fh:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "no HIP runtime loaded"
        _HIP = C.CDLL(path)
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def to_device(dst_ptr, host_array):
    a = np.ascontiguousarray(host_array)
    assert hip_runtime().hipMemcpy(dst_ptr, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice


def oracle_scene_from_accel(acc, materials, config):
    """Oracle scene over the library's committed records (its slot layout,
    unused instance ids included with no primitives)."""
    inst = acc.read_instances()
    W = acc.read_world()
    w2o = np.tile(np.eye(4, dtype=np.float32).reshape(16), (len(inst), 1))
    for i, r in enumerate(inst):
        if r.primitive_count:
            w2o[i] = acc.world_to_object(i)
    lights = W["lights"][: W["light_count"]]
    return O.OScene(list(inst), w2o, W["aabbs"], W["prims"], list(materials), lights, config=config), inst, W
