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


def all_buffers_equal(ctx, ofr, w, f):
    """Every buffer of frame f of a workload, GPU against the oracle, bit for
    bit (reservoirs of DI misses excluded: they hold no sample); returns the
    hit mask."""
    n = w.width * w.height
    for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
        assert np.array_equal(di_array(ctx.read(which, w.width, w.height)), di_array(ofr.read(which))), (f, which)
    hit = di_array(ofr.read(ctamd.CT_BUF_DI))[:, 3].view(np.float32) > 0
    for which in (ctamd.CT_BUF_INT_RESERVOIR, ctamd.CT_BUF_PREV_RESERVOIR):
        g, o = res_array(ctx.read(which, w.width, w.height)), res_array(ofr.read(which))
        assert np.array_equal(g[hit], o[hit]), (f, which)
    if w.restir.indirect:
        for which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
            assert np.array_equal(pres_array(ctx.read(which, w.width, w.height)), pres_array(ofr.read(which))), \
                (f, which)
        assert np.array_equal(bits(ctx.read(ctamd.CT_BUF_INDIRECT, w.width, w.height)),
                              bits(ofr.read(ctamd.CT_BUF_INDIRECT))), f
    assert np.array_equal(bits(ctx.read(ctamd.CT_BUF_VELOCITY, w.width, w.height)[hit]),
                          bits(ofr.read(ctamd.CT_BUF_VELOCITY)[hit])), f
    if w.taa and f > 0:
        assert np.array_equal(bits(ctx.read(ctamd.CT_BUF_TAA_FRAME, w.width, w.height)),
                              bits(ofr.read(ctamd.CT_BUF_TAA_FRAME))), f
    img = ctx.read(ctamd.CT_BUF_IMAGE, w.width, w.height)
    assert img.shape == (n, 4)
    assert np.array_equal(bits(img), bits(ofr.read(ctamd.CT_BUF_IMAGE))), f
    gs, os_ = ctx.stats(), ofr.stats()
    assert (gs.primary_rays, gs.shadow_rays, gs.bounce_rays, gs.shadow_rays_pt) == \
        (os_.primary_rays, os_.shadow_rays, os_.bounce_rays, os_.shadow_rays_pt), f
    return hit
