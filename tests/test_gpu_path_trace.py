"""Plain path tracing with the material callables (ct_path_trace) on the GPU:
bit for bit against the oracle on a scene mixing Lambertian, metal,
dielectric and constant-medium voxels, and equal to the adjoint's forward
render when every material is Lambertian with grey emission."""
import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd import adjoint as AD
from helpers import bits

pytestmark = pytest.mark.gpu


def _mixed_materials(sc, seed=5):
    rng = np.random.default_rng(seed)
    mats = []
    for k, m0 in enumerate(sc.materials):
        m = ctamd.Material.from_buffer_copy(bytes(m0))
        t = k % 4
        m.type = t
        if t == ctamd.CT_MATERIAL_METAL:
            m.specular[:] = m.diffuse[:]
            m.roughness = float(rng.uniform(0.0, 1.0))
        elif t == ctamd.CT_MATERIAL_DIELECTRIC:
            m.ior = float(rng.uniform(1.1, 1.8))
        mats.append(m)
    return mats


@pytest.mark.parametrize("depth", [1, 5])
def test_path_trace_matches_oracle(gpu, depth):
    w = ctamd.hd_di(width=40, height=24, grid=48, chunk=16, max_depth=2)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    ni = sc.info.instance_count
    acc = ctamd.Accel(max_instances=ni + 8, max_primitives=sc.info.primitive_count + 32 * ni + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(8, 8)
    ctamd.load_scene(sc, acc, ctx)
    mats = _mixed_materials(sc)
    extent = 48 * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(3, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    print("path trace", depth, sc.info.primitive_count, flush=True)
    img = AD.path_trace(acc, views, mats, spp=3, max_depth=depth, seed=0x77).cpu().numpy()
    osc = O.OScene.from_scene(sc)
    ref = osc.path_trace(ctamd.AdjointDesc(len(views), 3, depth, 0x77), views, mats)
    assert np.array_equal(bits(img), bits(ref))
    assert img.max() > 0 and np.isfinite(img).all()


def test_lambertian_path_trace_equals_adjoint_forward(gpu):
    import torch
    w = AD.AdjointWorkload(grid=48, chunk=16, width=32, height=24, views=2, spp=2, max_depth=3)
    sc, alb, emi = AD.build_scene(w)
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=max(sc.info.cube_light_count, 1) + 8)
    AD._load_geometry(sc, acc)
    slots = AD.VoxelSlots(sc, acc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(w.views, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    r = AD.DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    a = r.forward(torch.from_numpy(slots.scatter(alb)).to(dev), torch.from_numpy(slots.scatter(emi)).to(dev))
    # palette_size 0: voxel v has material v, Lambertian with diffuse alb[v] and emission (emi[v],) * 3
    mats = list(sc.materials)
    assert np.array_equal(sc.primitives[:, 0], np.arange(sc.info.primitive_count))
    p = AD.path_trace(acc, views, mats, spp=w.spp, max_depth=w.max_depth, seed=w.target_seed)
    assert np.array_equal(bits(a.cpu().numpy()), bits(p.cpu().numpy()))


def test_materials_kept_between_calls(gpu):
    w = ctamd.small_plain(width=24, height=16)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    acc = ctamd.Accel(max_instances=8, max_primitives=sc.info.primitive_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(8, 8)
    ctamd.load_scene(sc, acc, ctx)
    views = AD.orbit_views(1, 24, 16, (0, 0, 0), 3.0, 0.5)
    with pytest.raises(ctamd.CtError):
        AD.path_trace(acc, views, None, spp=1, max_depth=2)  # nothing uploaded yet
    mats = _mixed_materials(sc)
    a = AD.path_trace(acc, views, mats, spp=2, max_depth=2).cpu().numpy()
    b = AD.path_trace(acc, views, None, spp=2, max_depth=2).cpu().numpy()
    assert np.array_equal(bits(a), bits(b))
