"""Scenes loaded from .vox files on the HIP path: frames bit-identical to
the oracle, and the loaded voxels traced against brute force."""
import numpy as np
import pytest

import ctamd
from helpers import bits, gpu_scene, oracle_scene, random_rays
from test_gpu_parity import _frames_match
from ctamd.voxfile import monument

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chunk,indirect", [((0, 0, 0), False), ((16, 16, 16), True)])
def test_vox_scene_frames_match_oracle(gpu, chunk, indirect):
    sc = ctamd.Scene.from_vox(monument(n=40), chunk=chunk)
    assert sc.info.cube_light_count > 0
    w = ctamd.vox_view(sc, width=80, height=48, frames=3, indirect=indirect, max_depth=2)
    print("vox frames", chunk, indirect, sc.info.instance_count, sc.info.primitive_count, flush=True)
    _frames_match(w, 3, scene=sc)


def test_vox_scene_trace_closest_brute_force(gpu):
    sc = ctamd.Scene.from_vox(monument(n=48, seed=9), chunk=(16, 16, 16))
    w = ctamd.vox_view(sc, width=8, height=8, frames=1)
    _, acc, _ = gpu_scene(w, scene=sc)
    osc = oracle_scene(sc, acc)
    rng = np.random.default_rng(21)
    o, d = random_rays(rng, 4000, sc.grid_min, sc.grid_max)
    tg, ig = acc.trace_closest(o, d)
    tb, ib = osc.trace_closest(o[:500], d[:500], brute=True)
    assert np.array_equal(ib, ig[:500]) and np.array_equal(bits(tb), bits(tg[:500]))
    assert (ig[:, 0] != ctamd.CT_INVALID_INDEX).mean() > 0.2


def test_vox_model_loaded_with_transform(gpu):
    """load_model(name, transform): the transform goes into the build task;
    rays hit the moved voxels exactly as the oracle's brute force says."""
    sc = ctamd.Scene.from_vox(monument(n=24, seed=4), chunk=(8, 8, 8))
    w = ctamd.vox_view(sc, width=8, height=8, frames=1)
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 64 * sc.info.instance_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctx = ctamd.Ctx(w.width, w.height)
    tr = np.eye(4, dtype=np.float32)
    tr[3, :3] = (0.5, -0.25, 1.0)  # column-major: translation in the last column
    ctamd.load_scene(sc, acc, ctx, transform=tr)
    inst = acc.read_instances()
    assert len(inst) == sc.info.instance_count
    assert all(np.array_equal(np.array(i.transform[:], np.float32), tr.reshape(16)) for i in inst)
    osc = oracle_scene(sc, acc)
    rng = np.random.default_rng(5)
    off = np.array([0.5, -0.25, 1.0], np.float32)
    o, d = random_rays(rng, 3000, sc.grid_min + off, sc.grid_max + off)
    tg, ig = acc.trace_closest(o, d)
    tb, ib = osc.trace_closest(o[:400], d[:400], brute=True)
    assert np.array_equal(ib, ig[:400]) and np.array_equal(bits(tb), bits(tg[:400]))
    assert (ig[:, 0] != ctamd.CT_INVALID_INDEX).mean() > 0.2
