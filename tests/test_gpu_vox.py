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
