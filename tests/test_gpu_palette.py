"""The 512^3 palette workload on the GPU: ReSTIR PT with the hybrid shift, 4
bounces, path M-cap 20 and 3 spatial neighbours, on a sparse 512^3 scene
(16,106,127 voxels in 512 chunks of 64^3, a 256-entry palette with 3
emissive entries, 187,748 cube lights):

* three 1920x1080 orbit frames of the full workload, every buffer bit for
  bit against the oracle, with reconnection lengths above 1 among the
  frame's path reservoirs (the hybrid shift's random replay in use);
* the same workload on its 128^3 scene at 192x108 for 24 frames, where
  temporal reuse runs into the M-cap: bit for bit against the oracle; a
  path reservoir's confidence grows by one a frame until the history's
  share is capped at 20 (indirect_illumination.glsl:161), then stays at
  21.
"""
import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import all_buffers_equal, gpu_scene, oracle_scene, pres_array

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _k(pres):
    """Reconnection lengths of the reservoirs holding a path."""
    flags = pres[:, 2]
    w = pres[:, 1].view(np.float32)
    return ((flags >> 4) & 0xF)[w > 0]


def test_palette_512_1080p_frames_match_oracle(gpu):
    w = ctamd.palette_pt()
    assert (w.width, w.height, w.grid, w.chunk, w.palette) == (1920, 1080, (512,) * 3, (64,) * 3, (256, 3))
    assert w.restir.pt_shift_mapping == ctamd.CT_SHIFT_HYBRID and w.restir.max_depth == 4
    _say("palette 512^3: building the scene")
    sc, acc, ctx = gpu_scene(w)
    assert sc.info.primitive_count == 16106127 and sc.info.instance_count == 512
    assert sc.info.material_count == 256 and sc.info.cube_light_count == 187748
    _say("palette 512^3: oracle LBVH")
    ofr = O.OFrame(oracle_scene(sc, acc), w.width, w.height)
    for f in range(3):
        _say("palette 1080p frame", f)
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        hit = all_buffers_equal(ctx, ofr, w, f)
        assert hit.mean() > 0.2
        k = _k(pres_array(ofr.read(ctamd.CT_BUF_PATH_RESERVOIR)))
        assert (k == 1).sum() > 10000 and (k >= 2).sum() > 100, np.bincount(k)


def test_palette_128_frames_reach_the_m_cap(gpu):
    w = ctamd.palette_pt(grid=128, chunk=32, width=192, height=108, frames=24)
    sc, acc, ctx = gpu_scene(w)
    ofr = O.OFrame(oracle_scene(sc, acc), w.width, w.height)
    for f in range(24):
        if f % 8 == 0:
            _say("palette 128^3 frame", f)
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        all_buffers_equal(ctx, ofr, w, f)
        m = pres_array(ctx.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR, w.width, w.height))[:, 0].view(np.float32)
        assert m.max() == min(f + 1, 21), (f, m.max())
