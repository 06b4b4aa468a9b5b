"""Edge cases on the GPU: zero rays, tiny and ragged frames bit-exact
against the oracle, and the argument and state errors of the frame, brush
and replay entry points."""
import ctypes as C

import numpy as np
import pytest
import torch

import ctamd
from helpers import gpu_scene
from test_gpu_parity import _frames_match

pytestmark = pytest.mark.gpu


def test_zero_rays(gpu):
    sc, acc, ctx = gpu_scene(ctamd.small_plain(width=8, height=8))
    z = np.zeros((0, 3), np.float32)
    t, ids = acc.trace_closest(z, z)
    assert t.shape == (0,) and ids.shape == (0, 2)
    assert acc.trace_any(z, z, np.zeros(0, np.float32)).shape == (0,)
    assert acc.trace_bounce(z, z)["ids"].shape[0] == 0


@pytest.mark.parametrize("width,height", [(1, 1), (17, 9), (3, 40)])
def test_tiny_and_ragged_frames_match_oracle(gpu, width, height):
    print("frame", width, height, flush=True)
    _frames_match(ctamd.small_plain(width=width, height=height, frames=3, max_depth=2), 3)


def test_state_and_argument_errors(gpu):
    L = ctamd.lib()
    w = ctamd.small_plain(width=16, height=8, frames=2)
    empty = ctamd.Accel(max_instances=8, max_primitives=1024, max_cube_lights=8)
    ctx = ctamd.Ctx(16, 8)
    p = ctamd.frame_params(w, 0)
    assert L.ct_frame_render(ctx.handle, empty.handle, C.byref(p), None) == ctamd.CT_E_STATE  # nothing built
    sc, acc, ctx2 = gpu_scene(w)
    big = ctamd.frame_params(ctamd.small_plain(width=32, height=8), 0)
    assert L.ct_frame_render(ctx2.handle, acc.handle, C.byref(big), None) == ctamd.CT_E_INVALID  # above the context
    p1 = ctamd.frame_params(w, 1)
    p1.is_active |= ctamd.CT_TAA_BIT
    p1.has_previous = 0
    assert L.ct_frame_render(ctx2.handle, acc.handle, C.byref(p1), None) == ctamd.CT_E_INVALID  # TAA without history
    # brush and replay before any frame, and outside the frame
    assert L.ct_brush_apply(ctx2.handle, acc.handle, 0, 0, 1, None) == ctamd.CT_E_STATE
    ctx2.render(acc, p)
    assert L.ct_brush_apply(ctx2.handle, acc.handle, 16, 0, 1, None) == ctamd.CT_E_INVALID
    p_di = ctamd.frame_params(w, 0)
    p_di.restir.indirect = 0  # parameters without paths: refused before the output is touched
    assert L.ct_pt_random_replay(ctx2.handle, acc.handle, C.byref(p_di), 0, 0, None, None) == ctamd.CT_E_INVALID
    # A device buffer of the full output size, so that a refusal that came
    # too late would still write in bounds.
    out = torch.zeros(16 * 8 * 8, dtype=torch.int32, device="cuda")
    optr = C.c_void_p(out.data_ptr())
    assert L.ct_pt_random_replay(ctx2.handle, acc.handle, C.byref(p_di), 0, 0, optr, None) == ctamd.CT_E_STATE
    # the frame's own argument checks apply to a replay as well
    for field, value, code in (("max_depth", 99, ctamd.CT_E_INVALID), ("pt_neighbor_count", 99, ctamd.CT_E_INVALID),
                               ("pt_shift_mapping", 7, ctamd.CT_E_UNSUPPORTED),
                               ("max_neighbors", 99, ctamd.CT_E_INVALID)):
        bad = ctamd.frame_params(w, 0)
        setattr(bad.restir, field, value)
        assert L.ct_pt_random_replay(ctx2.handle, acc.handle, C.byref(bad), 0, 0, optr, None) == code, field
    torch.cuda.synchronize()
    assert int(out.abs().sum()) == 0
    assert L.ct_brush_mark(acc.handle, None, 3, None) == ctamd.CT_E_INVALID
