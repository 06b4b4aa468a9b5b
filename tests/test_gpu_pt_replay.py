"""Random replay (trace_random_replay_path_hybrid_simple) on the GPU: a
path reservoir's own seed replayed from its own primary hit reaches the
reservoir's reconnection vertex exactly; a neighbour's seed either reaches
a reconnection vertex or is reported not invertible with nothing set."""
import numpy as np
import pytest

import ctamd
from helpers import gpu_scene, pres_array

pytestmark = pytest.mark.gpu


def _frame():
    w = ctamd.hd_di(width=64, height=40, grid=48, chunk=16, frames=2, max_depth=4)
    w.restir.pt_temporal = 0
    w.restir.pt_spatial = 0  # the output reservoir is each pixel's own path
    sc, acc, ctx = gpu_scene(w)
    p = ctamd.frame_params(w, 0)
    ctx.render(acc, p)
    return w, sc, acc, ctx, p


def test_own_seed_replays_the_own_reconnection_vertex(gpu):
    w, sc, acc, ctx, p = _frame()
    res = pres_array(ctx.read(ctamd.CT_BUF_PATH_RESERVOIR, w.width, w.height))
    out = ctamd.pt_random_replay(ctx, acc, p)
    plen, rlen = res[:, 2] & 0xF, (res[:, 2] >> 4) & 0xF
    assert np.array_equal(out[:, 2][out[:, 3] > 0], plen[out[:, 3] > 0])
    has_rc = (res[:, 0].view(np.float32) > 0) & (rlen >= 1) & (rlen <= plen) & (res[:, 12] < ctamd.CT_MAX_INSTANCES)
    assert has_rc.mean() > 0.05, has_rc.mean()
    sel = np.nonzero(has_rc)[0]
    print("pixels with a reconnection vertex", sel.size, flush=True)
    assert (out[sel, 0] == 1).all()
    assert np.array_equal(out[sel, 1], rlen[sel])
    assert np.array_equal(out[sel, 4:6], res[sel, 12:14])   # instance, primitive
    assert np.array_equal(out[sel, 6:9], res[sel, 14:17])   # hit position, bit for bit


def test_neighbour_seed_is_replayed_or_rejected(gpu):
    w, sc, acc, ctx, p = _frame()
    out = ctamd.pt_random_replay(ctx, acc, p, 1, 0)
    ok = out[:, 0] == 1
    assert ok.any()
    assert np.array_equal(out[ok, 1], out[ok, 3])             # reached exactly the reconnection length
    bad = ~ok & (out[:, 3] > 0)
    assert (out[bad, 4] == ctamd.CT_MAX_INSTANCES).all() and (out[bad, 9] == ctamd.CT_MAX_INSTANCES).all()
    # the last column has no right neighbour: nothing replayed
    assert not out.reshape(w.height, w.width, 12)[:, -1, 3].any()
