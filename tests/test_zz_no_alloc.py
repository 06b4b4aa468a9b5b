"""ct_trace_any_noalloc, the allocation-free any-hit entry point, against
ct_trace_any on the same rays: the same occlusion bytes on two calls in a
row (the second smaller, so stale scratch contents would show), the size
limit refused with CT_E_CAPACITY, and nothing written for an empty call."""
import ctypes as C

import numpy as np
import pytest
import torch

import ctamd
from helpers import gpu_scene, random_rays

pytestmark = pytest.mark.gpu


def _noalloc(acc, o, d, dist, tg):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    to, td, tdist = dev(o), dev(d), dev(dist)
    ttg = None if tg is None else dev(tg.view(np.int32))
    occ = torch.full((o.shape[0],), 7, dtype=torch.uint8, device="cuda")
    rc = ctamd.lib().ct_trace_any_noalloc(acc.handle, o.shape[0], to.data_ptr(), td.data_ptr(), tdist.data_ptr(),
                                          None if ttg is None else ttg.data_ptr(), occ.data_ptr(), None)
    assert rc == ctamd.CT_OK, ctamd.lib().ct_last_error_message()
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def test_noalloc_any_hit_equals_the_allocating_entry_point(gpu):
    w = ctamd.hd_di(width=32, height=32, grid=64, chunk=32, frames=1)
    sc, acc, ctx = gpu_scene(w)
    rng = np.random.default_rng(5)
    for n in (200000, 3001):
        print("noalloc any-hit rays", n, flush=True)
        o, d = random_rays(rng, n, sc.grid_min, sc.grid_max, inside_frac=0.3)
        dist = rng.uniform(0.05, 3.0, n).astype(np.float32)
        t, ids = acc.trace_closest(o, d)
        tg = np.where(rng.random((n, 1)) < 0.5, ids, np.uint32(0xFFFFFFFF)).astype(np.uint32)
        for targets in (None, tg):
            ref = acc.trace_any(o, d, dist, targets)
            got = _noalloc(acc, o, d, dist, targets)
            assert np.array_equal(ref, got)
            assert 0.05 < ref.mean() < 0.95


def test_noalloc_limits(gpu):
    sc, acc, ctx = gpu_scene(ctamd.small_plain(width=8, height=8))
    L = ctamd.lib()
    occ = torch.full((4,), 7, dtype=torch.uint8, device="cuda")
    z = torch.zeros(12, dtype=torch.float32, device="cuda")
    assert L.ct_trace_any_noalloc(acc.handle, (1 << 20) + 1, z.data_ptr(), z.data_ptr(), z.data_ptr(), None,
                                  occ.data_ptr(), None) == ctamd.CT_E_CAPACITY
    assert L.ct_trace_any_noalloc(acc.handle, 0, None, None, None, None, None, None) == ctamd.CT_OK
    torch.cuda.synchronize()
    assert occ.cpu().tolist() == [7, 7, 7, 7]
