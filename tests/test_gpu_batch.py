"""Batched frames: one launch per pass over several views of several
scenes (with and without hipGraph capture) gives, view by view, the same
bits as rendering each view on its own."""
import numpy as np
import pytest

import ctamd
from helpers import bits, di_array, pres_array

pytestmark = pytest.mark.gpu


def _scene(w):
    sc = ctamd.Scene(ctamd.scene_desc(w))
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    first = ctamd.Ctx(8, 8)
    ctamd.load_scene(sc, acc, first)  # the accel holds the scene; contexts bind to it
    return sc, acc


def _views():
    scenes, views = ctamd.batched_views(n_views=5, seed=0xC3, grids=(64, 96), max_depth=2, size_scale=0.08)
    built = [_scene(s) for s in scenes]
    return scenes, views, built


def _ctx(sc, acc, w):
    ctx = ctamd.Ctx(w.width, w.height)
    ctamd.bind_scene(sc, ctx)
    return ctx


def _read(ctx, w):
    out = {"img": bits(ctx.read(ctamd.CT_BUF_IMAGE, w.width, w.height)),
           "di": di_array(ctx.read(ctamd.CT_BUF_DI, w.width, w.height)),
           "pdi": di_array(ctx.read(ctamd.CT_BUF_PREV_DI, w.width, w.height))}
    if w.restir.indirect:
        out["pres"] = pres_array(ctx.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR, w.width, w.height))
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_per_view(gpu, use_graph):
    scenes, views, built = _views()
    plain = [_ctx(*built[k], w) for k, w in views]
    # fresh contexts for the batch: both sides start with the same history
    batch_ctx = [_ctx(*built[k], w) for k, w in views]
    b = ctamd.Batch(batch_ctx, [built[k][1] for k, _ in views])
    for f in range(3):
        params = [ctamd.frame_params(w, f) for _, w in views]
        b.render(params, use_graph=use_graph)
        bs = b.stats()
        tot = {"shadow": 0, "bounce": 0, "pt": 0, "prim": 0}
        for (k, w), ctx, p in zip(views, plain, params):
            ctx.render(built[k][1], p)
            st = ctx.stats()
            tot["shadow"] += st.shadow_rays
            tot["bounce"] += st.bounce_rays
            tot["pt"] += st.shadow_rays_pt
            tot["prim"] += st.primary_rays
        for (k, w), pc, bc in zip(views, plain, batch_ctx):
            a, c = _read(pc, w), _read(bc, w)
            for key in a:
                assert np.array_equal(a[key], c[key]), (f, w.name, key)
        assert (bs.shadow_rays, bs.bounce_rays, bs.shadow_rays_pt, bs.primary_rays) == \
            (tot["shadow"], tot["bounce"], tot["pt"], tot["prim"])
        assert bs.ms_total > 0


@pytest.mark.skip(reason="the launch-shape tuner is switched off: one run of this test ended in a segmentation "
                         "fault of the host process inside ct_batch_autotune, cause not found")
def test_batch_autotune_keeps_results(gpu):
    """The tuner tries 1-4 view groups (concurrent streams) x 3-5 blocks/CU
    x 4 far/near splits of the ray queue; with its choice, a batch still
    gives each view's bits."""
    scenes, views, built = _views()
    ctxs = [_ctx(*built[k], w) for k, w in views]
    plain = [_ctx(*built[k], w) for k, w in views]
    b = ctamd.Batch(ctxs, [built[k][1] for k, _ in views])
    p0 = [ctamd.frame_params(w, 0) for _, w in views]
    t = b.autotune(p0, repeats=1)
    assert t.blocks_per_cu in (3, 4, 5) and t.groups in (1, 2, 3, 4) and t.ms > 0 and t.candidates == 48
    assert t.far_ray in (np.float32(1.0), np.float32(2.0), np.float32(4.0), np.float32(1e30))
    # the tuner rendered frame 0 twice per candidate; give the plain views the same history
    for _ in range(2 * t.candidates):
        for (k, w), ctx, p in zip(views, plain, p0):
            ctx.render(built[k][1], p)
    for f in (1, 2):
        ps = [ctamd.frame_params(w, f) for _, w in views]
        b.render(ps, use_graph=True)
        for (k, w), ctx, p in zip(views, plain, ps):
            ctx.render(built[k][1], p)
        for (k, w), pc, bc in zip(views, plain, ctxs):
            a, c = _read(pc, w), _read(bc, w)
            for key in a:
                assert np.array_equal(a[key], c[key]), (f, t.groups, w.name, key)


def test_batch_mixing_shift_mappings(gpu):
    """Views with the hybrid shift beside views with the reconnection shift
    in one batch (the hybrid kernels, their retrace requests in the group's
    segment queues and the retrace kernel over several frames' pixels): four
    frames, each view's bits equal its own render."""
    scenes, views = ctamd.batched_views(n_views=6, seed=0xC3, grids=(64, 96), max_depth=3, size_scale=0.15)
    built = [_scene(s) for s in scenes]
    for n, (k, w) in enumerate(views):
        w.restir = ctamd.RestirParams.from_buffer_copy(bytes(w.restir))
        w.restir.pt_shift_mapping = ctamd.CT_SHIFT_HYBRID if n % 2 == 0 else ctamd.CT_SHIFT_RECONNECTION
    plain = [_ctx(*built[k], w) for k, w in views]
    batch_ctx = [_ctx(*built[k], w) for k, w in views]
    b = ctamd.Batch(batch_ctx, [built[k][1] for k, _ in views])
    long_rc = 0
    for f in range(4):
        params = [ctamd.frame_params(w, f) for _, w in views]
        b.render(params, use_graph=True)
        for (k, w), ctx, p in zip(views, plain, params):
            ctx.render(built[k][1], p)
        for n, ((k, w), pc, bc) in enumerate(zip(views, plain, batch_ctx)):
            a, c = _read(pc, w), _read(bc, w)
            for key in a:
                assert np.array_equal(a[key], c[key]), (f, n, w.name, key)
            if n % 2 == 0:
                pres = a["pres"]
                long_rc += int((((pres[:, 2] >> 4) & 0xF) >= 2)[pres[:, 1].view(np.float32) > 0].sum())
    assert long_rc > 0  # the hybrid views hold paths with reconnection lengths above 1
