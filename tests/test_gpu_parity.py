"""HIP path against the oracle, bit for bit (integer ids, sorted arrays and
fp32 values compared as bit patterns)."""
import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import bits, gpu_scene, oracle_scene, random_rays, di_array, res_array, pres_array

pytestmark = pytest.mark.gpu


def _compare_bvh(g, o, label):
    n = o["n"]
    assert g["n"] == n, label
    assert np.array_equal(g["keys"], o["keys"]), f"{label}: sorted Morton keys differ"
    assert np.array_equal(g["ids"], o["ids"]), f"{label}: sorted primitive ids differ"
    assert np.array_equal(g["child"], o["child"]), f"{label}: Karras topology differs"
    assert np.array_equal(bits(g["boxes"]), bits(o["boxes"])), f"{label}: refit boxes differ"
    assert np.array_equal(g["parent"], o["parent"]), f"{label}: parents differ"
    assert g["depth"] == o["depth"] and g["root_ref"] == o["root_ref"], label


def test_lbvh_arrays_small(gpu):
    w = ctamd.small_plain(width=32, height=32)
    sc, acc, ctx = gpu_scene(w)
    osc = oracle_scene(sc, acc)
    n = sc.info.primitive_count
    _compare_bvh(acc.read_bvh(0, n), osc.read_bvh(0, n), "blas 0")
    _compare_bvh(acc.read_bvh(ctamd.CT_INVALID_INDEX, 1), osc.read_bvh(ctamd.CT_INVALID_INDEX, 1), "tlas")


def test_lbvh_arrays_chunked(gpu):
    w = ctamd.hd_di(width=64, height=36, grid=96, chunk=32)
    sc, acc, ctx = gpu_scene(w)
    osc = oracle_scene(sc, acc)
    ni = sc.info.instance_count
    assert ni == 27
    for i in range(ni):
        n = sc.instances[i].primitive_count
        _compare_bvh(acc.read_bvh(i, n), osc.read_bvh(i, n), f"blas {i}")
    _compare_bvh(acc.read_bvh(ctamd.CT_INVALID_INDEX, ni), osc.read_bvh(ctamd.CT_INVALID_INDEX, ni), "tlas")


@pytest.mark.parametrize("grid,chunk", [(64, 64), (96, 32)])
def test_trace_closest_bit_exact(gpu, grid, chunk):
    w = ctamd.hd_di(width=32, height=32, grid=grid, chunk=chunk)
    sc, acc, ctx = gpu_scene(w)
    osc = oracle_scene(sc, acc)
    rng = np.random.default_rng(7)
    o, d = random_rays(rng, 20000, sc.grid_min, sc.grid_max)
    tg, ig, sg = acc.trace_closest(o, d, stats=True)
    to, io, so = osc.trace_closest(o, d, stats=True)
    assert np.array_equal(ig, io), "hit ids differ"
    assert np.array_equal(bits(tg), bits(to)), "hit t differs"
    assert (sg.nodes, sg.leaves, sg.instances) == (so.nodes, so.leaves, so.instances)
    # brute force over every primitive on a subset
    tb, ib = osc.trace_closest(o[:400], d[:400], brute=True)
    assert np.array_equal(ib, ig[:400]) and np.array_equal(bits(tb), bits(tg[:400]))


def test_trace_any_bit_exact(gpu):
    w = ctamd.hd_di(width=32, height=32, grid=96, chunk=32)
    sc, acc, ctx = gpu_scene(w)
    osc = oracle_scene(sc, acc)
    rng = np.random.default_rng(11)
    o, d = random_rays(rng, 20000, sc.grid_min, sc.grid_max, inside_frac=0.5)
    dist = (rng.random(o.shape[0]) * 3.0).astype(np.float32)
    tg = np.full((o.shape[0], 2), ctamd.CT_INVALID_INDEX, np.uint32)
    lights = np.array([[l.instance_info.instance_id, l.instance_info.primitive_id] for l in sc.lights], np.uint32)
    tg[::3] = lights[rng.integers(0, len(lights), size=tg[::3].shape[0])]
    g, sg = acc.trace_any(o, d, dist, tg, stats=True)
    r, so = osc.trace_any(o, d, dist, tg, stats=True)
    assert np.array_equal(g, r)
    assert (sg.nodes, sg.leaves, sg.instances) == (so.nodes, so.leaves, so.instances)
    b = osc.trace_any(o[:400], d[:400], dist[:400], tg[:400], brute=True)
    assert np.array_equal(b, g[:400])
    assert 0 < g.mean() < 1


def _frames_match(w, frames, scene=None, check=None):
    sc, acc, ctx = gpu_scene(w, scene=scene)
    assert ctx.light_config().cube_light_count == sc.info.cube_light_count
    osc = oracle_scene(sc, acc)
    ofr = O.OFrame(osc, w.width, w.height)
    for f in range(frames):
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        gi, oi = ctx.read(ctamd.CT_BUF_IMAGE, w.width, w.height), ofr.read(ctamd.CT_BUF_IMAGE)
        for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
            assert np.array_equal(di_array(ctx.read(which, w.width, w.height)), di_array(ofr.read(which))), (f, which)
        for which in (ctamd.CT_BUF_INT_RESERVOIR, ctamd.CT_BUF_PREV_RESERVOIR):
            ga, oa = res_array(ctx.read(which, w.width, w.height)), res_array(ofr.read(which))
            hit = di_array(ofr.read(ctamd.CT_BUF_DI))[:, 3].view(np.float32) > 0
            assert np.array_equal(ga[hit], oa[hit]), (f, which)
        if w.restir.indirect:
            for which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
                gp = pres_array(ctx.read(which, w.width, w.height))
                op = pres_array(ofr.read(which))
                bad = np.nonzero((gp != op).any(1))[0]
                assert bad.size == 0, (f, which, bad[:8], gp[bad[:2]], op[bad[:2]])
            gind, oind = ctx.read(ctamd.CT_BUF_INDIRECT, w.width, w.height), ofr.read(ctamd.CT_BUF_INDIRECT)
            assert np.array_equal(bits(gind), bits(oind)), f"frame {f}: indirect colour differs"
        assert np.array_equal(bits(gi), bits(oi)), f"frame {f}: radiance differs"
        gs, os_ = ctx.stats(), ofr.stats()
        assert gs.primary_rays == os_.primary_rays
        assert gs.shadow_rays == os_.shadow_rays, (f, gs.shadow_rays, os_.shadow_rays)
        assert gs.bounce_rays == os_.bounce_rays, (f, gs.bounce_rays, os_.bounce_rays)
        assert gs.shadow_rays_pt == os_.shadow_rays_pt, (f, gs.shadow_rays_pt, os_.shadow_rays_pt)
        if check:
            check(f, ctx, ofr)
    return sc


def test_frame_parity_small(gpu):
    _frames_match(ctamd.small_plain(width=64, height=64, frames=4), 4)


def test_frame_parity_reference_mode_taa(gpu):
    _frames_match(ctamd.small_plain(reference_mode=True, width=72, height=40, frames=4, taa=True), 4)


def test_frame_parity_chunked_orbit(gpu):
    _frames_match(ctamd.hd_di(width=96, height=64, grid=96, chunk=32, frames=3), 3)


@pytest.mark.parametrize("kind", ["direct", "pt"])
def test_frame_parity_ris_brdf(gpu, kind):
    """RIS with its BRDF strategy (CT_RIS_BRDF_BIT, restir_resampling.glsl:
    190-252): a closest-hit ray per shaded pixel inside the RIS pass, every
    buffer and ray count bit for bit against the oracle.  A fifth of the
    voxels emit (many lights make the strategy's candidates competitive), so
    frame 0 (no temporal history) holds reservoirs whose BRDF candidate won
    the merge: W_y = 0 with F > 0, no visibility pass."""
    w = ctamd.hd_di(width=96, height=64, grid=96, chunk=32, frames=3, indirect=kind == "pt")
    w.emissive *= 10
    w.features |= ctamd.CT_RIS_BRDF_BIT

    def check(f, ctx, ofr):
        st = ctx.stats()
        hits = int((di_array(ofr.read(ctamd.CT_BUF_DI))[:, 3].view(np.float32) > 0).sum())
        if not w.restir.indirect:
            assert st.bounce_rays == hits  # brdf_count = 1 ray per shaded pixel
        if f == 0:
            r = res_array(ctx.read(ctamd.CT_BUF_INT_RESERVOIR, w.width, w.height))
            wy, fr = r[:, 3].view(np.float32), r[:, 6:9].view(np.float32)
            assert ((wy == 0) & (fr.max(axis=1) > 0)).sum() > 10

    _frames_match(w, w.frames, check=check)


@pytest.mark.parametrize("grid,chunk", [(64, 64), (96, 32)])
def test_trace_bounce_bit_exact(gpu, grid, chunk):
    """intersect() path segments: ids, hit record and node counts."""
    w = ctamd.hd_di(width=32, height=32, grid=grid, chunk=chunk)
    sc, acc, ctx = gpu_scene(w)
    osc = oracle_scene(sc, acc)
    rng = np.random.default_rng(13)
    o, d = random_rays(rng, 20000, sc.grid_min, sc.grid_max, inside_frac=0.5)
    g = acc.trace_bounce(o, d, stats=True)
    r = osc.trace_bounce(o, d, stats=True)
    assert np.array_equal(g["ids"], r["ids"]), "hit ids differ"
    for k in ("pos", "nrm", "dist", "wo"):
        assert np.array_equal(bits(g[k]), bits(r[k])), k
    gs, rs = g["stats"], r["stats"]
    assert (gs.nodes, gs.leaves, gs.instances) == (rs.nodes, rs.leaves, rs.instances)
    b = osc.trace_bounce(o[:400], d[:400], brute=True)
    assert np.array_equal(b["ids"], g["ids"][:400])
    assert np.array_equal(bits(b["pos"]), bits(g["pos"][:400]))
    assert 0.2 < (g["ids"][:, 0] != ctamd.CT_INVALID_INDEX).mean() < 1.0


def _pt_workload(max_depth=3, radius=1, shift=ctamd.CT_SHIFT_RECONNECTION, rr=0, temporal=1, spatial=1):
    w = ctamd.hd_di(width=48, height=40, grid=96, chunk=32, frames=3, max_depth=max_depth)
    w.restir.pt_neighbor_radius = radius
    w.restir.pt_shift_mapping = shift
    w.restir.pt_russian_roulette = rr
    w.restir.pt_temporal = temporal
    w.restir.pt_spatial = spatial
    return w


@pytest.mark.parametrize("kw", [dict(max_depth=1), dict(max_depth=2), dict(max_depth=4, radius=3),
                                dict(shift=ctamd.CT_SHIFT_HYBRID), dict(rr=1), dict(temporal=0),
                                dict(spatial=0)],
                         ids=["depth1", "depth2", "depth4_radius3", "hybrid", "russian_roulette", "no_temporal",
                              "no_spatial"])
def test_frame_parity_pt_variants(gpu, kw):
    _frames_match(_pt_workload(**kw), 3)


@pytest.mark.parametrize("brdf", [False, True], ids=["nee", "ris_brdf"])
def test_split_frame_equals_whole_frame(gpu, brdf):
    """A frame of 2^20 pixels or more runs as two halves (even and odd tile
    rows) on two streams by default; it gives the bits of the same frame run
    whole on one stream (per-pass events on). The height leaves a partial
    last tile row."""
    w = ctamd.hd_di(width=1024, height=1032, grid=64, chunk=32, frames=4, indirect=True, max_depth=2)
    if brdf:
        w.features |= ctamd.CT_RIS_BRDF_BIT
    sc, acc, whole = gpu_scene(w)
    split = ctamd.Ctx(w.width, w.height)
    ctamd.bind_scene(sc, split)
    whole.set_instrumentation(1)
    for f in range(4):
        p = ctamd.frame_params(w, f)
        whole.render(acc, p)
        split.render(acc, p)
        for which in (ctamd.CT_BUF_IMAGE, ctamd.CT_BUF_TAA_FRAME, ctamd.CT_BUF_INDIRECT):
            assert np.array_equal(bits(whole.read(which, w.width, w.height)),
                                  bits(split.read(which, w.width, w.height))), (f, which)
        for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
            assert np.array_equal(di_array(whole.read(which, w.width, w.height)),
                                  di_array(split.read(which, w.width, w.height))), (f, which)
        for which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
            assert np.array_equal(pres_array(whole.read(which, w.width, w.height)),
                                  pres_array(split.read(which, w.width, w.height))), (f, which)
        a, b = whole.stats(), split.stats()
        assert (a.shadow_rays, a.bounce_rays, a.shadow_rays_pt, a.primary_rays) == \
            (b.shadow_rays, b.bounce_rays, b.shadow_rays_pt, b.primary_rays), f
