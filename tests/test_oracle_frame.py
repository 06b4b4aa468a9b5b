"""Invariants of the oracle's ReSTIR DI frame (CPU only)."""
import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import di_array, res_array, bits


def _run(w, frames, mutate=None):
    sc = ctamd.Scene(ctamd.scene_desc(w))
    osc = O.OScene.from_scene(sc)
    fr = O.OFrame(osc, w.width, w.height)
    out = []
    for f in range(frames):
        p = ctamd.frame_params(w, f)
        if mutate:
            mutate(p)
        fr.render(p, threads=4)
        out.append({k: fr.read(k) for k in (ctamd.CT_BUF_IMAGE, ctamd.CT_BUF_DI, ctamd.CT_BUF_INT_RESERVOIR,
                                           ctamd.CT_BUF_PREV_RESERVOIR, ctamd.CT_BUF_VELOCITY)})
        out[-1]["stats"] = fr.stats()
    return sc, out


@pytest.fixture(scope="module")
def small():
    return _run(ctamd.small_plain(width=48, height=48), 3)


def test_static_camera_reprojects_to_own_pixel(small):
    sc, out = small
    for o in out:
        di = di_array(o[ctamd.CT_BUF_DI])
        hit = di[:, 3].view(np.float32) > 0
        v = o[ctamd.CT_BUF_VELOCITY][hit]
        assert np.all(np.abs(v) < 1e-3)  # motion vector ~ 0 for a static camera and scene


def test_sky_pixels_are_the_daylight_gradient(small):
    sc, out = small
    w = ctamd.small_plain(width=48, height=48)
    p = ctamd.frame_params(w, 0)
    img = out[0][ctamd.CT_BUF_IMAGE]
    di = di_array(out[0][ctamd.CT_BUF_DI])
    miss = di[:, 3].view(np.float32) <= 0
    iv = np.array(p.inv_view[:], np.float64).reshape(4, 4).T
    ip = np.array(p.inv_proj[:], np.float64).reshape(4, 4).T
    ys, xs = np.divmod(np.arange(48 * 48), 48)
    ndc = np.stack([(xs + 0.5) / 48 * 2 - 1, (ys + 0.5) / 48 * 2 - 1, np.ones_like(xs), np.ones_like(xs)], 1)
    t = (ip @ ndc.T).T[:, :3]
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    d = (iv[:3, :3] @ t.T).T
    u = d[:, 1] / np.linalg.norm(d, axis=1)
    s = 0.5 * (u + 1.0)
    expect = np.stack([1 - 0.5 * s, 1 - 0.3 * s, np.ones_like(s)], 1)
    assert miss.sum() > 100
    assert np.allclose(img[miss, :3], expect[miss], atol=2e-6)


def test_temporal_m_growth_law(small):
    """M after temporal reuse = M_RIS + min(M_prev, 20 * M_RIS) with M_RIS = candidates."""
    sc, out = small
    r1 = res_array(out[1][ctamd.CT_BUF_INT_RESERVOIR]).view(np.float32)
    r0p = res_array(out[0][ctamd.CT_BUF_PREV_RESERVOIR]).view(np.float32)
    di1 = di_array(out[1][ctamd.CT_BUF_DI])
    hit = di1[:, 3].view(np.float32) > 0
    conf = di1[:, 14].view(np.float32)
    W_prev = r0p[:, 3]
    # 32 cube candidates (workload value), one temporal merge with the own pixel (static camera)
    had_prev = hit & (W_prev > 0)
    expect = 32.0 + np.minimum(r0p[:, 5], 20.0 * 32.0)
    assert had_prev.sum() > 0
    assert np.array_equal(r1[had_prev, 5], expect[had_prev].astype(np.float32))
    assert np.all(r1[hit & ~had_prev, 5] == 32.0)
    assert np.allclose(conf[had_prev], np.clip(r0p[had_prev, 5] / 20.0, 0, 1))


def test_zero_neighbours_leaves_reservoir_unchanged():
    w = ctamd.small_plain(width=40, height=32)

    def no_neighbours(p):
        p.restir.min_neighbors = p.restir.max_neighbors = 0

    sc, out = _run(w, 2, no_neighbours)
    for o in out:
        di = di_array(o[ctamd.CT_BUF_DI])
        hit = di[:, 3].view(np.float32) > 0
        a = res_array(o[ctamd.CT_BUF_INT_RESERVOIR])[hit]
        b = res_array(o[ctamd.CT_BUF_PREV_RESERVOIR])[hit]
        assert np.array_equal(a, b)
        assert o["stats"].shadow_rays_spatial == 0


def test_frame_is_deterministic(small):
    sc, out2 = _run(ctamd.small_plain(width=48, height=48), 3)
    for a, b in zip(small[1], out2):
        assert np.array_equal(bits(a[ctamd.CT_BUF_IMAGE]), bits(b[ctamd.CT_BUF_IMAGE]))


def test_emission_added_on_emissive_voxels():
    w = ctamd.small_plain(width=64, height=64)
    d = ctamd.scene_desc(w)
    d.emissive_count = 4000  # many lights so some are seen directly
    sc = ctamd.Scene(d)
    osc = O.OScene.from_scene(sc)
    fr = O.OFrame(osc, w.width, w.height)
    p = ctamd.frame_params(w, 0)
    fr.render(p)
    img = fr.read(ctamd.CT_BUF_IMAGE)
    di = di_array(fr.read(ctamd.CT_BUF_DI))
    hit = di[:, 3].view(np.float32) > 0
    mat = di[:, 13]
    em = np.array([sc.materials[m].emission[0] if h else 0.0 for m, h in zip(mat, hit)], np.float32)
    seen = hit & (em > 0)
    assert seen.sum() > 10 and np.all(img[seen, 0] >= em[seen])
