"""The oracle's ReSTIR PT pieces against what fixes them without the
reference: libm for the sin/cos the build defines, the cosine-weighted
hemisphere density for the scatter directions, brute force for the path
segment closest hit, and the rules of the reference's path builder that
make some contributions vanish."""
import math

import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import bits, oracle_scene, random_rays, pres_array


def test_sincos_against_libm():
    xs = np.linspace(0.0, 2.0 * math.pi, 200001, endpoint=False).astype(np.float32)
    # quadrant boundaries and their float neighbours
    edges = np.array([k * math.pi / 4 for k in range(8)], np.float32)
    xs = np.concatenate([xs, edges, np.nextafter(edges, np.float32(10)), np.nextafter(edges, np.float32(-10))])
    xs = xs[(xs >= 0) & (xs < np.float32(2 * math.pi))]
    err = 0.0
    for x in xs[::7]:
        s, c = O.sincos(x)
        err = max(err, abs(s - math.sin(float(x))), abs(c - math.cos(float(x))))
        assert abs(s * s + c * c - 1.0) < 5e-7
    assert err < 2e-7, err


@pytest.mark.parametrize("n", [(0, 0, 1), (0, 0, -1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                               (0.48, -0.6, 0.64)])
def test_cosine_direction_density(n):
    """random_cosine_direction: unit vectors in the normal's hemisphere with
    p(w) = cos(theta)/pi, so z = cos(theta) has density 2z (E[z] = 2/3,
    E[z^2] = 1/2) and the azimuth is uniform."""
    n = np.asarray(n, np.float32)
    n = n / np.linalg.norm(n)
    s = 2024
    zs, ds = [], []
    for _ in range(20000):
        d, s = O.random_cosine_direction(s, n)
        ds.append(d)
        zs.append(float(np.dot(d, n)))
    ds = np.array(ds)
    zs = np.array(zs)
    assert np.all(np.abs(np.linalg.norm(ds, axis=1) - 1.0) < 2e-6)
    assert zs.min() >= 0.0
    assert abs(zs.mean() - 2.0 / 3.0) < 0.01
    assert abs((zs ** 2).mean() - 0.5) < 0.01
    # the tangential part averages to zero in every direction (uniform azimuth)
    tang = ds - zs[:, None] * n[None, :]
    assert np.all(np.abs(tang.mean(0)) < 0.02)


@pytest.mark.parametrize("grid,chunk", [(64, 64), (96, 32)])
def test_bounce_traversal_equals_brute_force(grid, chunk):
    w = ctamd.hd_di(width=16, height=16, grid=grid, chunk=chunk)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    osc = oracle_scene(sc)
    rng = np.random.default_rng(5)
    o, d = random_rays(rng, 1500, sc.grid_min, sc.grid_max, inside_frac=0.5)
    t = osc.trace_bounce(o, d)
    b = osc.trace_bounce(o, d, brute=True)
    for k in ("ids", "pos", "nrm", "dist", "wo"):
        assert np.array_equal(bits(t[k]) if k != "ids" else t[k], bits(b[k]) if k != "ids" else b[k]), k
    hit = t["ids"][:, 0] != ctamd.CT_INVALID_INDEX
    assert 0.2 < hit.mean() < 1.0
    # the hit lies on the voxel's face the normal names, at the reported distance
    nrm = t["nrm"][hit]
    assert np.all(np.abs(nrm).sum(1) == 1.0)
    assert np.allclose(np.linalg.norm(t["pos"][hit] - o[hit], axis=1), t["dist"][hit], rtol=1e-5, atol=1e-6)


def _render(w, frames):
    sc = ctamd.Scene(ctamd.scene_desc(w))
    osc = oracle_scene(sc)
    fr = O.OFrame(osc, w.width, w.height)
    out = []
    for f in range(frames):
        fr.render(ctamd.frame_params(w, f))
        out.append({"img": fr.read(ctamd.CT_BUF_IMAGE), "ind": fr.read(ctamd.CT_BUF_INDIRECT),
                    "pres": fr.read(ctamd.CT_BUF_PATH_RESERVOIR), "stats": fr.stats()})
    return out


def test_depth_one_adds_no_indirect_light():
    """With max_depth 1 the only path segment ends at the first secondary
    vertex, whose emission the path tracer leaves to direct lighting
    (path_tracer.glsl:609-611), and an escape before it is no candidate
    (path_builder.glsl:46): the indirect term is exactly zero."""
    w = ctamd.hd_di(width=24, height=20, grid=96, chunk=32, frames=2, max_depth=1)
    for r in _render(w, 2):
        assert r["stats"].bounce_rays > 0
        assert not np.any(r["ind"])
    w3 = ctamd.hd_di(width=24, height=20, grid=96, chunk=32, frames=1, max_depth=3)
    assert np.any(_render(w3, 1)[0]["ind"] > 0)


def test_path_segments_bounded_by_depth():
    for depth in (1, 2, 3):
        w = ctamd.hd_di(width=24, height=20, grid=96, chunk=32, frames=1, max_depth=depth)
        r = _render(w, 1)[0]
        hits = r["stats"].primary_rays  # upper bound on paths
        assert 0 < r["stats"].bounce_rays <= depth * hits


def test_temporal_history_grows_by_one_per_frame():
    """Static camera: the reprojected history is the pixel itself and the
    Talbot merge adds the history's M (capped at 20 x 1) to the new path's 1."""
    w = ctamd.small_plain(width=24, height=24, frames=5, max_depth=3)
    out = _render(w, 5)
    for f, r in enumerate(out):
        M = pres_array(r["pres"])[:, 0].view(np.float32)
        assert M.max() == f + 1, (f, M.max())


def test_pt_deterministic_across_threads():
    w = ctamd.hd_di(width=32, height=24, grid=96, chunk=32, frames=2, max_depth=3)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    osc = oracle_scene(sc)
    res = []
    for threads in (1, 8):
        fr = O.OFrame(osc, w.width, w.height)
        for f in range(2):
            fr.render(ctamd.frame_params(w, f), threads=threads)
        res.append((bits(fr.read(ctamd.CT_BUF_IMAGE)), pres_array(fr.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR))))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
