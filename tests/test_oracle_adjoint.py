"""The oracle's differentiable path tracer: its path-replay gradients
against central finite differences of an L2 loss (the definition of the
derivative), and the closed form of the radiance of one path."""
import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd import adjoint as AD


def _setup(spp=2, depth=3, size=24, views=2):
    w = AD.AdjointWorkload(grid=64, chunk=64, width=size, height=size, views=views, spp=spp, max_depth=depth)
    sc, alb, emi = AD.build_scene(w)
    osc = O.OScene.from_scene(sc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    v = AD.orbit_views(w.views, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    desc = ctamd.AdjointDesc(len(v), w.spp, w.max_depth, w.target_seed)
    return w, sc, osc, v, desc, alb, emi


def _loss(osc, desc, v, alb, emi, target):
    img = osc.adjoint_render(desc, v, alb, emi).astype(np.float64)
    return float(((img - target) ** 2).mean())


@pytest.mark.parametrize("depth", [1, 3])
def test_gradients_match_central_differences(depth):
    w, sc, osc, v, desc, alb, emi = _setup(depth=depth)
    rng = np.random.default_rng(0xC6)
    # target from perturbed parameters so the loss gradient is not zero
    alb_t = np.clip(alb + rng.normal(0, 0.05, alb.shape), 0, 1).astype(np.float32)
    emi_t = (emi * rng.uniform(0.8, 1.2, emi.shape)).astype(np.float32)
    target = osc.adjoint_render(desc, v, alb_t, emi_t).astype(np.float64)
    img = osc.adjoint_render(desc, v, alb, emi).astype(np.float64)
    n = img.size
    gimg = (2.0 * (img - target) / n).astype(np.float32)
    ga, ge = osc.adjoint_backward(desc, v, alb, emi, gimg)
    eps = 1e-3
    checked = 0
    for kind, grad in (("e", ge), ("a", ga.reshape(-1))):
        idx = np.argsort(-np.abs(grad))[:12]
        for i in idx:
            if grad[i] == 0.0:
                continue
            a2, e2 = alb.copy(), emi.copy()
            arr = e2 if kind == "e" else a2.reshape(-1)
            base = arr[i]
            arr[i] = base + eps
            lp = _loss(osc, desc, v, a2, e2, target)
            arr[i] = base - eps
            lm = _loss(osc, desc, v, a2, e2, target)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grad[i]) <= 1e-3 * abs(grad[i]) + 1e-9, (kind, i, fd, grad[i])
            checked += 1
    assert checked >= 12


def test_emission_is_linear_and_albedo_enters_as_a_product():
    """Radiance is linear in every emission (zeroing all of them gives a
    black image, doubling them doubles it); depth 0 sees emission only."""
    w, sc, osc, v, desc, alb, emi = _setup(depth=3)
    img = osc.adjoint_render(desc, v, alb, emi)
    assert not np.any(osc.adjoint_render(desc, v, alb, np.zeros_like(emi)))
    img2 = osc.adjoint_render(desc, v, alb, 2 * emi)
    assert np.allclose(img2, 2 * img, rtol=1e-6, atol=0)
    d0 = ctamd.AdjointDesc(desc.n_views, desc.spp, 0, desc.seed)
    img0 = osc.adjoint_render(d0, v, alb, emi)
    img0b = osc.adjoint_render(d0, v, np.zeros_like(alb), emi)
    assert np.array_equal(img0, img0b)  # no scattering: albedo unused
    assert np.all(img >= img0 - 1e-7)   # more bounces only add light
