"""The frame adjoint's model on the oracle: the image of the last rendered
ReSTIR frame with its final reservoirs held fixed (the checkpoint the frame
leaves in its buffers), as a function of every material's diffuse albedo d
and emission e; the image is the shading output (the TAA pass's input).  A pixel shows F_y(d, Le) W_y + I_ind + e, with the light
sample y replayed from the reservoir's seed, W_y and the visibility the
frame found kept, Le the emission of the light voxel's material, and the
indirect term I_ind kept.

* At the scene's own values the model reproduces the frame's image bit for
  bit (so the replay of each final sample is the one the frame shaded).
* The adjoint of L = sum(g * I) matches central finite differences of the
  model on the materials with the largest gradients.
* Euler's identity for the bilinear model: sum(d * dL/dd) = sum(g * (I(d, e)
  - I(0, e))) and sum(e * dL/de) = sum(g * (I(d, e) - I(d, 0))), over all
  materials at once.
"""
import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import oracle_scene


def _workload(indirect):
    w = ctamd.hd_di(width=64, height=48, grid=64, chunk=32, frames=3, indirect=indirect, max_depth=2)
    w.emissive *= 4
    return w


def _materials(sc):
    mats = sc.materials
    d = np.array([m.diffuse[:] for m in mats], np.float32)
    e = np.array([m.emission[:] for m in mats], np.float32)
    return d, e


@pytest.mark.parametrize("indirect", [False, True], ids=["di", "di_pt"])
def test_frame_adjoint_model_and_gradients(indirect):
    w = _workload(indirect)
    sc = ctamd.Scene(ctamd.scene_desc(w))
    ofr = O.OFrame(oracle_scene(sc), w.width, w.height)
    for f in range(w.frames):
        p = ctamd.frame_params(w, f)
        ofr.render(p)
    d0, e0 = _materials(sc)
    img = ofr.read(ctamd.CT_BUF_TAA_FRAME)  # the shading output TAA blends
    assert np.array_equal(ofr.reshade(p, d0, e0).view(np.uint32), img.view(np.uint32))
    rng = np.random.default_rng(5)
    g = rng.standard_normal((w.width * w.height, 4)).astype(np.float32)
    g[:, 3] = 0.0
    gd, ge = ofr.adjoint(p, g, len(d0))
    assert np.count_nonzero(gd) > 100 and np.count_nonzero(ge) > 10

    def loss(d, e):
        return float((g.astype(np.float64) * ofr.reshade(p, d, e).astype(np.float64)).sum())

    # Euler's identity, every material at once
    lhs_d = float((d0.astype(np.float64) * gd).sum())
    rhs_d = loss(d0, e0) - loss(np.zeros_like(d0), e0)
    assert abs(lhs_d - rhs_d) <= 1e-4 * max(abs(rhs_d), 1.0), (lhs_d, rhs_d)
    lhs_e = float((e0.astype(np.float64) * ge).sum())
    rhs_e = loss(d0, e0) - loss(d0, np.zeros_like(e0))
    assert abs(lhs_e - rhs_e) <= 1e-4 * max(abs(rhs_e), 1.0), (lhs_e, rhs_e)
    # central differences on the largest gradients
    top_d = np.argsort(-np.abs(gd).sum(1))[:12]
    top_e = np.argsort(-np.abs(ge).sum(1))[:8]
    for which, top, grad in (("d", top_d, gd), ("e", top_e, ge)):
        for m in top:
            for k in range(3):
                base = d0 if which == "d" else e0
                h = 1e-2 * max(abs(float(base[m, k])), 1e-2)
                lo, hi = base.copy(), base.copy()
                lo[m, k] -= h
                hi[m, k] += h
                fd = ((loss(hi, e0) - loss(lo, e0)) if which == "d" else (loss(d0, hi) - loss(d0, lo))) / (2 * h)
                assert abs(fd - grad[m, k]) <= 1e-3 * max(abs(fd), 1e-3), (which, m, k, fd, grad[m, k])
