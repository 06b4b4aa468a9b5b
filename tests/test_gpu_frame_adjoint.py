"""ct_frame_adjoint on the GPU: the gradient of a rendered ReSTIR frame's
shading output with its final reservoirs held fixed, with respect to every
material's diffuse albedo and emission.

* The GPU frames equal the oracle's bit for bit (so both sides hold the same
  checkpoint), and the GPU gradients equal the oracle adjoint's up to the
  order of the float atomics (relative 1e-5).
* Euler's identity for the bilinear model, evaluated with the oracle's
  re-shading: sum(d * dL/dd) = L(d, e) - L(0, e) and sum(e * dL/de) =
  L(d, e) - L(d, 0) for L = sum(g * image).
* Without a frame rendered on the accel's current build the call is
  refused (CT_E_STATE), and parameters of another size are rejected.
* After a batch render each view's context holds the same checkpoint as
  after rendering the view alone.
"""
import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd.adjoint import frame_adjoint
from helpers import all_buffers_equal, gpu_scene, oracle_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _workload(indirect):
    w = ctamd.hd_di(width=96, height=64, grid=96, chunk=32, frames=3, indirect=indirect, max_depth=2)
    w.emissive *= 4
    return w


@pytest.mark.parametrize("indirect", [False, True], ids=["di", "di_pt"])
def test_frame_adjoint_matches_oracle(gpu, indirect):
    import torch
    w = _workload(indirect)
    sc, acc, ctx = gpu_scene(w)
    ofr = O.OFrame(oracle_scene(sc, acc), w.width, w.height)
    for f in range(w.frames):
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        all_buffers_equal(ctx, ofr, w, f)
    nm = sc.info.material_count
    rng = np.random.default_rng(9)
    g = rng.standard_normal((w.width * w.height, 4)).astype(np.float32)
    _say("frame adjoint", w.name, "materials", nm)
    gd, ge = frame_adjoint(ctx, acc, p, torch.from_numpy(g).cuda(), nm)
    torch.cuda.synchronize()
    gd, ge = gd.cpu().numpy(), ge.cpu().numpy()
    od, oe = ofr.adjoint(p, g, nm)
    for a, b in ((gd, od), (ge, oe)):
        assert np.count_nonzero(b) > 50
        tol = 1e-5 * np.abs(b) + 1e-6 * np.abs(b).max()
        assert np.all(np.abs(a - b) <= tol), np.abs(a - b).max()
    mats = sc.materials
    d0 = np.array([m.diffuse[:] for m in mats], np.float32)
    e0 = np.array([m.emission[:] for m in mats], np.float32)

    def loss(d, e):
        return float((g.astype(np.float64) * ofr.reshade(p, d, e).astype(np.float64)).sum())

    lhs = float((d0.astype(np.float64) * gd).sum())
    rhs = loss(d0, e0) - loss(np.zeros_like(d0), e0)
    assert abs(lhs - rhs) <= 1e-4 * max(abs(rhs), 1.0), (lhs, rhs)
    lhs = float((e0.astype(np.float64) * ge).sum())
    rhs = loss(d0, e0) - loss(d0, np.zeros_like(e0))
    assert abs(lhs - rhs) <= 1e-4 * max(abs(rhs), 1.0), (lhs, rhs)


def test_frame_adjoint_refusals(gpu):
    import torch
    w = _workload(False)
    sc, acc, ctx = gpu_scene(w)
    nm = sc.info.material_count
    g = torch.zeros((w.width * w.height, 4), dtype=torch.float32, device="cuda")
    p = ctamd.frame_params(w, 0)
    with pytest.raises(ctamd.CtError) as ei:  # nothing rendered yet
        frame_adjoint(ctx, acc, p, g, nm)
    assert ei.value.code == ctamd.CT_E_STATE
    ctx.render(acc, p)
    small = ctamd.frame_params(ctamd.hd_di(width=32, height=32, grid=96, chunk=32), 0)
    with pytest.raises(ctamd.CtError) as ei:
        frame_adjoint(ctx, acc, small, torch.zeros((32 * 32, 4), dtype=torch.float32, device="cuda"), nm)
    assert ei.value.code == ctamd.CT_E_INVALID
    gd, ge = frame_adjoint(ctx, acc, p, g, nm)
    torch.cuda.synchronize()
    assert float(gd.abs().sum()) == 0.0 and float(ge.abs().sum()) == 0.0


def test_frame_adjoint_after_a_batch_render(gpu):
    """A context rendered as part of a batch holds the same checkpoint as one
    that rendered its frame alone: the frame adjoint of each view after two
    batched frames equals the adjoint after the same frames rendered view by
    view (float atomics: relative 1e-5)."""
    import torch
    scenes, views = ctamd.batched_views(n_views=3, seed=0xC5, grids=(64,), max_depth=2, size_scale=0.1)
    sw = scenes[0]
    sc = ctamd.Scene(ctamd.scene_desc(sw))
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=sc.info.cube_light_count + 8)
    ctamd.load_scene(sc, acc, ctamd.Ctx(8, 8))

    def ctx_for(w):
        c = ctamd.Ctx(w.width, w.height)
        ctamd.bind_scene(sc, c)
        return c

    plain = [ctx_for(w) for _, w in views]
    batched = [ctx_for(w) for _, w in views]
    b = ctamd.Batch(batched, [acc] * len(views))
    for f in range(2):
        params = [ctamd.frame_params(w, f) for _, w in views]
        b.render(params)
        b.stats()
        for ctx, p in zip(plain, params):
            ctx.render(acc, p)
            ctx.stats()
    nm = sc.info.material_count
    rng = np.random.default_rng(4)
    for (_, w), cp, cb, p in zip(views, plain, batched, params):
        g = torch.from_numpy(rng.standard_normal((w.width * w.height, 4)).astype(np.float32)).cuda()
        gp = [t.cpu().numpy() for t in frame_adjoint(cp, acc, p, g, nm)]
        gb = [t.cpu().numpy() for t in frame_adjoint(cb, acc, p, g, nm)]
        for a, c in zip(gp, gb):
            assert np.count_nonzero(a) > 0
            assert np.all(np.abs(a - c) <= 1e-5 * np.abs(a) + 1e-6 * np.abs(a).max())
