"""Every workload at its full size on the GPU:

* the 64^3 plain workload at 256x256, four frames, in both parameter modes,
  bit for bit against the oracle;
* the 1080p workload (256^3 in 64 BLAS of 64^3, ReSTIR DI + 3-bounce path,
  TAA 0.1, orbiting camera) rendered as the library renders it by default
  (two interleaved halves on two streams), three frames, every buffer bit
  for bit against the oracle; then two frames with the RIS BRDF strategy
  on, and the frame adjoint of the last one against the oracle's;
* the batched workload (64 views over the 64^3, 128^3, 256^3 and 384^3
  scenes, DI + 2 bounces): the graph-replayed batch equals each view
  rendered on its own, and one view of each scene equals the oracle;
* the inverse-rendering workload (128^3, 8 views of 512x512, 16 spp, 3
  bounces): the path-replay gradient against central finite differences
  (eps 1e-3 of the parameter's magnitude) on 64 voxels.
The 512^3 palette workload has its own file (test_gpu_palette.py).
"""
import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd import adjoint as AD
from helpers import all_buffers_equal, bits, di_array, gpu_scene, oracle_scene, pres_array

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("reference_mode", [False, True], ids=["workload", "reference"])
def test_plain_256_four_frames_match_oracle(gpu, reference_mode):
    w = ctamd.small_plain(reference_mode=reference_mode, width=256, height=256, frames=4)
    sc, acc, ctx = gpu_scene(w)
    assert sc.info.primitive_count == round(0.12 * 64 ** 3) and sc.info.cube_light_count == 8
    ofr = O.OFrame(oracle_scene(sc, acc), w.width, w.height)
    for f in range(4):
        _say("plain 256x256 frame", f)
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        all_buffers_equal(ctx, ofr, w, f)


def test_hd_1080p_frames_match_oracle(gpu):
    w = ctamd.hd_di()
    assert (w.width, w.height, w.grid, w.chunk) == (1920, 1080, (256,) * 3, (64,) * 3)
    _say("1080p: building the scene")
    sc, acc, ctx = gpu_scene(w)
    assert sc.info.instance_count == 64 and sc.info.primitive_count == 4194304
    _say("1080p: oracle LBVH")
    ofr = O.OFrame(oracle_scene(sc, acc), w.width, w.height)
    for f in range(3):
        _say("1080p frame", f)
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        hit = all_buffers_equal(ctx, ofr, w, f)
        assert hit.mean() > 0.3
    # two more frames with the RIS BRDF strategy on (temporal history carried
    # over), then the frame adjoint of the last one against the oracle's
    w.features |= ctamd.CT_RIS_BRDF_BIT
    for f in range(3, 5):
        _say("1080p frame", f, "with the RIS BRDF strategy")
        p = ctamd.frame_params(w, f)
        ctx.render(acc, p)
        ofr.render(p)
        all_buffers_equal(ctx, ofr, w, f)
    import torch
    nm = sc.info.material_count
    g = np.random.default_rng(2).standard_normal((w.width * w.height, 4)).astype(np.float32)
    _say("1080p frame adjoint")
    gd, ge = AD.frame_adjoint(ctx, acc, p, torch.from_numpy(g).cuda(), nm)
    torch.cuda.synchronize()
    od, oe = ofr.adjoint(p, g, nm)
    for a, b in ((gd.cpu().numpy(), od), (ge.cpu().numpy(), oe)):
        assert np.count_nonzero(b) > 10000
        assert np.all(np.abs(a - b) <= 1e-5 * np.abs(b) + 1e-6 * np.abs(b).max())


def test_batched_64_views_equal_per_view_and_oracle(gpu):
    scenes, views = ctamd.batched_views()
    assert len(views) == 64 and sorted({k for k, _ in views}) == [0, 1, 2, 3]
    built = []
    for s in scenes:
        _say("batched: scene", s.name)
        sc = ctamd.Scene(ctamd.scene_desc(s))
        acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                          max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                          max_cube_lights=sc.info.cube_light_count + 8)
        ctamd.load_scene(sc, acc, ctamd.Ctx(8, 8))
        built.append((sc, acc))

    def ctx_for(k, w):
        c = ctamd.Ctx(w.width, w.height)
        ctamd.bind_scene(built[k][0], c)
        return c

    plain = [ctx_for(k, w) for k, w in views]
    batch_ctx = [ctx_for(k, w) for k, w in views]
    b = ctamd.Batch(batch_ctx, [built[k][1] for k, _ in views])
    # one view of each scene against the oracle
    picks = [next(i for i, (k, _) in enumerate(views) if k == s) for s in range(4)]
    oracle = {}
    for i in picks:
        k, w = views[i]
        _say("batched: oracle scene for view", i, "scene", k)
        oracle[i] = O.OFrame(oracle_scene(built[k][0], built[k][1]), w.width, w.height)
    for f in range(2):
        _say("batched frame", f)
        params = [ctamd.frame_params(w, f) for _, w in views]
        b.render(params, use_graph=True)
        for (k, w), c, p in zip(views, plain, params):
            c.render(built[k][1], p)
        for v, ((k, w), pc, bc) in enumerate(zip(views, plain, batch_ctx)):
            for which in (ctamd.CT_BUF_IMAGE, ctamd.CT_BUF_TAA_FRAME, ctamd.CT_BUF_INDIRECT):
                assert np.array_equal(bits(pc.read(which, w.width, w.height)),
                                      bits(bc.read(which, w.width, w.height))), (f, v, which)
            for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
                assert np.array_equal(di_array(pc.read(which, w.width, w.height)),
                                      di_array(bc.read(which, w.width, w.height))), (f, v, which)
            assert np.array_equal(pres_array(pc.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR, w.width, w.height)),
                                  pres_array(bc.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR, w.width, w.height))), (f, v)
        for i in picks:
            _say("batched: oracle view", i, "frame", f)
            k, w = views[i]
            oracle[i].render(params[i])
            bc = batch_ctx[i]
            for which in (ctamd.CT_BUF_DI, ctamd.CT_BUF_PREV_DI):
                assert np.array_equal(di_array(bc.read(which, w.width, w.height)), di_array(oracle[i].read(which))), \
                    (f, i, which)
            for which in (ctamd.CT_BUF_PATH_RESERVOIR, ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR):
                assert np.array_equal(pres_array(bc.read(which, w.width, w.height)),
                                      pres_array(oracle[i].read(which))), (f, i, which)
            assert np.array_equal(bits(bc.read(ctamd.CT_BUF_IMAGE, w.width, w.height)),
                                  bits(oracle[i].read(ctamd.CT_BUF_IMAGE))), (f, i)


def test_inverse_rendering_workload_gradients_match_finite_differences(gpu):
    import torch
    w = AD.AdjointWorkload()
    assert (w.grid, w.views, w.width, w.height, w.spp, w.max_depth) == (128, 8, 512, 512, 16, 3)
    _say("adjoint: building the 128^3 scene")
    sc, alb, emi = AD.build_scene(w)
    assert 450_000 < sc.info.primitive_count < 550_000
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=8)
    AD._load_geometry(sc, acc)
    slots = AD.VoxelSlots(sc, acc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(w.views, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    r = AD.DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    a0, e0 = AD.initial_parameters(w, alb, emi)
    target = r.forward(torch.from_numpy(slots.scatter(alb)).to(dev), torch.from_numpy(slots.scatter(emi)).to(dev))
    target = target.double()
    a = torch.from_numpy(slots.scatter(a0)).to(dev).requires_grad_(True)
    e = torch.from_numpy(slots.scatter(e0)).to(dev).requires_grad_(True)
    loss = ((AD.render(r, a, e).double() - target) ** 2).mean()
    loss.backward()
    ga, ge = a.grad.cpu().numpy(), e.grad.cpu().numpy()

    def L(av, ev):
        with torch.no_grad():
            return float(((r.forward(av, ev).double() - target) ** 2).mean())

    eps = 1e-3
    rng = np.random.default_rng(w.fd_seed)
    seen = np.nonzero(np.abs(ga).sum(axis=1) > 0)[0]
    # 64 voxels: the 16 largest emission and 16 largest albedo gradients, and
    # 32 random voxels the views see (seed 0xC6)
    checks = [("e", int(i)) for i in np.argsort(-np.abs(ge))[:16]]
    checks += [("a", int(i)) for i in np.argsort(-np.abs(ga.reshape(-1)))[:16]]
    checks += [("a", int(3 * s + rng.integers(0, 3))) for s in rng.choice(seen, 32, replace=False)]
    gmax = max(np.abs(ga).max(), np.abs(ge).max())
    for n, (kind, i) in enumerate(checks):
        if n % 16 == 0:
            _say("adjoint: finite differences", n)
        base_a, base_e = a.detach().clone(), e.detach().clone()
        arr = base_e if kind == "e" else base_a.view(-1)
        grad = ge[i] if kind == "e" else ga.reshape(-1)[i]
        v0 = float(arr[i])
        h = eps * max(1.0, abs(v0))  # relative to the parameter: emission runs to 20
        arr[i] = v0 + h
        lp = L(base_a, base_e)
        arr[i] = v0 - h
        lm = L(base_a, base_e)
        fd = (lp - lm) / (2 * h)
        # 1e-3 relative; voxels whose gradient is near zero are held to the
        # central difference's own resolution on this loss
        assert abs(fd - grad) <= 1e-3 * abs(grad) + 2e-6 * gmax, (kind, i, fd, float(grad))
    assert len(checks) == 64


def test_inverse_rendering_workload_200_adam_steps(gpu):
    """The 200-step fit of the inverse-rendering workload: the loss falls at
    every one of the first 10 steps (learning rates ramped over them) and
    ends below 1 % of where it starts."""
    _say("adjoint: 200 Adam steps")
    hist = AD.inverse_rendering_c4(steps=200, log=_say)
    assert len(hist) == 200
    assert all(hist[i + 1] < hist[i] for i in range(10)), hist[:11]
    assert hist[-1] < 0.01 * hist[0], (hist[0], hist[-1])
