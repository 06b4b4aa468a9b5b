"""Adjoint kernels on the GPU: forward image bit for bit against the
oracle, gradients against the oracle's double-precision accumulation and
against central finite differences of the L2 loss (eps 1e-3, 64 voxels),
through the torch.autograd binding."""
import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd import adjoint as AD
from helpers import bits

pytestmark = pytest.mark.gpu


def _gpu_setup(w):
    sc, alb, emi = AD.build_scene(w)
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024,
                      max_cube_lights=max(sc.info.cube_light_count, 1) + 8)
    AD._load_geometry(sc, acc)
    slots = AD.VoxelSlots(sc, acc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(w.views, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    return sc, acc, slots, views, alb, emi


@pytest.mark.parametrize("depth", [0, 3])
def test_adjoint_forward_and_gradients_match_oracle(gpu, depth):
    import torch
    w = AD.AdjointWorkload(grid=64, chunk=32, width=40, height=32, views=3, spp=3, max_depth=depth)
    sc, acc, slots, views, alb, emi = _gpu_setup(w)
    r = AD.DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    a_t = torch.from_numpy(slots.scatter(alb)).to(dev)
    e_t = torch.from_numpy(slots.scatter(emi)).to(dev)
    img = r.forward(a_t, e_t).cpu().numpy()
    osc = O.OScene.from_scene(sc)
    desc = ctamd.AdjointDesc(len(views), w.spp, w.max_depth, w.target_seed)
    oimg = osc.adjoint_render(desc, views, alb, emi)
    assert np.array_equal(bits(img), bits(oimg))
    rng = np.random.default_rng(3)
    g = rng.normal(size=oimg.shape).astype(np.float32)
    ga, ge = r.backward(a_t, e_t, torch.from_numpy(g).to(dev))
    oga, oge = osc.adjoint_backward(desc, views, alb, emi, g)
    ga = slots.gather(ga.cpu().numpy())
    ge = slots.gather(ge.cpu().numpy())
    assert np.allclose(ge, oge, rtol=1e-4, atol=1e-6 * np.abs(oge).max())
    assert np.allclose(ga, oga, rtol=1e-4, atol=1e-6 * max(np.abs(oga).max(), 1e-30))
    assert np.abs(oge).sum() > 0


def test_autograd_against_finite_differences(gpu):
    import torch
    w = AD.AdjointWorkload(grid=96, chunk=32, width=96, height=96, views=4, spp=4, max_depth=3)
    sc, acc, slots, views, alb, emi = _gpu_setup(w)
    r = AD.DiffRenderer(acc, views, w.spp, w.max_depth, w.target_seed)
    dev = torch.device("cuda")
    rng = np.random.default_rng(w.fd_seed)
    alb_t = np.clip(alb + rng.normal(0, 0.05, alb.shape), 0, 1).astype(np.float32)
    emi_t = (emi * rng.uniform(0.8, 1.2, emi.shape)).astype(np.float32)
    target = r.forward(torch.from_numpy(slots.scatter(alb_t)).to(dev),
                       torch.from_numpy(slots.scatter(emi_t)).to(dev)).double()
    a = torch.from_numpy(slots.scatter(alb)).to(dev).requires_grad_(True)
    e = torch.from_numpy(slots.scatter(emi)).to(dev).requires_grad_(True)
    loss = ((AD.render(r, a, e).double() - target) ** 2).mean()
    loss.backward()
    ga, ge = a.grad.cpu().numpy(), e.grad.cpu().numpy()

    def L(av, ev):
        with torch.no_grad():
            return float(((r.forward(av, ev).double() - target) ** 2).mean())

    eps = 1e-3
    # 64 voxels: the 32 largest emission and 32 largest albedo gradients
    checked = 0
    for kind, grad in (("e", ge), ("a", ga.reshape(-1))):
        for i in np.argsort(-np.abs(grad))[:32]:
            base_a, base_e = a.detach().clone(), e.detach().clone()
            arr = base_e if kind == "e" else base_a.view(-1)
            v0 = float(arr[i])
            arr[i] = v0 + eps
            lp = L(base_a, base_e)
            arr[i] = v0 - eps
            lm = L(base_a, base_e)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grad[i]) <= 1e-3 * abs(grad[i]) + 1e-12, (kind, int(i), fd, float(grad[i]))
            checked += 1
    assert checked == 64


def test_inverse_rendering_reduces_loss(gpu):
    w = AD.AdjointWorkload(grid=64, chunk=32, width=64, height=64, views=4, spp=4, max_depth=3)
    hist = AD.inverse_rendering_c4(steps=15, w=w)
    assert hist[-1] < 0.9 * hist[0], hist
