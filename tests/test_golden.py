"""The oracle reproduces the committed golden fixtures bit for bit; on the
GPU the HIP path reproduces the frame fixtures."""
import os

import numpy as np
import pytest

import ctamd
import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_rng_fixture():
    k = load("rng_kat.npz")
    assert [O.tea(int(a), int(b)) for a, b in k["pairs"]] == k["tea"].tolist()
    s = int(k["seed"])
    for st, r in zip(k["lcg_state"], k["rnd"]):
        rr, s = O.rnd(s)
        assert s == st and np.float32(rr) == r


def test_lbvh_fixture():
    k = load("lbvh_300.npz")
    r = np.float32(0.015625 + 0.000244140625)
    c = k["centers"]
    b = O.lbvh_build(c, c - r, c + r)
    for name in ("keys", "ids", "child", "parent"):
        assert np.array_equal(b[name], k[name]), name
    assert np.array_equal(b["boxes"].view(np.uint32), k["boxes"].view(np.uint32))
    assert b["depth"] == int(k["depth"])


def _frames_oracle(w):
    from golden.make_golden import frames
    return frames(w)


@pytest.mark.parametrize("variant", ["", "ref_"])
def test_frames_fixture_oracle(variant):
    import sys
    sys.path.insert(0, os.path.dirname(G))
    k = load("frames_32.npz")
    w = ctamd.small_plain(width=32, height=32, frames=4) if not variant else \
        ctamd.small_plain(reference_mode=True, width=32, height=32, frames=4, taa=True)
    img, di, res = _frames_oracle(w)
    assert np.array_equal(img.view(np.uint32), k[variant + "img"].view(np.uint32))
    assert np.array_equal(di, k[variant + "di"])


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["", "ref_"])
def test_frames_fixture_gpu(gpu, variant):
    from helpers import gpu_scene
    k = load("frames_32.npz")
    w = ctamd.small_plain(width=32, height=32, frames=4) if not variant else \
        ctamd.small_plain(reference_mode=True, width=32, height=32, frames=4, taa=True)
    sc, acc, ctx = gpu_scene(w)
    for f in range(4):
        ctx.render(acc, ctamd.frame_params(w, f))
        img = ctx.read(ctamd.CT_BUF_IMAGE, 32, 32)
        di = np.frombuffer(bytes(ctx.read(ctamd.CT_BUF_DI, 32, 32)), np.uint32).reshape(-1, 15)
        assert np.array_equal(img.view(np.uint32), k[variant + "img"][f].view(np.uint32)), f
        assert np.array_equal(di, k[variant + "di"][f]), f


def test_pt_fixture_oracle():
    import sys
    sys.path.insert(0, os.path.dirname(G))
    from golden.make_golden import pt_frames, pt_workload
    k = load("pt_32.npz")
    out = pt_frames(pt_workload())
    for name in ("img", "ind"):
        assert np.array_equal(out[name].view(np.uint32), k[name].view(np.uint32)), name
    for name in ("pres", "tres"):
        assert np.array_equal(out[name], k[name]), name


@pytest.mark.gpu
def test_pt_fixture_gpu(gpu):
    import sys
    sys.path.insert(0, os.path.dirname(G))
    from golden.make_golden import pt_workload
    from helpers import gpu_scene
    k = load("pt_32.npz")
    w = pt_workload()
    sc, acc, ctx = gpu_scene(w)
    for f in range(w.frames):
        ctx.render(acc, ctamd.frame_params(w, f))
        img = ctx.read(ctamd.CT_BUF_IMAGE, 32, 32)
        ind = ctx.read(ctamd.CT_BUF_INDIRECT, 32, 32)
        pres = np.frombuffer(bytes(ctx.read(ctamd.CT_BUF_PATH_RESERVOIR, 32, 32)), np.uint32).reshape(-1, 23)
        tres = np.frombuffer(bytes(ctx.read(ctamd.CT_BUF_TEMPORAL_PATH_RESERVOIR, 32, 32)), np.uint32).reshape(-1, 23)
        assert np.array_equal(img.view(np.uint32), k["img"][f].view(np.uint32)), f
        assert np.array_equal(ind.view(np.uint32), k["ind"][f].view(np.uint32)), f
        assert np.array_equal(pres, k["pres"][f]), f
        assert np.array_equal(tres, k["tres"][f]), f
