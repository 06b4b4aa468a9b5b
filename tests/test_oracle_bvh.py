"""LBVH arrays and BVH traversal of the oracle against invariants and
brute force over every primitive (CPU only)."""
import numpy as np
import pytest

import ctamd
import oracle as O
from helpers import random_rays, bits

LEAF = 0x80000000
R = np.float32(0.015625 + 0.000244140625)


def _lbvh_for(n, seed):
    rng = np.random.default_rng(seed)
    cells = rng.choice(32 ** 3, size=n, replace=False)
    c = np.stack([cells % 32, (cells // 32) % 32, cells // 1024], 1).astype(np.float32) * np.float32(0.03125)
    c = c + np.float32(0.015625)
    return c, c - R, c + R, O.lbvh_build(c, c - R, c + R)


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (17, 3), (500, 4), (4000, 5)])
def test_lbvh_invariants(n, seed):
    c, lo, hi, b = _lbvh_for(n, seed)
    keys, ids, child, boxes, parent = b["keys"], b["ids"], b["child"], b["boxes"], b["parent"]
    # sorted (key, id): stable sort of Morton keys
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert sorted(ids.tolist()) == list(range(n))
    for i in range(n - 1):
        if keys[i] == keys[i + 1]:
            assert ids[i] < ids[i + 1]
    if n == 1:
        assert b["depth"] == 0
        return
    # every leaf exactly once, every internal node except the root exactly once
    refs = child.reshape(-1, 2)
    leaves = sorted(int(r) & 0x7FFFFFFF for r in refs.ravel() if int(r) & LEAF)
    internals = sorted(int(r) for r in refs.ravel() if not int(r) & LEAF)
    assert leaves == list(range(n)) and internals == list(range(1, n - 1))
    # parent links
    for i, (a, bb) in enumerate(refs):
        for r in (int(a), int(bb)):
            assert parent[(n - 1) + (r & 0x7FFFFFFF) if r & LEAF else r] == i
    assert parent[0] == 0xFFFFFFFF
    # child boxes contain the leaf boxes below them; node box = union of children
    bx = boxes.reshape(-1, 2, 6)

    def box_of(r):
        r = int(r)
        if r & LEAF:
            k = ids[r & 0x7FFFFFFF]
            return np.concatenate([lo[k], hi[k]])
        u = bx[r]
        return np.concatenate([np.minimum(u[0, :3], u[1, :3]), np.maximum(u[0, 3:], u[1, 3:])])

    for i in range(n - 1):
        for s in range(2):
            assert np.array_equal(bx[i, s], box_of(refs[i, s]))
    # Karras split: the children of node i split its key range at the highest differing bit
    def span(r):
        r = int(r)
        if r & LEAF:
            return r & 0x7FFFFFFF, r & 0x7FFFFFFF
        a, bb = span(refs[r, 0]), span(refs[r, 1])
        return a[0], bb[1]

    for i in range(n - 1):
        (l0, l1), (r0, r1) = span(refs[i, 0]), span(refs[i, 1])
        assert l1 + 1 == r0
        k = [(int(keys[j]) << 32) | j for j in range(l0, r1 + 1)]
        top = (k[0] ^ k[-1]).bit_length()
        assert all(((x ^ k[0]) >> (top - 1)) == 0 for x in k[: l1 - l0 + 1])
        assert all(((x ^ k[0]) >> (top - 1)) == 1 for x in k[l1 - l0 + 1:])


def _scene(grid, chunk, occ, em, seed=0x5EED):
    w = ctamd.hd_di(width=16, height=16, grid=grid, chunk=chunk)
    d = ctamd.scene_desc(w)
    d.occupied_count = occ
    d.emissive_count = em
    d.seed = seed
    sc = ctamd.Scene(d)
    return sc, O.OScene.from_scene(sc)


@pytest.mark.parametrize("grid,chunk", [(24, 24), (32, 16)])
def test_closest_bvh_equals_brute_force(grid, chunk):
    sc, osc = _scene(grid, chunk, occ=grid ** 3 // 5, em=4)
    rng = np.random.default_rng(grid)
    o, d = random_rays(rng, 3000, sc.grid_min, sc.grid_max)
    t1, i1 = osc.trace_closest(o, d)
    t2, i2 = osc.trace_closest(o, d, brute=True)
    assert np.array_equal(i1, i2) and np.array_equal(bits(t1), bits(t2))
    assert 0.2 < (i1[:, 0] != 0xFFFFFFFF).mean() < 1.0


def test_any_bvh_equals_brute_force():
    sc, osc = _scene(32, 16, occ=32 ** 3 // 6, em=8)
    rng = np.random.default_rng(9)
    o, d = random_rays(rng, 3000, sc.grid_min, sc.grid_max, inside_frac=0.6)
    dist = (rng.random(o.shape[0]) * 1.5).astype(np.float32)
    tg = np.full((o.shape[0], 2), 0xFFFFFFFF, np.uint32)
    tg[::2, 0] = rng.integers(0, sc.info.instance_count, size=tg[::2].shape[0])
    tg[::2, 1] = 0
    a = osc.trace_any(o, d, dist, tg)
    b = osc.trace_any(o, d, dist, tg, brute=True)
    assert np.array_equal(a, b) and 0.05 < a.mean() < 0.95


def test_any_hit_excludes_only_the_target():
    sc, osc = _scene(16, 16, occ=16 ** 3 // 4, em=2)
    # A ray from outside the grid aimed through a face interior of each light
    # voxel (a ray through an edge or corner misses: strict < in the slab
    # test).  Excluding the target leaves exactly the other voxels.
    for L in sc.lights:
        c = np.array(L.position[:], np.float32)
        aim = c + np.array([0.004, -0.003, 0.002], np.float32)
        o = np.array([3.0, 0.1, 0.2], np.float32) + aim
        d = (aim - o) / np.linalg.norm(aim - o)
        o, d = o[None].astype(np.float32), d[None].astype(np.float32)
        dist = np.array([np.linalg.norm(aim - o[0]) + 0.05], np.float32)
        tgt = np.array([[L.instance_info.instance_id, L.instance_info.primitive_id]], np.uint32)
        t, ids = osc.trace_closest(o, d, brute=True)
        assert ids[0, 0] != 0xFFFFFFFF
        blocked = osc.trace_any(o, d, dist, None, brute=True)[0]
        free = osc.trace_any(o, d, dist, tgt, brute=True)[0]
        assert blocked == 1
        first_is_light = tuple(ids[0]) == tuple(tgt[0])
        if first_is_light:
            # nothing in front of the light: only voxels behind it could block
            assert free == osc.trace_any(o, d, np.array([t[0] - 1e-4], np.float32), None, brute=True)[0]
        else:
            assert free == 1
