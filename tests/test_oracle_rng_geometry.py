"""Pins of the oracle's RNG and geometry against published values, closed
forms and brute force (CPU only)."""
import math

import numpy as np
import pytest

import oracle as O

M32 = 0xFFFFFFFF
TEA_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)  # random.glsl:35-36


def tea_ref(v0, v1, key, rounds):
    """TEA (Wheeler & Needham 1994) with a variable round count."""
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) & M32) + key[0]) ^ (v1 + s) ^ ((v1 >> 5) + key[1]))) & M32
        v1 = (v1 + ((((v0 << 4) & M32) + key[2]) ^ (v0 + s) ^ ((v0 >> 5) + key[3]))) & M32
    return v0, v1


def test_tea_reference_matches_published_vector():
    # Published TEA known answer: zero key, zero plaintext, 32 rounds.
    assert tea_ref(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_tea16_matches_reference():
    rng = np.random.default_rng(1)
    for v0, v1 in rng.integers(0, 2 ** 32, size=(200, 2), dtype=np.uint64):
        assert O.tea(int(v0), int(v1)) == tea_ref(int(v0), int(v1), TEA_KEY, 16)[0]


def test_lcg_numerical_recipes_sequence():
    # Numerical Recipes quick generator from idum = 0.
    expect = [1013904223, 1196435762, 3519870697, 2868466484]
    s = 0
    for e in expect:
        r, s = O.lcg(s)
        assert s == e and r == (e & 0x00FFFFFF)


def test_rnd_is_24bit_fraction():
    s = 12345
    for _ in range(100):
        s_prev = s
        r, s = O.rnd(s)
        assert r == np.float32(((1664525 * s_prev + 1013904223) & M32 & 0xFFFFFF) / 16777216.0)
        assert 0.0 <= r < 1.0


def test_urnd_interval_equal_bounds_consumes_nothing():
    v, s = O.urnd_interval(99, 3, 3)
    assert v == 3 and s == 99
    v, s = O.urnd_interval(99, 0, 6)
    r, s2 = O.rnd(99)
    assert v == int(np.float32(r) * np.float32(6.0)) and s == s2


def test_intersect_box_axis_ray():
    # Unit-voxel box at the origin, ray along +x from x = -1: enters at x = -h.
    h = 0.015625
    hit, t, n = O.intersect_box((0, 0, 0), h, (-1.0, 0.0, 0.0), (1.0, 0.0, 0.0), True, True)
    assert hit and t == np.float32(1.0 - h) and tuple(n) == (-1.0, 0.0, 0.0)
    # Miss: offset beyond the half extent.
    hit, _, _ = O.intersect_box((0, 0, 0), h, (-1.0, 0.02, 0.0), (1.0, 0.0, 0.0), True, True)
    assert not hit
    # Pointing away.
    hit, _, _ = O.intersect_box((0, 0, 0), h, (-1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), True, True)
    assert not hit


def test_intersect_box_inside_origin_exit_face():
    h = 0.015625
    hit, t, n = O.intersect_box((0, 0, 0), h, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), True, True)
    assert hit and t == np.float32(h)          # winding -1: leaving through +y
    hit, t, _ = O.intersect_box((0, 0, 0), h, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), False, False)
    assert not hit                              # rayCanStartInBox = false: front faces are behind


def test_intersect_box_edge_is_strict():
    # A ray through the shared edge of the faces misses (strict < in TEST, primitives.glsl:317-319).
    h = 0.015625
    hit, _, _ = O.intersect_box((0, 0, 0), h, (-1.0, h, 0.0), (1.0, 0.0, 0.0), True, True)
    assert not hit


def test_intersect_box_random_against_slab_interval():
    """Closed form: for an outside origin the hit distance is the entry of
    the slab interval and the hit point lies on the box surface."""
    rng = np.random.default_rng(3)
    h = np.float32(0.015625)
    n_hit = 0
    for _ in range(2000):
        o = (rng.random(3) * 0.2 - 0.1).astype(np.float32)
        if np.all(np.abs(o) < h):
            continue
        d = (rng.random(3) * 0.05 - 0.025) - o  # aim near the box
        d = (d / np.linalg.norm(d)).astype(np.float32)
        hit, t, n = O.intersect_box((0, 0, 0), float(h), o, d, True, True)
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (-h - o.astype(np.float64)) / d
            t2 = (h - o.astype(np.float64)) / d
        tn = np.max(np.minimum(t1, t2))
        tf = np.min(np.maximum(t1, t2))
        expect = tn < tf and tf > 0
        if abs(tf - tn) < 1e-5 or abs(tn) < 1e-5:
            continue  # grazing: float rounding may decide either way
        assert hit == expect
        if hit:
            n_hit += 1
            assert abs(t - tn) < 1e-5
            p = o + d * t
            assert abs(np.max(np.abs(p)) - h) < 1e-5
            assert n[np.argmax(np.abs(n))] == -np.sign(d[np.argmax(np.abs(n))])
    assert n_hit > 200


def test_compute_ray_origin():
    # Far from the origin: +/-256 ulps along the normal.
    p = np.float32(1.5)
    out = O.compute_ray_origin((p, 0.5, -0.75), (1.0, 0.0, 0.0))
    assert out.view(np.uint32)[0] == np.array([p], np.float32).view(np.uint32)[0] + 256
    assert out[1] == np.float32(0.5) and out[2] == np.float32(-0.75)
    out = O.compute_ray_origin((-1.5, 0.0, 0.0), (-1.0, 0.0, 0.0))
    assert out.view(np.uint32)[0] == np.array([-1.5], np.float32).view(np.uint32)[0] + 256  # moves away from 0
    # Near the origin: + n / 65536.
    out = O.compute_ray_origin((0.01, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert out[1] == np.float32(1.0 / 65536.0) and out[0] == np.float32(0.01)


def test_cube_like_normal():
    assert tuple(O.cube_like_normal((0.1, -0.9, 0.3))) == (0.0, -1.0, 0.0)
    assert tuple(O.cube_like_normal((0.5, 0.5, 0.1))) == (1.0, 0.0, 0.0)  # tie -> x first
    assert tuple(O.cube_like_normal((0.0, 0.0, -2.0))) == (0.0, 0.0, -1.0)


@pytest.mark.parametrize("dy", [-1.0, -0.3, 0.0, 0.6, 1.0])
def test_sky_daylight_closed_form(dy):
    # time = 1, afternoon: mix(sunset, day, 1) = day gradient (mat.glsl:111-116, 174-176).
    d = np.array([0.3, dy, -0.8], np.float64)
    t = 0.5 * (d[1] / np.linalg.norm(d) + 1.0)
    expect = np.array([1.0, 1.0, 1.0]) * (1 - t) + np.array([0.5, 0.7, 1.0]) * t
    assert np.allclose(O.sky_color(1.0, True, d), expect, atol=1e-6)


def test_morton_interleave():
    assert O.morton3(1, 0, 0) == 0b100 and O.morton3(0, 1, 0) == 0b010 and O.morton3(0, 0, 1) == 0b001
    assert O.morton3(1023, 1023, 1023) == (1 << 30) - 1
    rng = np.random.default_rng(5)
    for x, y, z in rng.integers(0, 1024, size=(100, 3)):
        m = 0
        for b in range(10):
            m |= ((int(x) >> b) & 1) << (3 * b + 2) | ((int(y) >> b) & 1) << (3 * b + 1) | ((int(z) >> b) & 1) << (3 * b)
        assert O.morton3(int(x), int(y), int(z)) == m
