"""Pins of the oracle's material callables (rcall_scatter.glsl:11-88) and of
its plain path tracer: closed forms for reflection, Snell refraction, the
Schlick reflectance, metal absorption and the reference's ior convention;
an isotropic medium; a Lambertian-only render equal, bit for bit, to the
adjoint's forward render (itself pinned by finite differences)."""
import ctypes as C
import math

import numpy as np
import pytest

import ctamd
import oracle as O
from ctamd import adjoint as AD
from helpers import bits


def _mat(type_, diffuse=(0.5, 0.6, 0.7), specular=(0.9, 0.8, 0.7), rough=0.0, ior=1.5, emission=(0, 0, 0)):
    m = ctamd.Material()
    m.type = type_
    m.diffuse[:] = diffuse
    m.specular[:] = specular
    m.emission[:] = emission
    m.roughness, m.ior = rough, ior
    return m


def _scatter(m, d, n, seed=7):
    f3 = C.c_float * 3
    s = C.c_uint32(seed)
    sd, att, ok = f3(), f3(), C.c_int(0)
    O.lib().oc_scatter_material(C.byref(m), f3(*d), f3(*n), C.byref(s), sd, att, C.byref(ok))
    return np.array(sd[:], np.float64), np.array(att[:], np.float32), bool(ok.value), s.value


def test_metal_mirror_and_absorption():
    r2 = 1 / math.sqrt(2)
    sd, att, ok, _ = _scatter(_mat(ctamd.CT_MATERIAL_METAL), (r2, -r2, 0.0), (0.0, 1.0, 0.0))
    assert ok and np.allclose(sd, (r2, r2, 0.0), atol=1e-7)
    assert np.array_equal(att, np.array([0.9, 0.8, 0.7], np.float32))  # specular
    # a grazing ray reflects into the surface plane: dot(scatter, n) = 0 absorbs
    _, _, ok, _ = _scatter(_mat(ctamd.CT_MATERIAL_METAL), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert not ok
    # roughness 1: a cosine lobe around the mirror direction
    sd, _, _, _ = _scatter(_mat(ctamd.CT_MATERIAL_METAL, rough=1.0), (r2, -r2, 0.0), (0.0, 1.0, 0.0))
    assert abs(np.linalg.norm(sd) - 1) < 1e-6 and sd @ np.array([r2, r2, 0.0]) >= -1e-6


def test_refraction_follows_snell():
    f3 = C.c_float * 3
    rng = np.random.default_rng(2)
    for _ in range(50):
        th = rng.uniform(0.0, 1.3)
        eta = rng.uniform(0.5, 0.95)
        d = np.array([math.sin(th), -math.cos(th), 0.0])
        out = f3()
        O.lib().oc_refraction(f3(*d), f3(0.0, 1.0, 0.0), eta, out)
        t = np.array(out[:], np.float64)
        assert abs(np.linalg.norm(t) - 1) < 2e-6
        assert abs(t[0] - eta * math.sin(th)) < 2e-6 and t[1] < 0  # sin(theta_t) = eta sin(theta_i)
    for eta in (0.5, 1.0, 1.5, 2.4):
        r0 = ((1 - eta) / (1 + eta)) ** 2
        assert abs(O.lib().oc_reflectance(1.0, eta) - r0) < 1e-7
        assert O.lib().oc_reflectance(0.0, eta) == pytest.approx(1.0, abs=1e-7)


def test_dielectric_ior_convention_and_total_reflection():
    # entering (dir . n < 0) uses eta = ior (rcall_scatter.glsl:29-34): with
    # ior 1.5, sin(theta) > 2/3 cannot refract and reflects
    th = math.asin(0.8)
    d = (math.sin(th), -math.cos(th), 0.0)
    for seed in range(20):
        sd, att, ok, _ = _scatter(_mat(ctamd.CT_MATERIAL_DIELECTRIC), d, (0.0, 1.0, 0.0), seed)
        assert ok and np.allclose(sd, (d[0], -d[1], 0.0), atol=1e-6) and np.array_equal(att, np.ones(3, np.float32))
    # leaving (dir . n > 0) uses 1/ior and refracts for most draws
    th = math.asin(0.3)
    d = (math.sin(th), math.cos(th), 0.0)
    refr = 0
    for seed in range(50):
        sd, _, ok, _ = _scatter(_mat(ctamd.CT_MATERIAL_DIELECTRIC), d, (0.0, 1.0, 0.0), seed)
        if sd[1] > 0:
            refr += 1
            assert abs(sd[0] - 0.3 / 1.5) < 2e-6
    assert refr > 40


def test_medium_is_isotropic_and_lambertian_is_cosine():
    dirs = []
    seed = 11
    for _ in range(4000):
        sd, att, ok, seed = _scatter(_mat(ctamd.CT_MATERIAL_CONSTANT_MEDIUM), (0, 0, -1.0), (0, 0, 1.0), seed)
        dirs.append(sd)
    dirs = np.array(dirs)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-6)
    assert np.abs(dirs.mean(0)).max() < 0.05 and abs((dirs[:, 2] > 0).mean() - 0.5) < 0.05
    f3 = C.c_float * 3
    for seed in range(20):
        sd, att, ok, _ = _scatter(_mat(ctamd.CT_MATERIAL_LAMBERTIAN), (0, 0, -1.0), (0, 0, 1.0), seed)
        s = C.c_uint32(seed)
        ref = f3()
        O.lib().oc_random_cosine_direction(C.byref(s), f3(0, 0, 1.0), ref)
        assert np.array_equal(np.array(sd, np.float32), np.array(ref[:], np.float32))
        assert np.array_equal(att, np.array([0.5, 0.6, 0.7], np.float32))


def test_lambertian_render_equals_the_adjoint_forward():
    w = AD.AdjointWorkload(grid=32, chunk=32, width=24, height=16, views=2, spp=2, max_depth=3)
    sc, alb, emi = AD.build_scene(w)
    osc = O.OScene.from_scene(sc)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(w.views, w.width, w.height, (0, 0, 0), 1.6 * extent, 0.4 * extent)
    desc = ctamd.AdjointDesc(len(views), w.spp, w.max_depth, w.target_seed)
    # one material per voxel carrying its albedo and (grey) emission
    mats = []
    for v in range(sc.info.primitive_count):
        mats.append(_mat(ctamd.CT_MATERIAL_LAMBERTIAN, diffuse=tuple(alb[v].tolist()), emission=(emi[v],) * 3))
    prims = sc.primitives.copy()
    prims[:, 0] = np.arange(sc.info.primitive_count)
    osc2 = O.OScene(list(sc.instances), np.tile(np.eye(4, dtype=np.float32).reshape(16), (sc.info.instance_count, 1)),
                    sc.aabbs, prims, mats, list(sc.lights))
    a = osc.adjoint_render(desc, views, alb, emi)
    p = osc2.path_trace(desc, views, mats)
    assert np.array_equal(bits(np.repeat(a[:, :1], 3, 1) if a.shape[1] == 1 else a), bits(p))
    assert p.max() > 0
