"""The hybrid shift of the oracle's ReSTIR PT (shift.glsl:352-444 with the
retrace of path_tracer.glsl:889-966), pinned by what fixes it without the
reference running it (the reference never selects it, rgen.glsl:282):

* the hybrid path builder picks the first connectable vertex (far field,
  rough, after a rough one; path_tracer.glsl:474-514), so reconnection
  lengths above 1 follow near-field first segments, and a reservoir's own
  seed replayed from its own primary vertex is invertible: it reaches the
  vertex before its reconnection vertex;
* reconnection length 1 is the reconnection shift itself (bit for bit),
  with near-field connections rejected;
* for longer reconnection lengths the Jacobian is the float64 product of
  the geometry ratio at the reconnection vertex against the reservoir's
  cached one and the cached pdf ratios (shift.glsl:146-280 with
  use_cached_jacobian), from the replayed vertex alone;
* shifts whose replay is not invertible give zero integrand and Jacobian.
"""
import numpy as np
import pytest

import ctamd
import oracle as O

INVALID = 0xFFFFFFFF
NEAR = 0.0078125  # NEAR_FIELD_DISTANCE, defines.glsl:52


def _frame(shift):
    w = ctamd.hd_di(width=128, height=96, grid=96, chunk=32, frames=1, max_depth=4)
    w.restir.pt_shift_mapping = shift
    sc = ctamd.Scene(ctamd.scene_desc(w))
    osc = O.OScene.from_scene(sc)
    fr = O.OFrame(osc, w.width, w.height)
    p = ctamd.frame_params(w, 0)
    fr.render(p, threads=4)
    return sc, osc, p, fr.read(ctamd.CT_BUF_DI), fr.read(ctamd.CT_BUF_PATH_RESERVOIR)


@pytest.fixture(scope="module")
def hybrid():
    return _frame(ctamd.CT_SHIFT_HYBRID)


def _k(r):
    return (r.path_flags >> 4) & 0xF


def _plen(r):
    return r.path_flags & 0xF


def test_hybrid_reconnection_lengths_and_own_replay(hybrid):
    sc, osc, p, di, pres = hybrid
    ks = [_k(r) for r in pres if r.weight > 0]
    assert ks.count(1) > 100 and sum(k >= 2 for k in ks) >= 5
    # the reconnection-shift frame of the same scene only ever uses length 1
    _, _, _, _, pres_rc = _frame(ctamd.CT_SHIFT_RECONNECTION)
    assert {_k(r) for r in pres_rc if r.weight > 0} == {1}
    n = 0
    for px, r in enumerate(pres):
        if r.weight > 0 and _k(r) >= 2 and _k(r) <= _plen(r):
            rt = osc.probe_retrace(p, di[px], r)
            assert rt["ok"], px
            assert np.all(rt["tp"] > 0) and rt["ids"][0] != INVALID, px
            n += 1
    assert n >= 3


def test_hybrid_length_one_is_the_reconnection_shift(hybrid):
    sc, osc, p, di, pres = hybrid
    prc = ctamd.FrameParams.from_buffer_copy(bytes(p))
    prc.restir.pt_shift_mapping = ctamd.CT_SHIFT_RECONNECTION
    W = p.width
    same = near = 0
    for px, r in enumerate(pres):
        if not (r.weight > 0 and _k(r) == 1 and di[px].distance > 0):
            continue
        for nb in (px - 1, px + 1, px - W, px + W):
            if not (0 <= nb < len(di) and di[nb].distance > 0):
                continue
            jh, fh = osc.probe_shift(p, di[nb], di[px], r)
            jr, fr_ = osc.probe_shift(prc, di[nb], di[px], r)
            rc = np.array(r.rc_vertex_hit.hit[:], np.float64)
            if np.linalg.norm(rc - np.array(di[nb].position[:], np.float64)) > 2 * NEAR:
                assert jh == jr and np.array_equal(fh, fr_), (px, nb)
                same += 1
            elif jh == 0.0 and jr != 0.0:
                near += 1  # a near-field connection: rejected by the hybrid shift only
    assert same > 200


def _pdf(n, wo, wi):
    n = n / np.linalg.norm(n)
    a, b = np.dot(n, wo), np.dot(n, wi)
    return 0.0 if min(a, b) < 1e-6 else b / np.pi


def _pushed(p, n):
    return O.compute_ray_origin(O.compute_ray_origin(np.asarray(p, np.float32), n), n).astype(np.float64)


def _entry(origin, target, centre, half):
    d = target - origin
    lo, hi = centre - half, centre + half
    with np.errstate(divide="ignore", invalid="ignore"):
        t1, t2 = (lo - origin) / d, (hi - origin) / d
    tn = np.minimum(t1, t2)
    ax = int(np.nanargmax(tn))
    nrm = np.zeros(3)
    nrm[ax] = -np.sign(d[ax])
    return origin + tn[ax] * d, nrm


def test_hybrid_jacobian_from_the_replayed_vertex(hybrid):
    """Own and neighbour shifts of reservoirs with reconnection length >= 2:
    the Jacobian from the replayed vertex y (the one before the
    reconnection vertex x) in float64 is
    (|cos(n_x, x->y)| / |x - y|^2) / cached.z * pdf(y -> x) / cached.x
    * pdf_x(y side, wi) / cached.y; replays that are not invertible give 0."""
    sc, osc, p, di, pres = hybrid
    W = p.width
    checked = rejected = 0
    for px, r in enumerate(pres):
        if not (r.weight > 0 and _k(r) >= 2 and _k(r) <= _plen(r)):
            continue
        for nb in (px, px - 1, px + 1, px - W, px + W):
            if not (0 <= nb < len(di) and di[nb].distance > 0):
                continue
            jac, f = osc.probe_shift(p, di[nb], di[px], r)
            rt = osc.probe_retrace(p, di[nb], r)
            if not rt["ok"]:
                assert jac == 0.0 and not np.any(f), (px, nb)
                rejected += 1
                continue
            if jac == 0.0:
                continue  # not visible, near field or behind a surface
            y = _pushed(rt["pos"], rt["nrm"])
            h = r.rc_vertex_hit
            slot = sc.instances[h.object.instance_id].first_primitive_index + h.object.primitive_id
            box = np.asarray(sc.aabbs[slot], np.float64).reshape(6)
            x, nx = _entry(y, np.array(h.hit[:], np.float64), 0.5 * (box[:3] + box[3:]), 0.015625)
            x = _pushed(x, nx)
            to_x = (x - y) / np.linalg.norm(x - y)
            cj = np.array(r.cached_jacobian[:], np.float64)
            geo = abs(np.dot(nx, -to_x)) / np.dot(x - y, x - y) / cj[2]
            wo_y = np.array(rt["wo"], np.float64)
            expect = geo * _pdf(np.array(rt["nrm"], np.float64), wo_y, to_x) / cj[0]
            wi = np.array(r.rc_vertex_wi[:], np.float64)
            is_final = _plen(r) == _k(r)
            if not is_final or not (r.path_flags >> 16) & 1:
                expect *= _pdf(nx, -to_x, wi) / cj[1]
            assert abs(jac - expect) <= 2e-5 * expect, (px, nb, jac, expect)
            checked += 1
    assert checked >= 5 and rejected > 0
