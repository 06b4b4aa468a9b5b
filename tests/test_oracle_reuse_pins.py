"""Pins of the oracle's reuse passes by what fixes them independently of its
code:

* temporal reuse (restir_resampling.glsl:311-416): a history that holds
  the pixel's own sample leaves W_y unchanged for any M split;
* pairwise MIS (pairwise_mis.glsl:8-114, restir_resampling.glsl:418-540):
  k neighbours that are copies of the pixel itself (same surface, same
  reservoir) give m_i = k/(k+1) per copy and m_c = 1 + k/(k+1), so the
  combined W_sum = p_hat W_y (k + 1) and, divided by p_hat (1 + k), W_y is
  unchanged; the sample and its F stay, and M becomes (k + 1) M;
* the reconnection shift (shift.glsl:25-350): from a primary vertex to
  itself its Jacobian is 1 and it returns the path's own integrand (times
  the primary vertex's albedo, which the reservoir's F lacks by the
  reference's prefix-throughput overwrite); from a
  neighbour's vertex its Jacobian is the float64 product of the geometry
  ratio (cos / d^2 at the reconnection vertex, GRIS) and the pdf ratios the
  shader forms, computed here from the vertices alone;
* TAA (taa.glsl:16-108): with blend alpha 1 the output is the 3x3
  edge-aware mean of the TAA input (float64 filter over the rejection
  rules); a scene of one material whose shaded value is constant stays
  constant under any alpha and camera motion.
"""
import ctypes as C

import numpy as np
import pytest

import ctamd
import oracle as O
import test_oracle_ris_expectation as R
from helpers import di_array

INVALID = 0xFFFFFFFF


def _frame(w, frames, mutate_params=None, scene=None):
    sc = ctamd.Scene(ctamd.scene_desc(w)) if scene is None else scene
    osc = O.OScene.from_scene(sc)
    fr = O.OFrame(osc, w.width, w.height)
    out = []
    for f in range(frames):
        p = ctamd.frame_params(w, f)
        if mutate_params:
            mutate_params(p)
        before = fr.read(ctamd.CT_BUF_IMAGE) if f > 0 else None
        prev_di = fr.read(ctamd.CT_BUF_PREV_DI) if f > 0 else None
        fr.render(p, threads=4)
        out.append(dict(p=p, prev_img=before, prev_di=prev_di, img=fr.read(ctamd.CT_BUF_IMAGE),
                        taa=fr.read(ctamd.CT_BUF_TAA_FRAME), di=fr.read(ctamd.CT_BUF_DI),
                        vel=fr.read(ctamd.CT_BUF_VELOCITY), res=fr.read(ctamd.CT_BUF_INT_RESERVOIR),
                        pres=fr.read(ctamd.CT_BUF_PATH_RESERVOIR)))
    return sc, osc, fr, out


# ------------------------------------------------------------ pairwise MIS
@pytest.fixture(scope="module")
def floor():
    """The lit floor of the RIS expectation test and RIS reservoirs at a few
    shading points on it (every light in view, so most samples are valid)."""
    cells, colours = R.CASES["three_lights"]
    osc = R._scene(cells, colours)
    p = R._params()
    points = []
    rng = np.random.default_rng(3)
    for j in range(6):
        pos = np.array([rng.uniform(-3, 3) * R.E, R.E, rng.uniform(-3, 3) * R.E], np.float32)
        di = R._di(pos, [0.4 - 0.1 * j, 0.9, 0.7])
        words = osc.probe_ris(p, di, rng.integers(0, 2 ** 32, 12, dtype=np.uint64).astype(np.uint32))
        for w in words:
            r = ctamd.Reservoir.from_buffer_copy(w.tobytes())
            if r.W_y > 0:
                points.append((di, r))
    assert len(points) > 40
    return osc, p, points


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_pairwise_mis_with_copies_of_the_own_sample_keeps_W(floor, k):
    osc, p, points = floor
    for n, (di, r) in enumerate(points):
        got = osc.probe_pairwise(p, di, r, [di] * k, [r] * k, seed=1234 + n)
        assert got.Y == r.Y and got.l_type == r.l_type and got.seed == r.seed
        assert list(got.F) == list(r.F)
        assert got.M == (k + 1) * r.M
        assert abs(got.W_y - r.W_y) <= 4e-6 * r.W_y, (n, got.W_y, r.W_y)


def test_pairwise_mis_with_a_neighbour_that_sees_nothing(floor):
    """A neighbour with no sample of its own, facing away from the lights (its
    target of the pixel's sample is 0): m_c = 1 + 1, the pixel's weight is
    doubled and so is the divisor 1 + k: W_y unchanged, M adds the
    neighbour's."""
    osc, p, points = floor
    for n, (di, r) in enumerate(points[:20]):
        nd = ctamd.DiInfo.from_buffer_copy(bytes(di))
        nd.normal[:] = [-v for v in di.normal]
        nr = ctamd.Reservoir.from_buffer_copy(bytes(r))
        nr.Y = INVALID
        got = osc.probe_pairwise(p, di, r, [nd], [nr], seed=99 + n)
        assert got.Y == r.Y and got.M == r.M + nr.M
        assert abs(got.W_y - r.W_y) <= 4e-6 * r.W_y, (n, got.W_y, r.W_y)


@pytest.mark.parametrize("m_prev", [32.0, 96.0, 640.0, 5000.0])
def test_temporal_reuse_of_the_own_sample_keeps_W(floor, m_prev):
    """The previous reservoir holds the same sample at the same surface
    (static pixel): the balance-heuristic weights are M/(M + M') and
    M'/(M + M') of the same p_hat, which sum to one, so W_y is unchanged for
    any M split (M' capped at 20 M, restir_resampling.glsl:317-319) and M
    becomes M + min(M', 20 M)."""
    osc, p, points = floor
    for n, (di, r) in enumerate(points):
        prev = ctamd.Reservoir.from_buffer_copy(bytes(r))
        prev.M = m_prev
        got = osc.probe_temporal(p, di, r, di, prev, seed=77 + n)
        assert got.Y == r.Y and list(got.F) == list(r.F)
        assert got.M == np.float32(r.M + min(m_prev, 20.0 * r.M))
        assert abs(got.W_y - r.W_y) <= 4e-6 * r.W_y, (n, got.W_y, r.W_y)


def test_temporal_reuse_with_a_history_without_sample(floor):
    """A previous reservoir with M' but no sample, at the same surface: the
    current sample's balance weight is M p / (M p + M' p) (the history could
    have produced it) and nothing else enters, so W_y scales by M / (M + M')
    and M becomes M + M'."""
    osc, p, points = floor
    for n, (di, r) in enumerate(points[:20]):
        prev = ctamd.Reservoir()
        prev.Y = INVALID
        prev.M = 5.0
        got = osc.probe_temporal(p, di, r, di, prev, seed=5 + n)
        assert got.Y == r.Y and got.M == r.M + 5.0
        expect = r.W_y * r.M / (r.M + 5.0)
        assert abs(got.W_y - expect) <= 4e-6 * expect, (got.W_y, expect)


# ------------------------------------------------------ reconnection shift
@pytest.fixture(scope="module")
def pt_frame():
    w = ctamd.small_plain(width=48, height=48, frames=1, indirect=True, max_depth=3)
    return _frame(w, 1)


def _pdf(n, wo, wi):
    """Lambertian sample_material_pdf with to_local's cosine, light.glsl:200-229."""
    n = n / np.linalg.norm(n)
    a, b = np.dot(n, wo), np.dot(n, wi)
    return 0.0 if min(a, b) < 1e-6 else b / np.pi


def _pushed(p, n):
    """compute_new_ray_origin(view side), primitives.glsl:259-264."""
    return O.compute_ray_origin(O.compute_ray_origin(p, n), n).astype(np.float64)


def _entry(origin, target, centre, half):
    """First point where the segment origin -> target enters the voxel box
    (load_vertex's re-intersection, primitives.glsl:567-611), and its face
    normal, in float64."""
    d = target - origin
    lo, hi = centre - half, centre + half
    with np.errstate(divide="ignore"):
        t1, t2 = (lo - origin) / d, (hi - origin) / d
    tn = np.minimum(t1, t2)
    ax = int(np.argmax(tn))
    t = tn[ax]
    nrm = np.zeros(3)
    nrm[ax] = -np.sign(d[ax])
    return origin + t * d, nrm


def test_identity_shift_has_jacobian_one(pt_frame):
    sc, osc, fr, out = pt_frame
    o = out[0]
    di, pres = o["di"], o["pres"]
    seen = 0
    for px in range(len(di)):
        r = pres[px]
        if not (di[px].distance > 0 and r.weight > 0 and r.rc_vertex_hit.object.instance_id != INVALID
                and (r.path_flags >> 4) & 0xF == 1):
            continue
        F = np.array(r.F[:], np.float32)
        if not np.any(F > 0):
            continue
        jac, f = osc.probe_shift(o["p"], di[px], di[px], r)
        if jac == 0.0:
            # not reconnectable from the primary vertex pushed along its normal
            # (the path left it pushed along its direction): the shift says so
            assert not np.any(f), px
            continue
        assert abs(jac - 1.0) <= 4e-7, (px, jac)
        # The shift recomputes f/pdf at the primary vertex (the albedo, for a
        # Lambertian surface) while the reservoir's F lacks it: the path
        # builder records the prefix throughput at every vertex up to the
        # reconnection vertex, so the reconnection vertex's overwrites the
        # primary's (path_tracer.glsl:104-107, path_state.glsl:160-163).
        albedo = np.array(sc.materials[di[px].mat_index].diffuse[:], np.float32)
        assert np.allclose(f, F * albedo, rtol=1e-5, atol=0), (px, f, F * albedo)
        seen += 1
    assert seen > 30


def test_neighbour_shift_jacobian_matches_fp64_geometry(pt_frame):
    sc, osc, fr, out = pt_frame
    o = out[0]
    di, pres = o["di"], o["pres"]
    W = o["p"].width
    checked = 0
    for px in range(len(di)):
        r = pres[px]
        if not (di[px].distance > 0 and r.weight > 0 and r.rc_vertex_hit.object.instance_id != INVALID
                and (r.path_flags >> 4) & 0xF == 1):
            continue
        nb = px + 1 if (px % W) + 1 < W else px - 1
        if not di[nb].distance > 0:
            continue
        jac, f = osc.probe_shift(o["p"], di[nb], di[px], r)
        if jac == 0.0:
            continue
        flags = r.path_flags
        path_length, last_nee = flags & 0xF, bool((flags >> 16) & 1)
        is_final = path_length == 1
        escaped = path_length + 1 == 1 and not last_nee
        rc_nee = is_final and last_nee
        src_n, dst_n = np.array(di[px].normal[:], np.float64), np.array(di[nb].normal[:], np.float64)
        src_P = _pushed(np.array(di[px].position[:], np.float32), src_n)
        dst_P = _pushed(np.array(di[nb].position[:], np.float32), dst_n)
        src_wo = np.array(di[px].ray_origin[:], np.float64) - np.array(di[px].position[:], np.float64)
        src_wo /= np.linalg.norm(src_wo)
        dst_wo = np.array(di[nb].ray_origin[:], np.float64) - np.array(di[nb].position[:], np.float64)
        dst_wo /= np.linalg.norm(dst_wo)
        h = r.rc_vertex_hit
        box = np.asarray(sc.aabbs[_slot(sc, h.object.instance_id, h.object.primitive_id)], np.float64).reshape(6)
        centre = 0.5 * (box[:3] + box[3:])
        rc_P, rc_n = _entry(dst_P, np.array(h.hit[:], np.float64), centre, 0.5 * 0.03125)
        rc_P = _pushed(rc_P.astype(np.float32), rc_n)  # load_vertex pushes the re-intersection off the face twice
        dst_cv = (rc_P - dst_P) / np.linalg.norm(rc_P - dst_P)
        src_cv = (rc_P - src_P) / np.linalg.norm(rc_P - src_P)
        # geometry ratio at the reconnection vertex (GRIS reconnection shift)
        geo = (abs(np.dot(rc_n, dst_cv)) / np.dot(rc_P - dst_P, rc_P - dst_P)) * \
              (np.dot(rc_P - src_P, rc_P - src_P) / abs(np.dot(rc_n, src_cv)))
        expect = geo * _pdf(dst_n, dst_wo, dst_cv) / _pdf(src_n, src_wo, src_cv)
        wi = np.array(r.rc_vertex_wi[:], np.float64)
        if (not is_final and not escaped) or (is_final and not rc_nee and ((flags >> 18) & 3) != 4):
            expect *= _pdf(rc_n, -dst_cv, wi) / _pdf(rc_n, -src_cv, wi)
        assert abs(jac - expect) <= 2e-6 * expect, (px, jac, expect)
        checked += 1
    assert checked > 20


def _slot(sc, inst, prim):
    return sc.instances[inst].first_primitive_index + prim


# ---------------------------------------------------------------------- TAA
def _taa_filter(di, taa, W, H):
    """taa.glsl:30-59 in float64: mean over the 3x3 neighbours with the same
    material, normals within 0.906 and depth within 10 %."""
    dist = np.array([d.distance for d in di]).reshape(H, W)
    mat = np.array([d.mat_index for d in di]).reshape(H, W)
    nrm = np.array([d.normal[:] for d in di], np.float64).reshape(H, W, 3)
    t = taa.astype(np.float64).reshape(H, W, 4)
    out = np.full((H, W, 4), np.nan)
    for y in range(H):
        for x in range(W):
            if not dist[y, x] > 0:
                continue
            acc, cnt = np.zeros(4), 0
            for j in (-1, 0, 1):
                for i in (-1, 0, 1):
                    nx, ny = x + i, y + j
                    if not (0 <= nx < W and 0 <= ny < H):
                        continue
                    if mat[ny, nx] != mat[y, x] or np.dot(nrm[y, x], nrm[ny, nx]) < 0.906 or \
                            abs(dist[y, x] - dist[ny, nx]) > 0.1 * dist[y, x]:
                        continue
                    acc += t[ny, nx]
                    cnt += 1
            out[y, x] = acc / cnt
    return out.reshape(H * W, 4)


def test_taa_alpha_one_is_the_spatial_filter():
    w = ctamd.hd_di(width=40, height=32, grid=48, chunk=16, frames=3, max_depth=1)

    def alpha1(p):
        p.restir.taa_alpha = 1.0

    sc, osc, fr, out = _frame(w, 3, alpha1)
    for o in out[1:]:
        W, H = o["p"].width, o["p"].height
        expect = _taa_filter(o["di"], o["taa"], W, H)
        hit = np.array([d.distance > 0 for d in o["di"]])
        tw, th = (W // 8) * 8, (H // 8) * 8  # taa dispatch truncation, main.cpp:1642-1643
        inside = np.zeros((H, W), bool)
        inside[:th, :tw] = True
        sel = hit & inside.ravel()
        assert sel.sum() > 200
        assert np.allclose(o["img"][sel], expect[sel], rtol=2e-6, atol=0)


def _one_material(w):
    """The workload's voxels with one material of albedo 0 and emission c:
    every hit pixel shades to c (DIRECT_EMITTANCE_ON, no indirect term)."""
    sc = ctamd.Scene(ctamd.scene_desc(w))
    prims = sc.primitives.copy()
    prims[:, 0] = 0
    m = ctamd.Material()
    m.type = 0
    m.emission[:] = [0.75, 0.375, 1.5]
    ni = sc.info.instance_count
    osc = O.OScene(list(sc.instances), np.tile(np.eye(4, dtype=np.float32).reshape(16), (ni, 1)), sc.aabbs, prims,
                   [m], list(sc.lights))
    return O.OFrame(osc, w.width, w.height)


def test_taa_keeps_a_constant_image_constant():
    """A static camera over one material whose shaded value is constant:
    filter, reprojection and the alpha 0.8 blend keep it constant."""
    w = ctamd.small_plain(width=40, height=32, frames=4, taa=True, indirect=False)
    fr = _one_material(w)
    for f in range(4):
        p = ctamd.frame_params(w, f)
        assert f == 0 or p.is_active & ctamd.CT_TAA_BIT
        fr.render(p, threads=4)
        img = fr.read(ctamd.CT_BUF_IMAGE)
        hit = np.array([d.distance > 0 for d in fr.read(ctamd.CT_BUF_DI)])
        assert hit.sum() > 300
        assert np.allclose(img[hit, :3], [0.75, 0.375, 1.5], rtol=1e-6, atol=0), f


def test_taa_history_check_reads_this_frames_di():
    """Defined by the reference's frame order: the shading pass writes the
    previous-DI buffer (rgen.glsl:503-506) before TAA reads it for its
    history check (taa.glsl:88-96), so the check sees this frame's DI at the
    reprojected pixel.  Under an orbiting camera a hit pixel whose
    reprojected pixel was sky in the last frame but is a hit of the same
    material now blends with the sky colour."""
    w = ctamd.hd_di(width=40, height=32, grid=48, chunk=16, frames=3, indirect=False)
    fr = _one_material(w)
    fr.render(ctamd.frame_params(w, 0), threads=4)
    prev_img = fr.read(ctamd.CT_BUF_IMAGE)
    prev_hit = np.array([d.distance > 0 for d in fr.read(ctamd.CT_BUF_DI)])
    p = ctamd.frame_params(w, 1)
    fr.render(p, threads=4)
    img, di, vel = fr.read(ctamd.CT_BUF_IMAGE), fr.read(ctamd.CT_BUF_DI), fr.read(ctamd.CT_BUF_VELOCITY)
    c = np.array([0.75, 0.375, 1.5], np.float32)
    a = np.float32(p.restir.taa_alpha)
    blended_with_sky = 0
    for i, d in enumerate(di):
        x, y = i % p.width, i // p.width
        if not d.distance > 0 or x >= (p.width // 8) * 8 or y >= (p.height // 8) * 8:
            continue
        s = d.seed
        jx, s = O.rnd(s)
        jy, s = O.rnd(s)
        px, py = int(np.float32(np.float32(x + vel[i, 0]) + jx)), int(np.float32(np.float32(y + vel[i, 1]) + jy))
        if not (0 <= px < p.width and 0 <= py < p.height):
            assert np.allclose(img[i, :3], c, rtol=1e-6), i
            continue
        q = py * p.width + px
        now = di[q]
        valid = now.distance > 0 and np.dot(now.normal[:], d.normal[:]) >= 0.99
        expect = a * c + (np.float32(1) - a) * prev_img[q, :3] if valid else c
        assert np.allclose(img[i, :3], expect, rtol=1e-5), i
        blended_with_sky += int(valid and not prev_hit[q])
    assert blended_with_sky > 0
