"""The oracle's RIS estimate against an FP64 model of its expectation, written
from the reference shaders (not from the oracle's code):

* RIS, restir_resampling.glsl:19-255, cube strategy: N_c candidates, each a
  uniform light index (urnd_interval) and a light sample; candidate weight
  w_i = luminance(F_i) * m_i with F_i = brdf * Le * G (no pdf division, no
  visibility) and m_i = 1 / max(nee * p_L + n_brdf * p_brdf * G, 1e-6);
  nee = point samples + cube samples and n_brdf = brdf_count + env samples,
  counting the inactive point and environment strategies (one each, their
  minimum counts) and brdf_count = 1 (main.cpp:1317-1320); p_L = 1/N_L
  divided by the cube's area 6 s^2 (light.glsl:350-353).
* The visibility pass, reservoir.glsl:114-126: W_y = W_sum / luminance(F)
  (no 1/M); luminance is Rec.709 (light.glsl:78-80).
* sample_lights, light.glsl:295-399 (KNOWN_LIGHT_POSITION): a face drawn
  uniformly, a point in ONE quadrant of it (random_quad from the face centre
  with half the face size, prng.glsl:156-173, tangents of
  calculate_orthonormal_basis), redrawn up to 11 times until the face
  faces P and lies above P's surface; all draws failing gives F = 0.
* Lambertian evaluate_material / sample_material_pdf with to_local reduced
  to the cosine with the normal (mat.glsl:24-27, 58-88; light.glsl:200-229).

With no occluders, resampling gives E[F_y W_y] = N_c E_x[m(x) F(x)] whatever
the selection weights are, and E[W_sum] = N_c E_x[m(x) luminance(F(x))];
the model integrates both over the light samples in float64 (midpoint rule
over each valid face's quadrant; face validity is constant over an
axis-aligned face).  Mistakes the test rejects (each asserted): W_y divided
by M, Rec.601 luminance, whole-face sampling, MIS denominators without the
inactive strategies.

The BRDF strategy (CT_RIS_BRDF_BIT, restir_resampling.glsl:190-252) has its
own model: the integral of its one cosine-distributed candidate over the
light faces it can reach, in area measure.  Mistakes it rejects: the light
term of m_b without G, F_b divided by the direction pdf, and the visibility
pass (which would shade the BRDF pick) run on its selection.
"""
import numpy as np
import pytest

import ctamd
import oracle as O

E = 0.03125  # VOXEL_EXTENT, shared.inl:45
INVALID = 0xFFFFFFFF
FACES = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], np.float64)


def _basis(n):
    """calculate_orthonormal_basis, prng.glsl:137-145."""
    if abs(n[0]) > abs(n[1]):
        t1 = np.array([-n[2], 0.0, n[0]])
    else:
        t1 = np.array([0.0, n[2], -n[1]])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    return t1, t2 / np.linalg.norm(t2)


def _lum709(c):
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]


def _lum601(c):
    return 0.299 * c[..., 0] + 0.587 * c[..., 1] + 0.114 * c[..., 2]


def _scene(light_cells, colours, hidden=0):
    """A 9x9 floor of voxels (top face at y = e) and light voxels at the given
    grid cells, in one instance; the lights are cube lights of size e.
    `hidden` more lights of radiance 1 fill a 10x10 block of layers under the
    floor (y <= -3 e), where no sample of a point above the floor reaches
    them: they only raise the light count."""
    under = [(j % 10 - 5, -3 - j // 100, (j // 10) % 10 - 5) for j in range(hidden)]
    light_cells = list(light_cells) + under
    colours = list(colours) + [(1.0, 1.0, 1.0)] * hidden
    cells = [(i, 0, k) for i in range(-4, 5) for k in range(-4, 5)] + light_cells
    aabbs = np.array([[c[0] * E, c[1] * E, c[2] * E, (c[0] + 1) * E, (c[1] + 1) * E, (c[2] + 1) * E] for c in cells],
                     np.float32)
    nf = 81
    prims = np.array([[0, INVALID]] * nf + [[1 + j, j] for j in range(len(light_cells))], np.uint32)
    mats = []
    for m in range(1 + len(light_cells)):
        mt = ctamd.Material()
        mt.type = 0
        mt.diffuse[:] = [0.6, 0.5, 0.4] if m == 0 else [0.5, 0.5, 0.5]
        if m > 0:
            mt.emission[:] = colours[m - 1]
        mats.append(mt)
    lights = []
    for j, c in enumerate(light_cells):
        L = ctamd.Light()
        L.position[:] = [(c[0] + 0.5) * E, (c[1] + 0.5) * E, (c[2] + 0.5) * E]
        L.emissive[:] = colours[j]
        L.instance_info.instance_id = 0
        L.instance_info.primitive_id = nf + j
        L.size = E
        L.type = ctamd.CT_LIGHT_CUBE
        lights.append(L)
    inst = ctamd.Instance()
    inst.transform[:] = np.eye(4, dtype=np.float32).reshape(16).tolist()
    inst.first_primitive_index = 0
    inst.primitive_count = len(cells)
    w2o = np.eye(4, dtype=np.float32).reshape(1, 16)
    return O.OScene([inst], w2o, aabbs, prims, mats, lights)


def _params():
    p = ctamd.FrameParams()
    p.width = p.height = 1
    p.is_active = ctamd.CT_RIS_CUBE_LIGHT_BIT
    p.restir = ctamd.restir_params_reference()
    return p


def _di(pos, origin):
    d = ctamd.DiInfo()
    d.position[:] = pos
    d.normal[:] = [0.0, 1.0, 0.0]
    d.ray_origin[:] = origin
    d.distance = float(np.linalg.norm(np.subtract(pos, origin)))
    d.seed = 0
    d.instance_hit.instance_id = 0
    d.instance_hit.primitive_id = 40  # the floor voxel (0, 0, 0)
    d.mat_index = 0
    d.confidence = 0.0
    return d


def _model(P, wo, light_cells, colours, albedo, n_c, n_lights, grid=160, quadrant=True, nee=None, n_brdf=2,
           lum=_lum709):
    """float64 E[F W_y] (rgb), E[W_sum] and the per-light E[F W_y 1{Y=k}]."""
    nrm = np.array([0.0, 1.0, 0.0])
    nee = 1 + n_c if nee is None else nee
    p_l = (1.0 / n_lights) / (6.0 * E * E)
    cos_o = np.dot(wo, nrm)
    u = (np.arange(grid) + 0.5) / grid
    uu, vv = np.meshgrid(u, u, indexing="ij")
    uu, vv = uu.ravel(), vv.ravel()
    if not quadrant:  # the whole face: [-1, 1)^2 times the half size
        uu, vv = 2.0 * uu - 1.0, 2.0 * vv - 1.0
    per_light, per_light_w = [], []
    for c, le in zip(light_cells, colours):
        centre = (np.array(c, np.float64) + 0.5) * E
        acc = np.zeros(3)
        acc_w = 0.0
        valid_faces = 0
        for n_f in FACES:
            base = centre + n_f * (E / 2)
            if np.dot(P - base, n_f) <= 0.0 or np.dot(P - base, nrm) >= 0.0:
                continue  # the face does not face P, or lies below P's surface
            valid_faces += 1
            t1, t2 = _basis(n_f)
            pts = base[None] + np.outer(uu * (E / 2), t1) + np.outer(vv * (E / 2), t2)
            d = pts - P[None]
            dist2 = np.einsum("ij,ij->i", d, d)
            wi = d / np.sqrt(dist2)[:, None]
            G = np.abs(wi @ n_f) / dist2
            cos_i = wi @ nrm
            ok = np.minimum(cos_o, cos_i) >= 1e-6
            brdf_s = np.where(ok, cos_i / np.pi, 0.0)       # albedo / pi * cos_i without the albedo
            pdf_b = np.where(ok, cos_i / np.pi, 0.0)
            m = 1.0 / np.maximum(nee * p_l + n_brdf * pdf_b * G, 1e-6)
            F = (brdf_s * G)[:, None] * (np.asarray(albedo) * np.asarray(le))[None]
            acc += (F * m[:, None]).mean(axis=0)
            acc_w += (lum(F) * m).mean()
        if valid_faces == 0:
            per_light.append(np.zeros(3))
            per_light_w.append(0.0)
            continue
        # the first valid draw is uniform over the valid faces; 11 draws in all
        p_ok = 1.0 - (1.0 - valid_faces / 6.0) ** 11
        per_light.append(p_ok * acc / valid_faces)
        per_light_w.append(p_ok * acc_w / valid_faces)
    scale = n_c / n_lights
    per_light = [scale * x for x in per_light]
    return sum(per_light), scale * sum(per_light_w), per_light


CASES = {
    "one_light": ([(1, 3, 0)], [(9.0, 2.0, 1.0)]),
    "three_lights": ([(1, 3, 0), (-3, 2, -2), (0, 2, 3)], [(9.0, 2.0, 1.0), (1.0, 8.0, 2.0), (1.0, 1.5, 12.0)]),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_ris_mean_matches_the_fp64_expectation(case):
    cells, colours = CASES[case]
    osc = _scene(cells, colours)
    p = _params()
    pos = np.array([0.3 * E, E, 0.45 * E], np.float32)
    di = _di(pos, [0.4, 0.9, 0.7])
    n = 1 << 17
    seeds = np.random.default_rng(7).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    res = osc.probe_ris(p, di, seeds)
    Y = res[:, 0]
    Wy = res[:, 3].view(np.float32).astype(np.float64)
    Wsum = res[:, 4].view(np.float32).astype(np.float64)
    M = res[:, 5].view(np.float32)
    F = res[:, 6:9].view(np.float32).astype(np.float64)
    n_l = len(cells)
    n_c = max(min(32, n_l), 4)                         # restir_resampling.glsl:54-57 at confidence 0
    assert np.all(M == n_c)
    est = F * Wy[:, None]
    mean, se = est.mean(axis=0), est.std(axis=0) / np.sqrt(n)
    wmean, wse = Wsum.mean(), Wsum.std() / np.sqrt(n)
    # the shading point as the pass sees it: hit_from_di pushes it off the surface twice
    P = O.compute_ray_origin(O.compute_ray_origin(pos, [0, 1, 0]), [0, 1, 0]).astype(np.float64)
    wo = -(pos - np.array(di.ray_origin[:], np.float64))
    wo /= np.linalg.norm(wo)
    albedo = (0.6, 0.5, 0.4)
    exp, exp_w, per = _model(P, wo, cells, colours, albedo, n_c, n_l)
    tol = 5.0 * se + 2e-4 * np.abs(exp)
    assert np.all(np.abs(mean - exp) < tol), (mean, exp, se)
    assert abs(wmean - exp_w) < 5.0 * wse + 2e-4 * exp_w, (wmean, exp_w, wse)
    for k in range(n_l):  # each light's share of the estimate
        ek = (est * (Y == k)[:, None])
        mk, sk = ek.mean(axis=0), ek.std(axis=0) / np.sqrt(n)
        assert np.all(np.abs(mk - per[k]) < 5.0 * sk + 2e-4 * np.abs(per[k]) + 1e-12), (k, mk, per[k])
    # each plausible slip is rejected by the same data
    assert np.any(np.abs(mean / n_c - exp) > 20 * se)                                   # W_y / M
    _, w601, _ = _model(P, wo, cells, colours, albedo, n_c, n_l, lum=_lum601)
    assert abs(wmean - w601) > 20 * wse                                                  # Rec.601 weights
    full, _, _ = _model(P, wo, cells, colours, albedo, n_c, n_l, quadrant=False)
    assert np.any(np.abs(mean - full) > 8 * se)                                          # whole-face samples
    active_only, _, _ = _model(P, wo, cells, colours, albedo, n_c, n_l, nee=n_c, n_brdf=1)
    assert np.any(np.abs(mean - active_only) > 8 * se)                                   # active strategies only


def _brdf_model(P, wo, light_cells, colours, albedo, n_lights, nee, n_brdf=2, grid=160, light_g=True,
                divide_pdf=False, lum=_lum709):
    """float64 E[F_b m_b] (rgb, per light) and E[luminance(F_b) m_b] of the
    BRDF strategy's one candidate: a cosine-distributed direction
    (random_cosine_direction, prng.glsl:84-108) that reaches a light face,
    F_b = brdf * Le * G and m_b = 1 / max(n_brdf * p_b * G + nee * p_L * G,
    1e-6) with p_L = (1 / N_L) / (6 e^2) (restir_resampling.glsl:205-224,
    light.glsl:272-293).  The direction density in area measure is
    (cos_i / pi) * G; every face whose outward normal points at P is hit
    first along its directions (one convex voxel, nothing in between)."""
    nrm = np.array([0.0, 1.0, 0.0])
    p_l = (1.0 / n_lights) / (6.0 * E * E)
    cos_o = np.dot(wo, nrm)
    u = ((np.arange(grid) + 0.5) / grid) * 2.0 - 1.0
    uu, vv = np.meshgrid(u, u, indexing="ij")
    uu, vv = uu.ravel(), vv.ravel()
    dA = (E / grid) ** 2
    per, per_w = [], []
    for c, le in zip(light_cells, colours):
        centre = (np.array(c, np.float64) + 0.5) * E
        acc, acc_w = np.zeros(3), 0.0
        for n_f in FACES:
            base = centre + n_f * (E / 2)
            if np.dot(P - base, n_f) <= 0.0:
                continue
            t1, t2 = _basis(n_f)
            pts = base[None] + np.outer(uu * (E / 2), t1) + np.outer(vv * (E / 2), t2)
            d = pts - P[None]
            dist2 = np.einsum("ij,ij->i", d, d)
            wi = d / np.sqrt(dist2)[:, None]
            G = np.abs(wi @ n_f) / dist2
            cos_i = wi @ nrm
            ok = np.minimum(cos_o, cos_i) >= 1e-6
            pdf_b = np.where(ok, cos_i / np.pi, 0.0)
            m = 1.0 / np.maximum(n_brdf * pdf_b * G + nee * p_l * (G if light_g else 1.0), 1e-6)
            F = (np.where(ok, cos_i / np.pi, 0.0) * G)[:, None] * (np.asarray(albedo) * np.asarray(le))[None]
            if divide_pdf:
                F = F / np.where(ok, pdf_b, 1.0)[:, None]
            dens = np.maximum(cos_i, 0.0) / np.pi * G
            acc += (dens[:, None] * F * m[:, None]).sum(axis=0) * dA
            acc_w += (dens * lum(F) * m).sum() * dA
        per.append(acc)
        per_w.append(acc_w)
    return sum(per), sum(per_w), per


def test_ris_brdf_strategy_matches_the_fp64_expectation():
    """RIS with CT_RIS_BRDF_BIT (restir_resampling.glsl:190-252) over one
    light above the floor and 1000 hidden ones under it, so that the BRDF
    candidate often wins the merge:

    * M counts the cube candidates and the one BRDF candidate;
    * a reservoir whose BRDF candidate was selected skips the visibility
      pass: W_y = 0 with F = F_b kept; its F_b W_sum / luminance(F_b) has
      mean E[F_b m_b] (the merge picks it with probability w_b / W_sum),
      and it names the light the ray hit;
    * the cube candidates' E[F W_y] is unchanged by the extra strategy (their
      W_y divides the whole W_sum), and E[W_sum] gains E[luminance(F_b) m_b].
    """
    cells, colours = [(1, 3, 0)], [(9.0, 2.0, 1.0)]
    hidden = 1000
    osc = _scene(cells, colours, hidden=hidden)
    assert osc.config.brdf_count == 1
    p = _params()
    p.is_active = ctamd.CT_RIS_CUBE_LIGHT_BIT | ctamd.CT_RIS_BRDF_BIT
    pos = np.array([0.3 * E, E, 0.45 * E], np.float32)
    di = _di(pos, [0.4, 0.9, 0.7])
    n = 1 << 17
    seeds = np.random.default_rng(11).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    res = osc.probe_ris(p, di, seeds)
    Y = res[:, 0]
    Wy = res[:, 3].view(np.float32).astype(np.float64)
    Wsum = res[:, 4].view(np.float32).astype(np.float64)
    M = res[:, 5].view(np.float32)
    F = res[:, 6:9].view(np.float32).astype(np.float64)
    n_l = 1 + hidden
    n_c = 32
    assert np.all(M == n_c + 1)
    P = O.compute_ray_origin(O.compute_ray_origin(pos, [0, 1, 0]), [0, 1, 0]).astype(np.float64)
    wo = -(pos - np.array(di.ray_origin[:], np.float64))
    wo /= np.linalg.norm(wo)
    albedo = (0.6, 0.5, 0.4)
    nee = 1 + n_c
    # cube candidates
    est = F * Wy[:, None]
    mean, se = est.mean(axis=0), est.std(axis=0) / np.sqrt(n)
    exp, exp_w, _ = _model(P, wo, cells, colours, albedo, n_c, n_l)
    assert np.all(np.abs(mean - exp) < 5.0 * se + 2e-4 * np.abs(exp)), (mean, exp, se)
    # the BRDF candidate
    lumF = _lum709(F)
    sel = (Wy == 0.0) & (lumF > 0.0)
    assert sel.sum() > 2000, sel.sum()
    assert np.all(Y[sel] == 0)
    est_b = np.where(sel[:, None], F * (Wsum / np.where(sel, lumF, 1.0))[:, None], 0.0)
    mean_b, se_b = est_b.mean(axis=0), est_b.std(axis=0) / np.sqrt(n)
    exp_b, exp_bw, _ = _brdf_model(P, wo, cells, colours, albedo, n_l, nee)
    assert np.all(np.abs(mean_b - exp_b) < 5.0 * se_b + 2e-4 * np.abs(exp_b)), (mean_b, exp_b, se_b)
    wmean, wse = Wsum.mean(), Wsum.std() / np.sqrt(n)
    assert abs(wmean - (exp_w + exp_bw)) < 5.0 * wse + 2e-4 * (exp_w + exp_bw), (wmean, exp_w, exp_bw, wse)
    assert exp_bw > 8 * wse  # the BRDF share of W_sum is resolved
    # each plausible slip is rejected by the same data
    no_g, _, _ = _brdf_model(P, wo, cells, colours, albedo, n_l, nee, light_g=False)
    assert np.any(np.abs(mean_b - no_g) > 8 * se_b)                                   # p_L without G
    by_pdf, _, _ = _brdf_model(P, wo, cells, colours, albedo, n_l, nee, divide_pdf=True)
    assert np.any(np.abs(mean_b - by_pdf) > 8 * se_b)                                 # F_b / p_b
    assert np.any(np.abs(mean - (exp + exp_b)) > 8 * se)                              # BRDF pick shaded
