// restir_pt.h -- ReSTIR PT on the device: the initial path tracer
// (trace_restir_path_tracing, path_tracer.glsl:854-887), the bounce
// closest hit it traces (intersect(), bounce.glsl:9-130), the reconnection
// shift (shift.glsl:25-350, 445-542), temporal path reuse
// (indirect_illumination.glsl:147-279) and spatial path reuse (:312-452).
//
// Included by frame.hip inside its anonymous namespace, after the DI
// helpers (Mat, Hit, sample_lights, material_pdf, evaluate_material).
// Each function restates the reference function named beside it for
// Lambertian materials; the arithmetic is written in the GLSL order.
//
// Traversals run inline (closest hit for path segments, any hit for the
// reconnection visibility), with the lane's traversal stack passed in.
#pragma once

constexpr float kHlfMax = 6.5504e+4f;         // HLF_MAX, shared.inl:56
constexpr float kHlfMin = 1.175494351e-38f;   // HLF_MIN, shared.inl:57
constexpr float kVoxelExtent = 0.03125f;      // VOXEL_EXTENT, shared.inl:45
constexpr uint32_t kMaxMaterials = 10000u;    // MAX_MATERIALS, shared.inl:20
constexpr uint32_t kGeomLightCube = 2u, kGeomLightAnalytic = 4u, kGeomLightEnv = 5u;  // shared.inl:336-341
// PATH_FLAG_*, defines.glsl:64-76
constexpr uint32_t kPathActive = 0x1u, kPathHit = 0x2u, kPathTransmission = 0x4u, kPathDelta = 0x10u,
                   kPathLightUpper = 0x80u, kPathLightLower = 0x100u, kPathDiffusePrimary = 0x200u;

struct ObjHit {  // OBJECT_HIT
  uint32_t inst, prim;
  F3 hit;
};
struct Isect {  // INTERSECT, shared.inl:323-334
  bool is_hit;
  float distance;
  F3 P, n, wo, wi;
  uint32_t inst, prim, mat_idx;
  Mat mat;
};
struct PRes {  // PATH_RESERVOIR, shared.inl:405-423
  float M, weight;
  uint32_t flags, rc_seed;
  F3 F;
  float light_pdf;
  F3 cj;
  uint32_t init_seed;
  ObjHit rc;
  F3 rc_wi, rc_irr;
};
struct PBuilder {  // PATH_BUILDER, path_builder.glsl:8-17
  ObjHit rc;
  F3 rc_wi;
  uint32_t cached_seed, seed, rc_len, flags;
  F3 cj;
};
constexpr uint32_t kDirtyRc = 1u, kDirtyRes = 2u;
struct PState {  // PATH_STATE, path_state.glsl:23-78 (fields the Lambertian path reads)
  uint32_t max_depth, flags, path_length, bounce_counters;
  float scene_length;
  F3 origin, dir;
  float pdf;
  F3 normal;
  ObjHit hit;
  F3 thp, prefix_thp, L;
  float rr_pdf, prev_scatter_pdf;
  uint32_t seed;
  PBuilder pb;
  PRes res;
  bool last_rough;
  // Parts of the stored state a path-loop step changed besides those every
  // step rewrites (kDirtyRc: the path builder's reconnection vertex and
  // Jacobian terms; kDirtyRes: the path reservoir).  Not stored.
  uint32_t dirty;
  // Replay of a source path for the hybrid shift (is_replay_for_hybrid_shift):
  // stops at the vertex before the source's reconnection vertex replay_rc.
  bool replay, replay_done;
  uint32_t replay_rc;
  Isect replay_prev;
  F3 replay_tp;
};

// Traversal context of one lane.  The visibility tests of a pass that
// defers them to the any-hit queue use px / slot / bits / emitted as the DI
// passes' Vis does.
struct PtTrace {
  uint32_t *stk;
  TraceCnt cnt;
  uint32_t bounce_rays, shadow_rays;
  bool overflow;
  uint32_t px, slot, bits, emitted, far;  // far: emitted rays of at least F.far_ray
  // Hybrid retraces of temporal path reuse: which of the pixel's two (0: the
  // history shifted to this pixel's primary hit, 1: this pixel's path
  // shifted to the previous pixel's), the previous pixel, and the retraces
  // a request pass asked for (bit per slot).
  uint32_t rc_slot, rc_prev, rc_req;
};

// ---------------------------------------------------------------- helpers
CT_D ObjHit objhit_invalid() {
  ObjHit h;
  h.inst = kMaxInstances;
  h.prim = kMaxPrimitives;
  h.hit = f3(0, 0, 0);
  return h;
}
CT_D bool instance_hit_exists(uint32_t inst, uint32_t prim) {  // primitives.glsl:542-545
  return inst < kMaxInstances && prim < kMaxPrimitives;
}
// path_reservoir.glsl:64-147
CT_D uint32_t pf_path_length(uint32_t f) { return f & 0xFu; }
CT_D uint32_t pf_rc_length(uint32_t f) { return (f >> 4) & 0xFu; }
CT_D void pf_set_path_length(uint32_t &f, uint32_t l) { f = (f & ~0xFu) | (l & 0xFu); }
CT_D void pf_set_rc_length(uint32_t &f, uint32_t l) { f = (f & ~0xF0u) | ((l & 0xFu) << 4); }
CT_D bool pf_delta(uint32_t f, bool before) { return ((f >> (before ? 20 : 21)) & 1u) != 0u; }
CT_D bool pf_transmission(uint32_t f, bool before) { return ((f >> (before ? 22 : 23)) & 1u) != 0u; }
CT_D void pf_set_last_nee(uint32_t &f, bool b) { f = (f & ~0x10000u) | ((b ? 1u : 0u) << 16); }
CT_D bool pf_last_nee(uint32_t f) { return ((f >> 16) & 1u) != 0u; }
// insert_light_type keeps two bits, so GEOMETRY_LIGHT_ENV_MAP (5) reads
// back as 1 (path_reservoir.glsl:137-145).
CT_D void pf_set_light_type(uint32_t &f, uint32_t t) { f = (f & ~0xc0000u) | ((t & 3u) << 18); }
CT_D uint32_t pf_light_type(uint32_t f) { return (f >> 18) & 3u; }

CT_D float lum601(F3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }  // path_F_to_scalar
// eval_mis with the power heuristic, light.glsl:171-191
CT_D float eval_mis(float f_pdf, float g_pdf) {
  const float a = 1.0f * f_pdf, b = 1.0f * g_pdf;
  const float f = a * a, g = b * b;
  return f / (f + g);
}
CT_D bool has_nan_inf(F3 v) {
  return isnan(v.x) || isnan(v.y) || isnan(v.z) || isinf(v.x) || isinf(v.y) || isinf(v.z);
}
CT_D bool jacobian_invalid(float j) { return j <= 0.0f || isnan(j) || isinf(j); }  // shift.glsl:12-14

CT_D PRes pres_init(uint32_t max_depth) {  // path_reservoir_initialise, path_reservoir.glsl:151-165
  PRes r;
  r.M = 0.0f;
  r.weight = 0.0f;
  r.flags = 0u;
  pf_set_rc_length(r.flags, max_depth);
  r.rc_seed = 0u;
  r.F = f3(0, 0, 0);
  r.light_pdf = 0.0f;
  r.cj = f3(0, 0, 0);
  r.init_seed = 0u;
  r.rc = objhit_invalid();
  r.rc_wi = f3(0, 0, 0);
  r.rc_irr = f3(0, 0, 0);
  return r;
}
// 6 x float4 per pixel
CT_D void st_pres(float4 *B, uint32_t i, const PRes &r) {
  float4 *q = B + 6 * (size_t)i;
  q[0] = make_float4(r.M, r.weight, __uint_as_float(r.flags), __uint_as_float(r.rc_seed));
  q[1] = make_float4(r.F.x, r.F.y, r.F.z, r.light_pdf);
  q[2] = make_float4(r.cj.x, r.cj.y, r.cj.z, __uint_as_float(r.init_seed));
  q[3] = make_float4(__uint_as_float(r.rc.inst), __uint_as_float(r.rc.prim), r.rc.hit.x, r.rc.hit.y);
  q[4] = make_float4(r.rc.hit.z, r.rc_wi.x, r.rc_wi.y, r.rc_wi.z);
  q[5] = make_float4(r.rc_irr.x, r.rc_irr.y, r.rc_irr.z, 0.0f);
}
CT_D PRes ld_pres(const float4 *B, uint32_t i) {
  const float4 *q = B + 6 * (size_t)i;
  float4 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], g = q[5];
  PRes r;
  r.M = a.x; r.weight = a.y; r.flags = __float_as_uint(a.z); r.rc_seed = __float_as_uint(a.w);
  r.F = f3(b.x, b.y, b.z); r.light_pdf = b.w;
  r.cj = f3(c.x, c.y, c.z); r.init_seed = __float_as_uint(c.w);
  r.rc.inst = __float_as_uint(d.x); r.rc.prim = __float_as_uint(d.y); r.rc.hit = f3(d.z, d.w, e.x);
  r.rc_wi = f3(e.y, e.z, e.w);
  r.rc_irr = f3(g.x, g.y, g.z);
  return r;
}

// path_reservoir_update, path_reservoir.glsl:183-202
CT_D bool pres_update(PRes &r, F3 in_F, float p, uint32_t &seed) {
  r.M += 1.0f;
  const float w = lum601(in_F) / p;
  if (isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (rnd(seed) * r.weight <= w) {
    r.F = in_F;
    return true;
  }
  return false;
}
CT_D void pres_take(PRes &r, const PRes &in, F3 in_F) {
  r.flags = in.flags;
  r.rc_seed = in.rc_seed;
  r.init_seed = in.init_seed;
  r.cj = in.cj;
  r.light_pdf = in.light_pdf;
  r.rc_wi = in.rc_wi;
  r.rc = in.rc;
  r.rc_irr = in.rc_irr;
  r.F = in_F;
}
// path_reservoir_merge, path_reservoir.glsl:204-232
CT_D bool pres_merge(PRes &r, F3 in_F, float jac, const PRes &in, uint32_t &seed, float mis, bool force) {
  const float w = lum601(in_F) * jac * in.M * in.weight * mis;
  r.M += in.M;
  if (isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (force || rnd(seed) * r.weight <= w) {
    pres_take(r, in, in_F);
    return true;
  }
  return false;
}
// path_reservoir_merge_with_resampling_MIS, path_reservoir.glsl:234-263
CT_D bool pres_merge_mis(PRes &r, F3 in_F, float jac, const PRes &in, uint32_t &seed, float mis) {
  const float w = lum601(in_F) * jac * in.weight * mis;
  r.M += in.M;
  if (isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (rnd(seed) * r.weight <= w) {
    pres_take(r, in, in_F);
    return true;
  }
  return false;
}
CT_D void pres_finalize_ris(PRes &r) {  // path_reservoir.glsl:265-271
  const float p_hat = lum601(r.F);
  r.weight = (p_hat == 0.0f || r.M == 0.0f) ? 0.0f : r.weight / (p_hat * r.M);
}
CT_D void pres_finalize_gris(PRes &r) {  // path_reservoir.glsl:273-279
  const float p_hat = lum601(r.F);
  r.weight = (p_hat == 0.0f) ? 0.0f : r.weight / p_hat;
}

// ---------------------------------------------------------------- geometry
CT_D Isect isect_empty() {  // intersect_initiliaze, primitives.glsl:197-209
  Isect i;
  i.is_hit = false;
  i.distance = 0.0f;
  i.P = f3(0, 0, 0); i.n = f3(0, 0, 0); i.wo = f3(0, 0, 0); i.wi = f3(0, 0, 0);
  i.inst = kMaxInstances; i.prim = kMaxPrimitives; i.mat_idx = kMaxMaterials;
  i.mat.diffuse = f3(0, 0, 0); i.mat.emission = f3(0, 0, 0); i.mat.roughness = 0.0f;
  return i;
}
// transpose(world2obj) * vec4(n, 0) for the 3-row packing; the fourth row
// of an affine world-to-object matrix is (0, 0, 0, 1).
CT_D F3 xform_normal(const float *w, F3 n) {
  return f3(((w[0] * n.x + w[1] * n.y) + w[2] * n.z) + 0.0f * 0.0f,
            ((w[3] * n.x + w[4] * n.y) + w[5] * n.z) + 0.0f * 0.0f,
            ((w[6] * n.x + w[7] * n.y) + w[8] * n.z) + 0.0f * 0.0f);
}
CT_D F3 voxel_center(const DFrame &F, uint32_t gp) {
  const ct_aabb &bb = F.aabbs[gp];
  return muls(add(f3(bb.minimum[0], bb.minimum[1], bb.minimum[2]), f3(bb.maximum[0], bb.maximum[1], bb.maximum[2])),
              0.5f);
}
// is_hit_from_ray (primitives.glsl:412-453) with rayCanStartInBox = true
// on the current instance buffer; world-space hit position and normal.
CT_D bool hit_from_ray(const DFrame &F, uint32_t inst, uint32_t prim, F3 ro, F3 rd, float &t, F3 &pos, F3 &nrm) {
  if (inst >= F.S.n_inst) return false;
  const DInst &I = F.S.inst[inst];
  if (prim >= I.count) return false;
  float w[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) w[k] = I.w2o[k];
  F3 oo = xform3(w, ro, 1.0f), od = xform3(w, rd, 0.0f);
  F3 sinv = f3(safe_inverse(od.x), safe_inverse(od.y), safe_inverse(od.z));
  F3 sg = f3(-gsign(od.x), -gsign(od.y), -gsign(od.z));
  F3 on;
  bool h = intersect_box<true, false>(voxel_center(F, I.first + prim), oo, od, sg, sinv, t, &on);
  pos = add(ro, muls(rd, t));
  nrm = xform_normal(w, on);
  return h;
}
CT_D uint32_t material_of(const DFrame &F, uint32_t inst, uint32_t prim) {
  return F.prims[F.S.inst[inst].first + prim].material_index;
}

// intersect(), bounce.glsl:9-130: the box tests start at the origin pushed
// along the ray (compute_ray_origin(origin, dir)); distance and wo use the
// un-pushed origin.
template <bool kCount>
CT_D Isect bounce_intersect(const DFrame &F, F3 origin, F3 dir, PtTrace &tr) {
  const F3 ro = compute_ray_origin(origin, dir);
  float bt = kMaxDistance;
  uint32_t bi = kInvalid, bp = kInvalid;
  ++tr.bounce_rays;
  traverse_mode<kTraceBounce, kCount>(F.S, ro, dir, kMaxDistance, kInvalid, kInvalid, bt, bi, bp, tr.stk, kBlock,
                                      tr.cnt, tr.overflow);
  Isect i = isect_empty();
  i.mat_idx = 0u;
  F3 int_hit = f3(0, 0, 0);
  if (bi != kInvalid) {
    float t;
    F3 pos, nrm;
    hit_from_ray(F, bi, bp, ro, dir, t, pos, nrm);
    int_hit = pos;
    i.is_hit = true;
    i.P = pos;
    i.n = nrm;
    i.inst = bi;
    i.prim = bp;
    i.mat_idx = material_of(F, bi, bp);
    i.mat = ld_mat(F, i.mat_idx);
    i.distance = length(sub(origin, int_hit));
  }
  i.wo = normalize(sub(origin, int_hit));
  return i;
}

// Visibility modes of a pass: traversed in place, queued (taken as visible
// while the ray is recorded in the pixel's next slot), or read back from the
// pixel's occlusion bits after the any-hit launch.
constexpr int kVisInline = 0, kVisEmit = 1, kVisFinish = 2;
// A pass that only finds the hybrid retraces temporal path reuse needs:
// visibility is taken as given and nothing but the requests is stored.
constexpr int kVisRequest = 3;

// is_vertex_visible (light.glsl:82-142) on the current TLAS.
template <bool kCount, int kVis = kVisInline>
CT_D bool vertex_visible(const DFrame &F, F3 o, F3 d, float dist, uint32_t ti, uint32_t tp, PtTrace &tr) {
  if (kVis == kVisRequest) return true;
  if (kVis == kVisInline) {
    float bt = 0.0f;
    uint32_t bi = kInvalid, bp = kInvalid;
    ++tr.shadow_rays;
    return !traverse_mode<kTraceAny, kCount>(F.S, o, d, dist, ti, tp, bt, bi, bp, tr.stk, kBlock, tr.cnt, tr.overflow);
  }
  const uint32_t k = tr.slot++;
  if (k >= kRaySlots) return true;  // at most two tests per pixel; never taken
  if (kVis == kVisEmit) {
    const uint32_t idx = tr.px * kRaySlots + k;
    F.ray_o[idx] = make_float4(o.x, o.y, o.z, dist);
    F.ray_d[idx] = make_float4(d.x, d.y, d.z, __uint_as_float(ti));
    F.ray_t[idx] = tp;
    tr.emitted |= 1u << k;
    if (dist >= F.far_ray) tr.far |= 1u << k;
    return true;
  }
  return ((tr.bits >> k) & 1u) == 0u;
}

// load_intersection_data_vertex_position, primitives.glsl:567-611
CT_D Isect load_vertex(const DFrame &F, const ObjHit &h, F3 world_pos) {
  F3 rd = sub(h.hit, world_pos);
  float t;
  F3 pos, nor;
  if (!hit_from_ray(F, h.inst, h.prim, world_pos, rd, t, pos, nor)) return isect_empty();
  pos = compute_ray_origin(pos, nor);
  pos = compute_ray_origin(pos, nor);
  Isect i;
  i.is_hit = true;
  i.distance = length(sub(world_pos, pos));
  i.P = pos;
  i.n = nor;
  i.wo = normalize(sub(world_pos, pos));
  i.wi = f3(0, 0, 0);
  i.inst = h.inst;
  i.prim = h.prim;
  i.mat_idx = material_of(F, h.inst, h.prim);
  i.mat = ld_mat(F, i.mat_idx);
  return i;
}
// compute_new_ray_origin, primitives.glsl:259-264
CT_D F3 new_ray_origin(F3 pos, F3 n, bool view_side) {
  F3 nn = view_side ? n : neg(n);
  return compute_ray_origin(compute_ray_origin(pos, nn), nn);
}

// ---------------------------------------------------------------- sampling
// path_generate_scatter_ray, path_tracer.glsl:59-112, with the Lambertian
// scatter callable (rcall_scatter.glsl:73-84).
CT_D bool pt_scatter(PState &p, Isect &i) {
  F3 sd = random_cosine_direction(p.seed, i.n);
  if (normal_near_zero(sd)) sd = i.n;
  i.wi = sd;
  F3 weight = evaluate_material(i.mat, i.n, i.wo, i.wi);
  const float pdf = material_pdf(i.n, i.wo, i.wi);
  weight = divs(weight, pdf);
  p.dir = i.wi;
  if (p.path_length == 0u) p.flags |= kPathDiffusePrimary;
  p.thp = mulv(p.thp, weight);
  p.pdf = pdf;
  if (p.path_length == p.pb.rc_len) {
    p.pb.cj.y = pdf;
    p.dirty |= kDirtyRc;
  }
  p.prev_scatter_pdf = pdf;
  if (pdf <= 0.0f) return false;
  p.bounce_counters += 1u;
  p.normal = i.n;
  const bool valid = p.thp.x > 0.0f || p.thp.y > 0.0f || p.thp.z > 0.0f;
  if (valid && p.path_length <= p.pb.rc_len) {
    p.prefix_thp = p.thp;  // path_record_prefix_thp
    p.thp = f3(1.0f, 1.0f, 1.0f);
  }
  return valid;
}

// path_builder_add_escape_vertex, path_builder.glsl:36-75
CT_D bool add_escape_vertex(PBuilder &pb, uint32_t plen, F3 wi, F3 path_weight, F3 postfix, float rr_pdf, float mis,
                            float light_pdf, uint32_t light_type, PRes &res) {
  if (plen < 1u) return false;
  const bool is_rc = plen == pb.rc_len;
  if (!pres_update(res, path_weight, rr_pdf, pb.seed)) return false;
  res.flags = pb.flags;
  pf_set_path_length(res.flags, plen);
  pf_set_rc_length(res.flags, pb.rc_len);
  pf_set_last_nee(res.flags, false);
  res.rc_irr = postfix;
  res.F = path_weight;
  res.rc_wi = pb.rc_wi;
  res.rc = pb.rc;
  res.cj = pb.cj;
  res.rc_seed = pb.cached_seed;
  if (is_rc) {
    res.rc_wi = wi;
    res.rc_irr = divs(res.rc_irr, mis);
  }
  res.light_pdf = light_pdf;
  pf_set_light_type(res.flags, light_type);
  return true;
}

// path_builder_mark_escape_vertex_as_rc_vertex, path_builder.glsl:114-139:
// the selected escape vertex becomes the reconnection vertex of the
// reservoir's sample (Lambertian: no delta, transmission or specular bits).
CT_D void mark_escape_rc(PRes &res, uint32_t length, const ObjHit &hit, float light_pdf, uint32_t light_type,
                         F3 irradiance, F3 wi, float prev_scatter_pdf, float geometry_factor) {
  pf_set_rc_length(res.flags, length);
  res.rc = hit;
  res.flags &= ~((1u << 20) | (1u << 22));
  res.light_pdf = light_pdf;
  res.cj.x = prev_scatter_pdf;
  pf_set_light_type(res.flags, light_type);
  res.rc_irr = irradiance;
  res.rc_wi = wi;
  res.cj.z = geometry_factor;
}
CT_D bool hybrid_mode(const DFrame &F) { return F.rp.pt_shift_mapping == CT_SHIFT_HYBRID; }

// eval_voxel_pdf, path_tracer.glsl:185-211
CT_D float eval_voxel_pdf(F3 pos, const Isect &i) {
  const F3 to_light = sub(i.P, pos);
  const float d2 = dot(to_light, to_light);
  if (d2 <= kHlfMin) return 0.0f;
  const F3 L = divs(to_light, sqrtf(d2));
  const float cos_theta = dot(i.n, neg(L));
  if (cos_theta <= 0.0f) return 0.0f;
  const float area = kVoxelExtent * kVoxelExtent;
  const float denom = gmax(kHlfMin, cos_theta * area);
  return d2 / denom;
}

// path_handle_emissive_hit, path_tracer.glsl:258-337 (not a primary hit,
// no replay, reconnection shift).
CT_D void pt_emissive(const DFrame &F, PState &p, const Isect &i) {
  const float light_pdf = (1.0f / (float)F.n_cube) * eval_voxel_pdf(p.origin, i);
  const float mis = eval_mis(p.pdf, light_pdf);
  const F3 cur = mulv(p.thp, p.prefix_thp);
  const F3 Lr = muls(mulv(cur, i.mat.emission), mis);
  const F3 att = mulv(cur, i.mat.emission);
  if (p.replay) return;  // a replay only carries throughput to its stop vertex
  if (p.path_length > 1u) p.L = add(p.L, Lr);
  p.dirty |= kDirtyRes;
  const F3 postfix = muls(mulv(p.thp, att), mis);
  const uint32_t plen = p.path_length >= 1u ? p.path_length - 1u : 0u;
  const bool selected = add_escape_vertex(p.pb, plen, p.dir, Lr, postfix, p.rr_pdf, mis, light_pdf, kGeomLightCube, p.res);
  // an escaped vertex at length >= 2 can be the reconnection vertex
  // (path_tracer.glsl:303-335)
  if (selected && hybrid_mode(F) && p.path_length >= 2u && p.path_length < p.pb.rc_len) {
    const bool far = length(sub(i.P, p.origin)) >= F.rp.near_field_distance;
    if (far && p.last_rough) {
      const F3 disp = sub(p.origin, i.P);
      float g = fabsf(dot(i.n, i.wo)) / dot(disp, disp);
      if (isnan(g)) g = 0.0f;
      mark_escape_rc(p.res, p.path_length, p.hit, light_pdf, kGeomLightCube, i.mat.emission, f3(0, 0, 0),
                     p.prev_scatter_pdf, g);
    }
  }
}

// path_generate_light_sample (path_tracer.glsl:222-256) as far as it is
// observable.  The sample's MIS weight is eval_mis(1, ls.pdf, ...) with
// ls.pdf never assigned (0), so every NEE contribution is 0 or NaN and
// path_reservoir_update rejects it without drawing; only the random numbers
// the light sample consumes reach the outputs.  Its shadow ray is not
// traced.
CT_D void pt_light_sample(const DFrame &F, const Isect &i, uint32_t &seed) {
  const uint32_t n = F.n_cube;
  const uint32_t li = min(urnd_interval(seed, 0u, n), n - 1u);
  if (n == 0u) return;
  const PLight l = ld_light(F.cube, li);
  Hit h;
  h.P = i.P;
  h.n = i.n;
  h.distance = i.distance;
  h.scatter = neg(i.wo);
  h.inst = i.inst;
  h.prim = i.prim;
  h.mat = i.mat_idx;
  float pdf = 1.0f / (float)n;
  LS ls;
  sample_lights(F, h, l, pdf, seed, true, false, ls);
}

// path_handle_hit, path_tracer.glsl:599-727.  The path builder follows
// the frame's shift mapping (use_hybrid_shift, which generate_path leaves
// false, path_state.glsl:245-247, is taken from it); a replay stops at the
// vertex before the source's reconnection vertex.
CT_D void pt_handle_hit(const DFrame &F, PState &p, Isect &i) {
  const ct_restir_params &rp = F.rp;
  const bool compute_emissive =
      !(p.path_length == 1u && (p.flags & kPathTransmission) == 0u && (p.flags & kPathDelta) == 0u);
  if (compute_emissive && (i.mat.emission.x > 0.0f || i.mat.emission.y > 0.0f || i.mat.emission.z > 0.0f))
    pt_emissive(F, p, i);
  // path_has_finished_surface_bounces, path_state.glsl:151-159
  if ((p.bounce_counters & 0xFFu) > p.max_depth) {
    p.flags &= ~kPathActive;
    return;
  }
  const F3 prev_origin = p.origin;
  p.origin = compute_ray_origin(compute_ray_origin(i.P, i.n), i.n);
  const bool cur_rough = i.mat.roughness > rp.roughness_threshold;  // classify_as_rough
  // path_check_rc_vertex, path_tracer.glsl:474-597
  const bool far = length(sub(p.origin, prev_origin)) >= rp.near_field_distance;
  const bool connectable = far && cur_rough && p.last_rough;
  if (!p.replay) {
    const bool hybrid = hybrid_mode(F);
    const bool can_connect = p.path_length >= 1u && p.path_length < p.pb.rc_len && connectable;
    if ((!hybrid && p.path_length == 1u) || (hybrid && can_connect)) {
      p.pb.rc = p.hit;
      p.dirty |= kDirtyRc;
      p.pb.cj.x = p.prev_scatter_pdf;
      const F3 disp = sub(prev_origin, i.P);
      p.pb.cj.z = fabsf(dot(i.n, i.wo)) / dot(disp, disp);
      if (isnan(p.pb.cj.z)) p.pb.cj.z = 0.0f;
      p.pb.rc_len = p.path_length;
    }
  } else {
    // The source path's reconnection vertex is its first connectable one, so
    // no replayed vertex before it may be connectable, and the one right
    // before it must be rough to connect from (invertibility, GRIS).
    bool invertible = !(p.path_length >= 1u && p.path_length + 1u <= p.replay_rc && connectable);
    if (invertible && p.path_length + 1u == p.replay_rc) {
      if (!cur_rough) {
        invertible = false;
      } else {
        p.replay_done = true;
        p.replay_prev = i;
        p.replay_tp = mulv(p.thp, p.prefix_thp);
      }
    }
    if (!invertible || p.replay_done) {
      p.flags &= ~kPathActive;
      return;
    }
  }
  p.flags &= ~(kPathLightUpper | kPathLightLower);
  // path_handle_sample_light, path_tracer.glsl:339-438
  if (p.path_length == p.pb.rc_len) p.pb.cached_seed = p.seed;
  pt_light_sample(F, i, p.seed);
  p.flags = (p.flags | kPathLightUpper) & ~kPathLightLower;
  // Russian roulette, terminate_path_by_russian_roulette (path_tracer.glsl:165-177);
  // a replay draws its number and goes on (:696-699)
  if (rp.pt_russian_roulette && p.path_length >= 1u && p.replay) {
    (void)rnd(p.seed);
  } else if (rp.pt_russian_roulette && p.path_length >= 1u) {
    const float rr = luminance(mulv(p.thp, p.prefix_thp));
    const float prob = gmax(0.0f, 1.0f - rr);
    if (rnd(p.seed) < prob) {
      p.flags &= ~kPathActive;
      return;
    }
    p.rr_pdf *= 1.0f - prob;
  }
  bool valid = pt_scatter(p, i);
  if (p.path_length == p.pb.rc_len) p.pb.cached_seed = p.seed;
  p.pb.rc_wi = p.dir;
  p.last_rough = cur_rough;
  if ((p.flags & kPathActive) == 0u) return;
  if ((p.bounce_counters & 0xFFu) > p.max_depth) valid = false;
  if (!valid) p.flags &= ~kPathActive;
}

// path_handle_miss, path_tracer.glsl:729-815 (environment light on)
CT_D void pt_handle_miss(const DFrame &F, PState &p) {
  if (p.replay) {  // a replay that escapes before its stop vertex is not invertible
    p.flags &= ~kPathActive;
    return;
  }
  const float light_pdf = kInv4Pi;
  const float mis = eval_mis(p.pdf, light_pdf);
  const F3 Le = sky_color(F.time, F.afternoon != 0, p.dir);
  const F3 Lr = muls(mulv(mulv(p.thp, p.prefix_thp), Le), mis);
  if (p.path_length > 0u) p.L = add(p.L, Lr);
  p.dirty |= kDirtyRes;
  const bool selected = add_escape_vertex(p.pb, p.path_length, p.dir, Lr, muls(mulv(p.thp, Le), mis), p.rr_pdf, mis,
                                          light_pdf, kGeomLightEnv, p.res);
  // escaped to the sky: a reconnection vertex at infinity (path_tracer.glsl:805-833)
  if (selected && hybrid_mode(F) && p.path_length >= 1u && p.path_length + 1u < p.pb.rc_len && p.last_rough)
    mark_escape_rc(p.res, p.path_length + 1u, objhit_invalid(), light_pdf, kGeomLightEnv, Le, p.dir,
                   p.prev_scatter_pdf, 1.0f);
  p.flags &= ~kPathActive;
}

// trace_restir_path_tracing (path_tracer.glsl:854-887) split at its
// path_next_vertex calls, so each path segment can be traced by the
// wavefront closest-hit kernel: pt_begin is generate_path
// (path_state.glsl:208-249) plus path_handle_primary_hit (:440-469);
// pt_step is one loop iteration after the segment's hit is known; pt_end
// is path_finalize + path_output (:817-835).  Returns whether the path
// traces another segment.
CT_D bool pt_begin(const DFrame &F, PState &p, Isect i, uint32_t seed) {
  p.max_depth = F.rp.max_depth;
  p.flags = kPathActive | kPathHit;
  p.path_length = 0u;
  p.bounce_counters = 0u;
  p.scene_length = 0.0f;
  p.origin = f3(0, 0, 0);
  p.dir = f3(0, 0, 0);
  p.pdf = 0.0f;
  p.normal = f3(0, 0, 0);
  p.hit.inst = i.inst;
  p.hit.prim = i.prim;
  p.hit.hit = i.P;
  p.thp = f3(1.0f, 1.0f, 1.0f);
  p.prefix_thp = f3(1.0f, 1.0f, 1.0f);
  p.dirty = 0u;
  p.L = f3(0, 0, 0);
  p.rr_pdf = 1.0f;
  p.prev_scatter_pdf = 1.0f;
  p.seed = seed;
  p.pb.seed = seed;
  p.pb.rc = objhit_invalid();
  p.pb.rc_wi = f3(0, 0, 0);
  p.pb.cached_seed = 0u;
  p.pb.rc_len = F.rp.max_depth;
  p.pb.flags = 0u;
  p.pb.cj = f3(0, 0, 0);
  p.res = pres_init(F.rp.max_depth);
  p.last_rough = true;
  p.replay = false;
  p.replay_done = false;
  p.replay_rc = 0u;
  p.origin = compute_ray_origin(compute_ray_origin(i.P, i.n), i.n);
  p.flags &= ~(kPathLightUpper | kPathLightLower);
  p.res.init_seed = p.seed;
  const bool valid = pt_scatter(p, i);
  if (p.path_length == p.pb.rc_len) p.pb.cached_seed = p.seed;
  p.pb.rc_wi = p.dir;
  return valid;
}

// The INTERSECT record intersect() (bounce.glsl:9-130) builds once the
// segment from `origin` (boxes tested from `ro` = origin pushed along dir)
// has its closest voxel (inst = kInvalid on a miss).
CT_D Isect segment_record(const DFrame &F, F3 origin, F3 ro, F3 dir, uint32_t inst, uint32_t prim) {
  Isect i = isect_empty();
  i.mat_idx = 0u;
  F3 int_hit = f3(0, 0, 0);
  if (inst != kInvalid) {
    float t;
    F3 pos, nrm;
    hit_from_ray(F, inst, prim, ro, dir, t, pos, nrm);
    int_hit = pos;
    i.is_hit = true;
    i.P = pos;
    i.n = nrm;
    i.inst = inst;
    i.prim = prim;
    i.mat_idx = material_of(F, inst, prim);
    i.mat = ld_mat(F, i.mat_idx);
    i.distance = length(sub(origin, int_hit));
  }
  i.wo = normalize(sub(origin, int_hit));
  return i;
}

// One iteration of the path loop after path_next_vertex found `i`.
CT_D bool pt_step(const DFrame &F, PState &p, Isect &i) {
  if (i.is_hit) {
    p.hit.inst = i.inst;
    p.hit.prim = i.prim;
    p.hit.hit = i.P;
    p.flags |= kPathHit;
    p.path_length++;
    p.scene_length += i.distance;
    pt_handle_hit(F, p, i);
  } else {
    p.flags &= ~kPathHit;
    p.scene_length = kHlfMax;
    pt_handle_miss(F, p);
  }
  return (p.flags & kPathActive) != 0u;
}

// Random replay (trace_random_replay_path_hybrid_simple,
// path_tracer.glsl:889-966; generate_random_replay_path): the path a seed
// draws from a given primary intersection, through the same sampling steps
// as the frame's paths (pt_begin, then pt_step per segment with the
// segment's closest voxel), traced in place.  It reports the vertex at
// index rc_len (the reconnection vertex; index 0 is the primary hit), the
// vertex before it with its incoming direction and the throughput leaving
// it (after its scatter), and the length reached.  Replaying a path's own seed from its own
// primary hit reproduces the path; a replay that stops before rc_len is
// not invertible (ok = false, zero throughput).
struct Replay {
  uint32_t length;  // path length reached
  bool ok;          // the vertex at rc_len exists
  ObjHit rc, rc_prev;
  F3 rc_prev_wo, thp;
};

CT_D Replay pt_replay(const DFrame &F, Isect i0, uint32_t seed, uint32_t rc_len, PtTrace &tr) {
  Replay r;
  r.ok = false;
  r.rc = objhit_invalid();
  r.rc_prev = objhit_invalid();
  r.rc_prev_wo = f3(0, 0, 0);
  r.thp = f3(0, 0, 0);
  PState p;
  bool more = pt_begin(F, p, i0, seed);
  F3 wo = i0.wo;
  if (rc_len == 1u) {
    r.rc_prev = p.hit;
    r.rc_prev_wo = wo;
    r.thp = p.thp;
  }
  while (more && p.path_length < rc_len) {
    Isect i = bounce_intersect<false>(F, p.origin, p.dir, tr);
    wo = i.wo;
    more = pt_step(F, p, i);
    if (!i.is_hit) break;
    if (p.path_length + 1u == rc_len) {
      r.rc_prev = p.hit;
      r.rc_prev_wo = wo;
      r.thp = p.thp;
    }
  }
  r.length = p.path_length;
  if (p.path_length == rc_len && (p.flags & kPathHit)) {
    r.ok = true;
    r.rc = p.hit;
  } else {
    r.rc_prev = objhit_invalid();
    r.rc_prev_wo = f3(0, 0, 0);
    r.thp = f3(0, 0, 0);
  }
  return r;
}

CT_D void pt_end(PState &p, F3 &color) {
  p.res.M = 1.0f;  // path_builder_finilize
  pres_finalize_ris(p.res);
  color = muls(p.res.F, p.res.weight);
  if (has_nan_inf(color)) color = f3(0, 0, 0);
}

// Path state between the wavefront kernels: 14 float4 per path, stored
// field-major (field k of path p at S[k * np + p]).  Fields the outputs
// never read (L, normal, scene_length, hit) are not kept.
constexpr int kPathStateVec = 14;
// Stores the state; with `all` false only the vectors a path-loop step
// changes: the seeds, flags, ray, throughputs and the builder's latest
// direction always, the reconnection vertex and the reservoir when marked
// dirty.
CT_D void st_pstate(float4 *S, uint32_t np, uint32_t p, const PState &s, bool all = true) {
  const uint32_t packed = (s.path_length & 0xFFu) | ((s.bounce_counters & 0xFFu) << 8) | ((s.pb.rc_len & 0xFFu) << 16) |
                          ((s.last_rough ? 1u : 0u) << 24) | ((s.max_depth & 0x7Fu) << 25);
  S[0 * (size_t)np + p] = make_float4(__uint_as_float(s.seed), __uint_as_float(s.pb.seed), __uint_as_float(s.flags),
                                      __uint_as_float(packed));
  S[1 * (size_t)np + p] = make_float4(s.origin.x, s.origin.y, s.origin.z, s.pdf);
  S[2 * (size_t)np + p] = make_float4(s.dir.x, s.dir.y, s.dir.z, s.prev_scatter_pdf);
  S[3 * (size_t)np + p] = make_float4(s.thp.x, s.thp.y, s.thp.z, s.rr_pdf);
  S[4 * (size_t)np + p] = make_float4(s.prefix_thp.x, s.prefix_thp.y, s.prefix_thp.z, __uint_as_float(s.pb.cached_seed));
  if (all || (s.dirty & kDirtyRc))
    S[5 * (size_t)np + p] = make_float4(__uint_as_float(s.pb.rc.inst), __uint_as_float(s.pb.rc.prim), s.pb.rc.hit.x,
                                        s.pb.rc.hit.y);
  S[6 * (size_t)np + p] = make_float4(s.pb.rc.hit.z, s.pb.rc_wi.x, s.pb.rc_wi.y, s.pb.rc_wi.z);
  if (all || (s.dirty & kDirtyRc))
    S[7 * (size_t)np + p] = make_float4(s.pb.cj.x, s.pb.cj.y, s.pb.cj.z, __uint_as_float(s.pb.flags));
  if (!all && !(s.dirty & kDirtyRes)) return;
  const PRes &r = s.res;
  S[8 * (size_t)np + p] = make_float4(r.M, r.weight, __uint_as_float(r.flags), __uint_as_float(r.rc_seed));
  S[9 * (size_t)np + p] = make_float4(r.F.x, r.F.y, r.F.z, r.light_pdf);
  S[10 * (size_t)np + p] = make_float4(r.cj.x, r.cj.y, r.cj.z, __uint_as_float(r.init_seed));
  S[11 * (size_t)np + p] = make_float4(__uint_as_float(r.rc.inst), __uint_as_float(r.rc.prim), r.rc.hit.x, r.rc.hit.y);
  S[12 * (size_t)np + p] = make_float4(r.rc.hit.z, r.rc_wi.x, r.rc_wi.y, r.rc_wi.z);
  S[13 * (size_t)np + p] = make_float4(r.rc_irr.x, r.rc_irr.y, r.rc_irr.z, 0.0f);
}
// The stored path reservoir (vectors 8..13 of the state).
CT_D void ld_pstate_res(const float4 *S, uint32_t np, uint32_t p, PState &s) {
  float4 q[kPathStateVec];
#pragma unroll
  for (int k = 8; k < kPathStateVec; ++k) q[k] = S[(size_t)k * np + p];
  PRes &r = s.res;
  r.M = q[8].x; r.weight = q[8].y; r.flags = __float_as_uint(q[8].z); r.rc_seed = __float_as_uint(q[8].w);
  r.F = f3(q[9].x, q[9].y, q[9].z); r.light_pdf = q[9].w;
  r.cj = f3(q[10].x, q[10].y, q[10].z); r.init_seed = __float_as_uint(q[10].w);
  r.rc.inst = __float_as_uint(q[11].x); r.rc.prim = __float_as_uint(q[11].y); r.rc.hit = f3(q[11].z, q[11].w, q[12].x);
  r.rc_wi = f3(q[12].y, q[12].z, q[12].w);
  r.rc_irr = f3(q[13].x, q[13].y, q[13].z);
}
// Loads the state; with `res` false all but the path reservoir, which a
// caller that needs it then takes with ld_pstate_res.
CT_D void ld_pstate(const float4 *S, uint32_t np, uint32_t p, PState &s, bool res = true) {
  float4 q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = S[(size_t)k * np + p];
  s.seed = __float_as_uint(q[0].x);
  s.pb.seed = __float_as_uint(q[0].y);
  s.flags = __float_as_uint(q[0].z);
  const uint32_t packed = __float_as_uint(q[0].w);
  s.path_length = packed & 0xFFu;
  s.bounce_counters = (packed >> 8) & 0xFFu;
  s.pb.rc_len = (packed >> 16) & 0xFFu;
  s.last_rough = ((packed >> 24) & 1u) != 0u;
  s.max_depth = (packed >> 25) & 0x7Fu;
  s.origin = f3(q[1].x, q[1].y, q[1].z); s.pdf = q[1].w;
  s.dir = f3(q[2].x, q[2].y, q[2].z); s.prev_scatter_pdf = q[2].w;
  s.thp = f3(q[3].x, q[3].y, q[3].z); s.rr_pdf = q[3].w;
  s.prefix_thp = f3(q[4].x, q[4].y, q[4].z); s.pb.cached_seed = __float_as_uint(q[4].w);
  s.pb.rc.inst = __float_as_uint(q[5].x); s.pb.rc.prim = __float_as_uint(q[5].y);
  s.pb.rc.hit = f3(q[5].z, q[5].w, q[6].x);
  s.pb.rc_wi = f3(q[6].y, q[6].z, q[6].w);
  s.pb.cj = f3(q[7].x, q[7].y, q[7].z); s.pb.flags = __float_as_uint(q[7].w);
  if (res) ld_pstate_res(S, np, p, s);
  s.scene_length = 0.0f;
  s.normal = f3(0, 0, 0);
  s.L = f3(0, 0, 0);
  s.dirty = 0u;
  s.hit = objhit_invalid();
  s.replay = false;
  s.replay_done = false;
  s.replay_rc = 0u;
}

// compute_shifted_integrand_reconnection, shift.glsl:25-350.  `dst` and
// `src` are the vertices the connection leaves from (the primary ones for
// the reconnection shift); `s` the source reservoir.  With use_hybrid the
// reconnection vertex index comes from the path flags and a near-field
// connection is rejected; with use_cached the source side's geometry term
// and pdfs come from the reservoir's cached Jacobian, and a successful
// shift leaves the destination's there.
template <bool kCount, int kVis = kVisInline>
CT_D F3 shift_reconnection(const DFrame &F, float &dst_jac, Isect dst, Isect src, PRes &s, bool eval_vis,
                           PtTrace &tr, bool use_hybrid = false, bool use_cached = false) {
  dst_jac = 0.0f;
  F3 dst_cj = f3(0, 0, 0);
  const uint32_t rc_len = use_hybrid ? pf_rc_length(s.flags) : 1u;
  const F3 irr = s.rc_irr, wi = s.rc_wi;
  const bool exists = instance_hit_exists(s.rc.inst, s.rc.prim);
  const uint32_t plen = pf_path_length(s.flags);
  const bool nee = pf_last_nee(s.flags);
  const bool trans = pf_transmission(s.flags, true);
  src.P = new_ray_origin(src.P, src.n, !trans);
  dst.P = new_ray_origin(dst.P, dst.n, !trans);
  if (!exists) {
    F3 integrand = f3(0, 0, 0);
    if (pf_light_type(s.flags) == kGeomLightEnv && plen + 1u == rc_len && !nee) {
      if (vertex_visible<kCount, kVis>(F, dst.P, wi, kMaxDistance, kMaxInstances, kMaxPrimitives, tr)) {
        const float src_pdf1 = use_cached ? s.cj.x : material_pdf(src.n, src.wo, wi);
        const float dst_pdf1 = material_pdf(dst.n, dst.wo, wi);
        dst_cj.x = dst_pdf1;
        const F3 dst_f1 = evaluate_material(dst.mat, dst.n, dst.wo, wi);
        const float mis = eval_mis(dst_pdf1, s.light_pdf);
        integrand = mulv(muls(divs(dst_f1, dst_pdf1), mis), irr);
        dst_jac = dst_pdf1 / src_pdf1;
      }
    }
    if (use_cached) s.cj = dst_cj;
    if (jacobian_invalid(dst_jac)) dst_jac = 0.0f;
    if (has_nan_inf(integrand)) return f3(0, 0, 0);
    return integrand;
  }
  const bool is_final = plen == rc_len;
  const bool escaped = plen + 1u == rc_len && !nee;
  const bool rc_nee = is_final && nee;
  if (pf_delta(s.flags, true) || pf_delta(s.flags, false)) return f3(0, 0, 0);
  const Isect rc = load_vertex(F, s.rc, dst.P);
  const F3 dst_cv = neg(rc.wo);
  const F3 src_cv = normalize(sub(rc.P, src.P));
  const F3 sdisp = sub(rc.P, dst.P);
  const float sd2 = dot(sdisp, sdisp);
  const float scos = fabsf(dot(rc.n, neg(dst_cv)));
  if (use_hybrid && !(sqrtf(sd2) >= F.rp.near_field_distance)) return f3(0, 0, 0);
  const float cjz = scos / sd2;
  dst_cj.z = cjz;
  float jac;
  if (use_cached) {
    jac = cjz / s.cj.z;
  } else {
    const F3 odisp = sub(rc.P, src.P);
    const float od2 = dot(odisp, odisp);
    const float ocos = fabsf(dot(rc.n, neg(src_cv)));
    jac = cjz * od2 / ocos;
  }
  if (jacobian_invalid(jac)) return f3(0, 0, 0);
  const float dst_pdf1 = material_pdf(dst.n, dst.wo, dst_cv);
  dst_cj.x = dst_pdf1;
  const float src_pdf1 = use_cached ? s.cj.x : material_pdf(src.n, src.wo, src_cv);
  jac *= dst_pdf1 / src_pdf1;
  if (jacobian_invalid(jac)) return f3(0, 0, 0);
  const F3 dst_f1 = evaluate_material(dst.mat, dst.n, dst.wo, dst_cv);
  float dst_rc_all = 0.0f, dst_pdf2 = 1.0f, dst_rc = 1.0f, src_rc = 1.0f;
  if (!escaped) {
    dst_rc = material_pdf(rc.n, rc.wo, wi);
    dst_rc_all = dst_rc;
    dst_cj.y = dst_rc;
    src_rc = use_cached ? s.cj.y : material_pdf(rc.n, neg(src_cv), wi);
    dst_pdf2 = !rc_nee ? dst_rc : s.light_pdf;
  }
  F3 dst_f2 = f3(1.0f, 1.0f, 1.0f);
  if (!escaped) dst_f2 = evaluate_material(rc.mat, rc.n, rc.wo, wi);
  if ((dst_f1.x == 0.0f && dst_f1.y == 0.0f && dst_f1.z == 0.0f) ||
      (dst_f2.x == 0.0f && dst_f2.y == 0.0f && dst_f2.z == 0.0f))
    return f3(0, 0, 0);
  F3 no_f1 = mulv(divs(dst_f2, dst_pdf2), irr);
  F3 integrand = mulv(divs(dst_f1, dst_pdf1), no_f1);
  if (escaped) integrand = muls(integrand, eval_mis(dst_pdf1, s.light_pdf));
  if (is_final && pf_light_type(s.flags) != kGeomLightAnalytic) {
    const float lp = s.light_pdf;
    const float mis = eval_mis(rc_nee ? lp : dst_rc_all, rc_nee ? dst_rc_all : lp);
    integrand = muls(integrand, mis);
    no_f1 = muls(no_f1, mis);
    if (!rc_nee) jac *= dst_rc / src_rc;
  }
  if (!is_final && !escaped) jac *= dst_rc / src_rc;
  if (jacobian_invalid(jac)) return f3(0, 0, 0);
  if (eval_vis) {
    const F3 dir = normalize(sub(rc.P, dst.P));
    const float dist = length(sub(dst.P, rc.P));
    if (!vertex_visible<kCount, kVis>(F, dst.P, dir, dist, rc.inst, rc.prim, tr)) return f3(0, 0, 0);
  }
  if (has_nan_inf(integrand)) return f3(0, 0, 0);
  dst_jac = jac;
  if (use_cached) s.cj = dst_cj;
  return integrand;
}

// RECONNECTION_DATA (shared.inl:380-394) as the retrace produces it: the
// destination path's vertex before the reconnection vertex and the
// throughput that reaches it.
struct RcData {
  bool ok;
  Isect prev;
  F3 tp;
};

// The frame's retrace records: 5 float4 per pixel and slot.  A request pass
// writes (DI index of the destination, the source's init seed and flags)
// into the first; the retrace pass overwrites the record with its result.
constexpr uint32_t kRcVec = 5;
CT_D float4 *rc_rec(const DFrame &F, uint32_t px, uint32_t slot) { return F.rc + (size_t)kRcVec * (2u * px + slot); }
CT_D void st_rc_request(const DFrame &F, uint32_t px, uint32_t slot, uint32_t di_index, const PRes &src) {
  rc_rec(F, px, slot)[0] = make_float4(__uint_as_float(di_index), __uint_as_float(src.init_seed),
                                       __uint_as_float(src.flags), 0.0f);
}
CT_D void st_rc(const DFrame &F, uint32_t px, uint32_t slot, const RcData &r) {
  float4 *o = rc_rec(F, px, slot);
  const Isect &i = r.prev;
  o[0] = make_float4(i.P.x, i.P.y, i.P.z, i.distance);
  o[1] = make_float4(i.n.x, i.n.y, i.n.z, i.wo.x);
  o[2] = make_float4(i.wo.y, i.wo.z, i.wi.x, i.wi.y);
  o[3] = make_float4(i.wi.z, __uint_as_float(i.inst), __uint_as_float(i.prim), __uint_as_float(i.mat_idx));
  o[4] = make_float4(r.tp.x, r.tp.y, r.tp.z, __uint_as_float((r.ok ? 1u : 0u) | (i.is_hit ? 2u : 0u)));
}
CT_D RcData ld_rc(const DFrame &F, uint32_t px, uint32_t slot) {
  const float4 *o = rc_rec(F, px, slot);
  const float4 a = o[0], b = o[1], c = o[2], d = o[3], e = o[4];
  RcData r;
  const uint32_t fl = __float_as_uint(e.w);
  r.ok = (fl & 1u) != 0u;
  r.tp = f3(e.x, e.y, e.z);
  Isect &i = r.prev;
  i.is_hit = (fl & 2u) != 0u;
  i.P = f3(a.x, a.y, a.z);
  i.distance = a.w;
  i.n = f3(b.x, b.y, b.z);
  i.wo = f3(b.w, c.x, c.y);
  i.wi = f3(c.z, c.w, d.x);
  i.inst = __float_as_uint(d.y);
  i.prim = __float_as_uint(d.z);
  i.mat_idx = __float_as_uint(d.w);
  i.mat = ld_mat(F, i.mat_idx);
  return r;
}

// The retrace of trace_hybrid_shift_rays / trace_random_replay_path_hybrid_simple
// (indirect_illumination.glsl:12-28, path_tracer.glsl:889-966): replay the
// source reservoir's random numbers from the destination's primary
// intersection through the frame's own sampling steps (pt_begin, pt_step
// over in-place closest hits) until the vertex before the source's
// reconnection vertex.  Not invertible (ok = false) if it misses, stops
// early, reaches a connectable vertex before that one or that vertex is
// not rough.
CT_D RcData hybrid_retrace(const DFrame &F, Isect i0, const PRes &src, PtTrace &tr) {
  RcData out;
  out.ok = false;
  out.tp = f3(0, 0, 0);
  PState p;
  const uint32_t k = pf_rc_length(src.flags);
  // pt_begin's primary scatter with the path builder's reconnection length
  // at k instead of max_depth: both exceed 0, so the scatter is the same
  bool more = pt_begin(F, p, i0, src.init_seed);
  p.pb.rc_len = k;
  p.replay = true;
  p.replay_rc = k;
  p.last_rough = i0.mat.roughness > F.rp.roughness_threshold;  // generate_replay_primary_scatter_ray
  while (more) {
    Isect h = bounce_intersect<false>(F, p.origin, p.dir, tr);
    more = pt_step(F, p, h);
  }
  if (p.replay_done) {
    out.ok = true;
    out.prev = p.replay_prev;
    out.tp = p.replay_tp;
  }
  return out;
}

// compute_shifted_integrand_ (shift.glsl:446-483), reconnection or hybrid
// (the hybrid shift only in the kernels built with kHybrid).  `s` is the
// caller's copy of the source reservoir: the hybrid shift leaves the
// destination's cached Jacobian in it.
// The hybrid shift, compute_shifted_integrand_hybrid (shift.glsl:352-444):
// random replay up to the vertex before the reconnection vertex, then the
// reconnection from there with the cached Jacobian (or from the
// destination's primary vertex when the reconnection vertex is the first
// bounce).  Two points differ from the text, which as written rejects every
// shift: the Jacobian starts at 1 (the replayed prefix has unit Jacobian in
// primary-sample space) where the text sets 0 and multiplies into it (:363,
// :437), and reconnection length 1 reconnects from the destination's
// primary vertex where the text passes an invalid hit (:376-378).  A source
// whose reconnection vertex does not exist (the path ended first) is not
// shifted.
template <bool kCount, int kVis = kVisInline, bool kHybrid = false>
CT_D F3 shifted_integrand(const DFrame &F, float &dst_jac, const Isect &dst, const Isect &src, PRes &s,
                          bool eval_vis, PtTrace &tr) {
  dst_jac = 0.0f;
  if (s.weight == 0.0f) return f3(0, 0, 0);
  const uint32_t mapping = F.rp.pt_shift_mapping;
  if (!kHybrid) {
    if (mapping == CT_SHIFT_RECONNECTION) return shift_reconnection<kCount, kVis>(F, dst_jac, dst, src, s, eval_vis, tr);
    return f3(1.0f, 1.0f, 1.0f);
  }
  // The hybrid build reaches one reconnection for both mappings (it serves
  // frame sets that mix them): the hybrid shift's checks and retrace first,
  // then the reconnection from the retraced vertex or from dst.
  const bool hyb = mapping == CT_SHIFT_HYBRID;
  if (!hyb && mapping != CT_SHIFT_RECONNECTION) return f3(1.0f, 1.0f, 1.0f);
  F3 tp = f3(1.0f, 1.0f, 1.0f);
  RcData rc;
  bool retraced = false;
  if (hyb) {
    const uint32_t k = pf_rc_length(s.flags), plen = pf_path_length(s.flags);
    const bool escaped = plen + 1u == k;
    if (!(k <= plen || escaped) || k == 0u) return f3(0, 0, 0);
    if (k > 1u) {
      // Temporal path reuse (the passes around the any-hit launch) takes the
      // retrace from the frame's records, filled by a request pass and the
      // retrace kernel in between; spatial path reuse retraces in place.
      if (kVis == kVisRequest) {
        st_rc_request(F, tr.px, tr.rc_slot, tr.rc_slot == 0u ? tr.px : tr.rc_prev, s);
        tr.rc_req |= 1u << tr.rc_slot;
        return f3(0, 0, 0);
      }
      rc = kVis == kVisInline ? hybrid_retrace(F, dst, s, tr) : ld_rc(F, tr.px, tr.rc_slot);
      if (!rc.ok) return f3(0, 0, 0);
      tp = rc.tp;
      retraced = true;
    }
    if (!(tp.x > 0.0f || tp.y > 0.0f || tp.z > 0.0f)) return f3(0, 0, 0);
  }
  float rj = 0.0f;
  const F3 r = shift_reconnection<kCount, kVis>(F, rj, retraced ? rc.prev : dst, src, s, eval_vis, tr, hyb, retraced);
  dst_jac = rj;  // the hybrid shift's 1.0f * rj
  return hyb ? mulv(tp, r) : r;
}
// compute_shifted_integrand, shift.glsl:485-497: on a copy of the reservoir.
template <bool kCount, int kVis = kVisInline, bool kHybrid = false>
CT_D F3 shifted_integrand_copy(const DFrame &F, float &dst_jac, const Isect &dst, const Isect &src, const PRes &s,
                               bool eval_vis, PtTrace &tr) {
  PRes t = s;
  return shifted_integrand<kCount, kVis, kHybrid>(F, dst_jac, dst, src, t, eval_vis, tr);
}

// shift_and_merge_reservoir, shift.glsl:485-510 (force_merge, mis 1)
template <bool kCount, int kVis = kVisInline, bool kHybrid = false>
CT_D bool shift_and_merge(const DFrame &F, float &dst_jac, const Isect &dst, PRes &dst_res, const Isect &src,
                          const PRes &src_res, uint32_t &seed, PtTrace &tr) {
  PRes tmp = src_res;
  const F3 integrand = shifted_integrand<kCount, kVis, kHybrid>(F, dst_jac, dst, src, tmp, true, tr);
  const bool selected = pres_merge(dst_res, integrand, dst_jac, tmp, seed, 1.0f, true);
  if (!selected) dst_res.F = f3(0, 0, 0);
  dst_res.M = src_res.M;
  dst_res.weight = src_res.weight;
  return selected;
}

// temporal_path_reuse, indirect_illumination.glsl:147-279 (Talbot RMIS).
// Its two visibility tests (the central path shifted to the previous pixel,
// then the history shifted here) exist or not by data alone, never by an
// earlier test's outcome, so a queued pass records the same rays in the
// same slots as the finishing pass reads.
template <bool kCount, int kVis = kVisInline, bool kHybrid = false>
CT_D PRes pt_temporal(const DFrame &F, const PRes &central, PRes temporal, const Isect &cur_i, const Isect &prev_i,
                      uint32_t &seed, PtTrace &tr) {
  PRes dest = pres_init(F.rp.max_depth);
  const float current_M = central.M;
  temporal.M = gmin(F.rp.pt_m_cap * current_M, temporal.M);
  float dst_jac = 0.0f;
  for (int k = -1; k <= 0; ++k) {
    float p_sum = 0.0f, p_self = 0.0f;
    PRes tdst = dest;
    bool possible;
    if (k == -1) {
      tdst = central;
      dst_jac = 1.0f;
      possible = tdst.weight > 0.0f;
    } else {
      tr.rc_slot = 0u;  // the history shifted to this pixel's primary hit
      possible = shift_and_merge<kCount, kVis, kHybrid>(F, dst_jac, cur_i, tdst, prev_i, temporal, seed, tr);
    }
    if (possible) {
      // j = -1
      const float cur_p = lum601(tdst.F) * current_M;
      p_sum += cur_p;
      if (k == -1) p_self = cur_p;
      // j = 0
      if (k == 0) {
        p_self = lum601(temporal.F) / dst_jac * temporal.M;
        p_sum += p_self;
      } else {
        float tj = 0.0f;
        tr.rc_slot = 1u;  // this pixel's path shifted to the previous pixel's primary hit
        const F3 ti = shifted_integrand_copy<kCount, kVis, kHybrid>(F, tj, prev_i, cur_i, tdst, true, tr);
        const float pp = lum601(ti) * tj;
        p_sum += pp * temporal.M;
      }
    }
    const float mis = p_sum == 0.0f ? 0.0f : p_self / p_sum;
    pres_merge_mis(dest, tdst.F, dst_jac, tdst, seed, mis);
  }
  if (dest.weight > 0.0f) pres_finalize_gris(dest);
  if (dest.weight < 0.0f || isinf(dest.weight) || isnan(dest.weight)) dest.weight = 0.0f;
  return dest;
}

// get_next_neighbor_pixel, prng.glsl:272-296
CT_D void pt_neighbor(int32_t px, int32_t py, int32_t radius, uint32_t &seed, int32_t &nx, int32_t &ny) {
  const float ox = 2.0f * rnd(seed) - 1.0f;
  const float oy = 2.0f * rnd(seed) - 1.0f;
  nx = px + f2i_sat(ox * (float)radius);
  ny = py + f2i_sat(oy * (float)radius);
  if (nx == px && ny == py) {
    nx = -1;
    ny = -1;
  }
}

// The primary intersection of a pixel from a DI record
// (temporal_path_get_reprojected_primary_hit / path_get_neighbor_primary_hit,
// indirect_illumination.glsl:59-107): wo is not normalized there.
CT_D Isect isect_from_di_raw(const DFrame &F, const DI &d) {
  Isect i;
  i.is_hit = d.distance > 0.0f;
  i.distance = d.distance;
  i.P = d.pos;
  i.n = d.nrm;
  i.wo = sub(d.org, d.pos);
  i.wi = f3(0, 0, 0);
  i.inst = d.inst;
  i.prim = d.prim;
  i.mat_idx = d.mat;
  i.mat = ld_mat(F, d.mat);
  return i;
}
// The primary intersection rgen.glsl builds (rgen.glsl:275-278, 431-434).
CT_D Isect isect_from_di(const DFrame &F, const DI &d) {
  Isect i = isect_from_di_raw(F, d);
  i.wo = normalize(sub(d.org, d.pos));
  return i;
}

// indirect_illumination_spatial_reuse, indirect_illumination.glsl:312-452
template <bool kCount, bool kHybrid = false>
CT_D PRes pt_spatial(const DFrame &F, uint32_t x, uint32_t y, const Isect &ci, F3 camera_pos, uint32_t &seed,
                     PtTrace &tr) {
  const uint32_t cur = y * F.W + x;
  const PRes central = ld_pres(F.pres_o, cur);
  PRes dest = pres_init(F.rp.max_depth);
  const int32_t ncount = (int32_t)F.rp.pt_neighbor_count;
  int32_t valid = 0;
  float canonical = 1.0f;
  for (int32_t n = 0; n < ncount; ++n) {
    int32_t nx, ny;
    pt_neighbor((int32_t)x, (int32_t)y, F.rp.pt_neighbor_radius, seed, nx, ny);
    if (!(nx >= 0 && ny >= 0 && nx < (int32_t)F.W && ny < (int32_t)F.H)) continue;  // is_valid_screen_region
    const uint32_t ni = (uint32_t)ny * F.W + (uint32_t)nx;
    const DI nd = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, ni);
    const Isect nis = isect_from_di_raw(F, nd);
    if (!instance_hit_exists(nis.inst, nis.prim)) continue;
    {  // is_valid_geometry, primitives.glsl:547-556
      const float cd = length(sub(camera_pos, ci.P));
      const float nd2 = length(sub(camera_pos, nis.P));
      if (!(dot(ci.n, nis.n) >= 0.5f && fabsf(cd - nd2) < 0.1f * cd)) continue;
    }
    const PRes nres = ld_pres(F.pres_o, ni);
    valid++;
    float prefix_jac = 0.0f;
    const F3 prefix = shifted_integrand_copy<kCount, kVisInline, kHybrid>(F, prefix_jac, nis, ci, central, true, tr);
    const float prefix_pdf = lum601(prefix) * prefix_jac;
    canonical += 1.0f;
    if (prefix_pdf > 0.0f)
      canonical -= prefix_pdf * nres.M /
                   (prefix_pdf * nres.M + central.M * lum601(central.F) / (float)ncount);
    PRes tdst = dest;
    float dst_jac = 0.0f;
    const bool possible = shift_and_merge<kCount, kVisInline, kHybrid>(F, dst_jac, ci, tdst, nis, nres, seed, tr);
    float nw = 0.0f;
    if (possible) {
      nw = lum601(nres.F) / dst_jac * nres.M /
           ((lum601(nres.F) / dst_jac) * nres.M + lum601(tdst.F) * central.M / (float)ncount);
      if (isnan(nw) || isinf(nw)) nw = 0.0f;
    }
    pres_merge_mis(dest, tdst.F, dst_jac, tdst, seed, nw);
  }
  pres_merge_mis(dest, central.F, 1.0f, central, seed, canonical);
  if (dest.weight > 0.0f) {
    pres_finalize_gris(dest);
    dest.weight /= (float)(valid + 1);
  }
  if (dest.weight < 0.0f) dest.weight = 0.0f;
  if (isnan(dest.weight) || isinf(dest.weight)) dest.weight = 0.0f;
  return dest;
}
