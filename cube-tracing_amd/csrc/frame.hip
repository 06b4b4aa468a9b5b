// frame.hip -- the per-frame passes App::draw records (src/main.cpp:1436-1703):
// the prepass (rgen.glsl:46-320), the shading pass (rgen.glsl:329-510) and
// TAA (taa.glsl:16-108), on structure-of-arrays buffers in HBM.
//
// Wavefront form.  Every pass that traces shadow rays runs its per-pixel
// logic twice: an EMIT launch writes each needed shadow ray to a fixed slot
// (pixel * 16 + k) and appends the slot to a compacted queue, the any-hit
// traversal kernel resolves the queue into per-pixel occlusion bits, and a
// FINISH launch re-runs the same code reading the bits.  Both launches
// evaluate the same deterministic sequence of light samples, so the k-th
// ray of a pixel means the same query in both.  Queries of one pass never
// depend on each other's outcome (the reference code issues them before
// any result-dependent branch), which makes the split exact.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "ct_device.h"
#include "ct_internal.h"

namespace ct {
using namespace dev;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxInstances = CT_MAX_INSTANCES;
constexpr uint32_t kMaxPrimitives = CT_MAX_PRIMITIVES;
constexpr float kInvPi = 0.31830988618379067153776752674503f;
constexpr float kInv4Pi = 0.079577471545947667884441881686255f;
constexpr float kMaxDistance = 1e9f;
constexpr float kMinCosTheta = 1e-6f;

// Event marks of a frame (Ctx::ev indices).
enum {
  kMarkStart, kMarkPrimary, kMarkRis, kMarkTraceRis, kMarkTemporalFinish,
  kMarkPtEmit, kMarkTracePt, kMarkPt, kMarkSpatialEmit, kMarkTraceSpatial, kMarkSpatialFinish, kMarkPtSpatial, kMarkTaa,
  kMarkCount
};

// counters[] slots
// kCntQueue counts the far rays at the front of the ray queue, kCntQueueNear
// the near rays stored from its back end down (see queue_flush).
constexpr int kCntQueue = 0, kCntHead = 1, kCntRis = 2, kCntTemporal = 3, kCntSpatial = 4, kCntPtBounce = 5,
              kCntPtShadow = 6, kCntQueueNear = 7, kCntPtQueue = 8 /* and 9 */, kCntPtHead = 10;
#ifndef CT_FAR_RAY
#define CT_FAR_RAY 2.0f
#endif
// A shadow ray at least this long (world units; the 256^3 workload is 8
// wide) goes to the front of the queue and is traced first (the frame's
// far_ray; a batch's autotuner picks its own).
constexpr float kFarRay = CT_FAR_RAY;

struct DFrame {
  uint32_t W, H, npix, frame, active, afternoon;
  float time;
  uint32_t n_cube, n_point, n_env, n_mats;
  float inv_view[16], inv_proj[16];
  ct_restir_params rp;
  ct_light_config cfg;
  DScene S;
  DScene Sp;          // the scene before the last update (get_previous_tlas) when prev_set, else S
  uint32_t prev_set;  // CT_REMAP_BIT with a distinct previous set: temporal slot-1 queries trace Sp
  const ct_aabb *aabbs;
  const ct_primitive *prims;
  const float4 *mats;   // [2 per material]: (diffuse, type), (emission, roughness)
  const float4 *cube;   // [3 per light]: (pos, size), (emissive, type), (inst, prim, -, -)
  const float4 *point;
  const float4 *env;
  const float *geo_t;   // [16 per instance] geometry_T = previous model * inverse(model)
  const uint32_t *remap_prim, *remap_light, *prev_first;  // remapping tables, previous first indices
  uint32_t remap_cap, remap_light_cap;
  float prev_vp[16];    // inverse(prev_inv_proj) * inverse(prev_inv_view)
  float4 *dpos_c, *dnrm_c, *dorg_c;
  uint4 *dhit_c;
  float4 *dpos_p, *dnrm_p, *dorg_p;
  uint4 *dhit_p;
  uint4 *rA_i; float4 *rB_i; float *rC_i;
  uint4 *rA_p; float4 *rB_p; float *rC_p;
  float2 *vel;
  float4 *img;
  const float4 *img_prev;
  float4 *taa;
  float4 *hit;
  uint32_t *pidx;
  uint32_t *mask;
  float4 *ray_o, *ray_d;
  uint32_t *ray_t;
  uint32_t *queue;   // [queue_cap]: far rays from the front, near rays from the back
  uint32_t queue_cap;
  uint32_t *counters;
  float4 *gat;  // [8 per pixel] neighbour record, see st_gather
  float4 *spc;  // spatial reuse's per-neighbour results, see spc_at
  float4 *tpc;  // temporal reuse's two targets and flags, [2 per pixel], see temporal_emit
  float4 *pres_o;    // [6 per pixel] output path reservoirs (this frame)
  float4 *pres_t;    // [6 per pixel] temporal path reservoirs (history)
  float4 *indirect;  // indirect colour of the first pass
  float4 *pst;       // path state between the wavefront kernels (14 float4 per pixel)
  float4 *rc;        // hybrid shift: retrace records, 2 slots of 5 float4 per pixel
  uint32_t *ptq[2];  // path-segment queues (tagged pixel ids) of even / odd depths
  float4 *seg_o, *seg_d, *seg_hit;  // path segments and their hits by queue position
  uint32_t qtag;     // frame index << kViewShift when frames are batched, else 0
  uint32_t taa_w, taa_h;  // TAA dispatch size (main.cpp:1642-1643 truncation applied)
  float far_ray;     // shadow rays at least this long are queued (and traced) first
};

// Several frames in one launch: every block knows its frame and first
// pixel (or tile); queue entries carry the frame index in their top bits.
constexpr uint32_t kViewShift = 26, kSlotMask = (1u << kViewShift) - 1u;
struct FrameSet {
  const DFrame *fb;   // device array of frames
  const uint2 *blk;   // per block: (frame, first pixel or tile)
};

struct PLight {
  F3 pos;
  float size;
  F3 emissive;
  uint32_t type, inst, prim;
};
struct Mat {
  F3 diffuse, emission;
  float roughness;
};
struct Hit {  // HIT_INFO_INPUT, shared.inl:171-179
  F3 P, n;
  float distance;
  F3 scatter;
  uint32_t inst, prim, mat;
};
struct Res {  // RESERVOIR, shared.inl:368-377
  uint32_t Y, type, seed;
  float W_y, W_sum, M;
  F3 F;
};
struct DI {  // DIRECT_ILLUMINATION_INFO, shared.inl:430-440
  F3 pos;
  float distance;
  F3 nrm, org;
  uint32_t seed, inst, prim, mat;
  float conf;
};

CT_D Res res_init() {
  Res r;
  r.Y = kInvalid; r.type = CT_LIGHT_NONE; r.seed = 0;
  r.W_y = 0.0f; r.W_sum = 0.0f; r.M = 0.0f;
  r.F = f3(0.0f, 0.0f, 0.0f);
  return r;
}
CT_D Res ld_res(const uint4 *A, const float4 *B, const float *C, uint32_t i) {
  uint4 a = A[i];
  float4 b = B[i];
  Res r;
  r.Y = a.x; r.type = a.y; r.seed = a.z; r.W_y = __uint_as_float(a.w);
  r.W_sum = b.x; r.M = b.y; r.F = f3(b.z, b.w, C[i]);
  return r;
}
CT_D void st_res(uint4 *A, float4 *B, float *C, uint32_t i, const Res &r) {
  A[i] = make_uint4(r.Y, r.type, r.seed, __float_as_uint(r.W_y));
  B[i] = make_float4(r.W_sum, r.M, r.F.x, r.F.y);
  C[i] = r.F.z;
}
// Neighbour record, one 128 B line per pixel: everything spatial reuse reads
// from a neighbour (its DI minus seed/confidence, and its intermediate
// reservoir).  Written once the temporal pass is done.
constexpr uint32_t kGather = 8;
CT_D void st_gather(float4 *G, uint32_t i, const float4 &pos, const float4 &nrm, const float4 &org, uint32_t inst,
                    uint32_t prim, const Res &r) {
  float4 *g = G + (size_t)kGather * i;
  g[0] = pos;
  g[1] = nrm;
  g[2] = make_float4(org.x, org.y, org.z, __uint_as_float(inst));
  g[3] = make_float4(__uint_as_float(prim), __uint_as_float(r.Y), __uint_as_float(r.type), __uint_as_float(r.seed));
  g[4] = make_float4(r.W_y, r.W_sum, r.M, r.F.x);
  g[5] = make_float4(r.F.y, r.F.z, 0.0f, 0.0f);
}
CT_D void ld_gather(const float4 *G, uint32_t i, DI &d, Res &r) {
  const float4 *g = G + (size_t)kGather * i;
  float4 a = g[0], b = g[1], c = g[2], e = g[3], f = g[4], h = g[5];
  d.pos = xyz(a); d.distance = a.w;
  d.nrm = xyz(b); d.mat = __float_as_uint(b.w);
  d.org = xyz(c); d.inst = __float_as_uint(c.w);
  d.prim = __float_as_uint(e.x);
  d.seed = 0; d.conf = 0.0f;
  r.Y = __float_as_uint(e.y); r.type = __float_as_uint(e.z); r.seed = __float_as_uint(e.w);
  r.W_y = f.x; r.W_sum = f.y; r.M = f.z; r.F = f3(f.w, h.x, h.y);
}
CT_D float gather_distance(const float4 *G, uint32_t i) { return G[(size_t)kGather * i].w; }

CT_D DI ld_di(const float4 *P, const float4 *N, const float4 *O, const uint4 *Hh, uint32_t i) {
  float4 p = P[i], n = N[i], o = O[i];
  uint4 h = Hh[i];
  DI d;
  d.pos = xyz(p); d.distance = p.w;
  d.nrm = xyz(n); d.mat = __float_as_uint(n.w);
  d.org = xyz(o); d.seed = __float_as_uint(o.w);
  d.inst = h.x; d.prim = h.y; d.conf = __uint_as_float(h.z);
  return d;
}
CT_D void st_di(float4 *P, float4 *N, float4 *O, uint4 *Hh, uint32_t i, const DI &d) {
  P[i] = make_float4(d.pos.x, d.pos.y, d.pos.z, d.distance);
  N[i] = make_float4(d.nrm.x, d.nrm.y, d.nrm.z, __uint_as_float(d.mat));
  O[i] = make_float4(d.org.x, d.org.y, d.org.z, __uint_as_float(d.seed));
  Hh[i] = make_uint4(d.inst, d.prim, __float_as_uint(d.conf), 0u);
}
CT_D PLight ld_light(const float4 *L, uint32_t i) {
  float4 a = L[3 * i], b = L[3 * i + 1], c = L[3 * i + 2];
  PLight l;
  l.pos = xyz(a); l.size = a.w;
  l.emissive = xyz(b); l.type = __float_as_uint(b.w);
  l.inst = __float_as_uint(c.x); l.prim = __float_as_uint(c.y);
  return l;
}
CT_D Mat ld_mat(const DFrame &F, uint32_t i) {
  Mat m;
  if (i >= F.n_mats) {
    m.diffuse = f3(0, 0, 0);
    m.emission = f3(0, 0, 0);
    m.roughness = 0.0f;
    return m;
  }
  float4 a = F.mats[2 * i], b = F.mats[2 * i + 1];
  m.diffuse = xyz(a);
  m.emission = xyz(b);
  m.roughness = b.w;
  return m;
}
CT_D bool res_valid(const Res &r) { return r.M > 0.0f && r.Y != kInvalid && r.type != CT_LIGHT_NONE; }
CT_D bool light_by_type(const DFrame &F, uint32_t type, uint32_t idx, PLight &l) {
  if (type == CT_LIGHT_POINT && idx < F.n_point) { l = ld_light(F.point, idx); return true; }
  if (type == CT_LIGHT_ENV_MAP && idx < F.n_env) { l = ld_light(F.env, idx); return true; }
  if (type == CT_LIGHT_CUBE && idx < F.n_cube) { l = ld_light(F.cube, idx); return true; }
  return false;
}
CT_D Hit hit_from_di(const DI &d) {
  Hit h;
  F3 p = compute_ray_origin(d.pos, d.nrm);
  h.P = compute_ray_origin(p, d.nrm);
  h.n = d.nrm;
  h.distance = d.distance;
  h.scatter = f3(0, 0, 0);
  h.inst = d.inst; h.prim = d.prim; h.mat = d.mat;
  return h;
}

// reservoir.glsl:31-50
CT_D bool update_reservoir(Res &r, uint32_t X, uint32_t type, uint32_t rs, F3 Fv, float w, float c, uint32_t &seed) {
  if (w < 0.0f || isnan(w) || isinf(w)) return false;
  r.M += c;
  r.W_sum += w;
  if (rnd(seed) < (w / r.W_sum)) {
    r.Y = X; r.type = type; r.seed = rs; r.F = Fv;
    return true;
  }
  return false;
}

// ------------------------------------------------------- visibility source
// EMIT: a query that needs a shadow ray writes it to slot (pixel*16 + k)
// and reports "unoccluded".  FINISH: the k-th query reads bit k.
// `queries` counts the calls, made or taken from bits.
struct Vis {
  bool emit;
  uint32_t px, slot, bits, emitted, queries, far;
};
CT_D bool occluded(const DFrame &F, Vis &v, F3 o, F3 d, float dist, uint32_t ti, uint32_t tp) {
  ++v.queries;
  uint32_t k = v.slot++;
  if (k >= kRaySlots) return false;  // bounded by CT_MAX_NEIGHBORS; never taken
  if (v.emit) {
    uint32_t idx = v.px * kRaySlots + k;
    F.ray_o[idx] = make_float4(o.x, o.y, o.z, dist);
    F.ray_d[idx] = make_float4(d.x, d.y, d.z, __uint_as_float(ti));
    F.ray_t[idx] = tp;
    v.emitted |= 1u << k;
    if (dist >= F.far_ray) v.far |= 1u << k;
    return false;
  }
  return (v.bits >> k) & 1u;
}

// mat.glsl:24-27, 58-88; light.glsl:200-229 (Lambertian)
CT_D float local_z(F3 n, F3 v) { return dot(v, normalize(n)); }
CT_D float material_pdf(F3 n, F3 wo, F3 wi) {
  float a = local_z(n, wo), b = local_z(n, wi);
  if (gmin(a, b) < kMinCosTheta) return 0.0f;
  return kInvPi * b;
}
CT_D F3 evaluate_material(const Mat &m, F3 n, F3 wo, F3 wi) {
  float a = local_z(n, wo), b = local_z(n, wi);
  if (gmin(a, b) < kMinCosTheta) return f3(0, 0, 0);
  return muls(muls(m.diffuse, kInvPi), b);
}

// What calculate_sampled_light derives from the shading point alone, once
// per point instead of once per light sample: sn = normalize(hit.n), wo =
// -normalize(ray_d), the normal evaluate_material normalizes again and its
// cosine with wo, and the cosine material_pdf(hit.n, -ray_d, .) uses.  The
// same operations on the same inputs: values are unchanged.
struct Shade {
  F3 sn, nsn, wo;
  float a_eval, a_pdf;
};
CT_D Shade shade_of(F3 ray_d, const Hit &hit) {
  Shade s;
  s.sn = normalize(hit.n);
  s.nsn = normalize(s.sn);
  s.wo = neg(normalize(ray_d));
  s.a_eval = dot(s.wo, s.nsn);
  s.a_pdf = dot(neg(ray_d), s.sn);
  return s;
}
// evaluate_material(m, sn, wo, wi) with the shading point's terms precomputed
CT_D F3 evaluate_material(const Mat &m, const Shade &sh, F3 wi) {
  const float b = dot(wi, sh.nsn);
  if (gmin(sh.a_eval, b) < kMinCosTheta) return f3(0, 0, 0);
  return muls(muls(m.diffuse, kInvPi), b);
}
// material_pdf(hit.n, -ray_d, wi) with the shading point's terms precomputed
CT_D float material_pdf(const Shade &sh, F3 wi) {
  const float b = dot(wi, sh.sn);
  if (gmin(sh.a_pdf, b) < kMinCosTheta) return 0.0f;
  return kInvPi * b;
}
CT_D float geom_fact_sa(F3 P, F3 Ps, F3 ns) {
  F3 dir = normalize(sub(Ps, P));
  F3 dd = sub(Ps, P);
  return fabsf(dot(ns, dir)) / dot(dd, dd);
}
// prng.glsl:198-252
CT_D F3 random_cube_normal(uint32_t &seed) {
  uint32_t face = min(urnd_interval(seed, 0u, 6u), 5u);
  switch (face) {
    case 0: return f3(-1, 0, 0);
    case 1: return f3(1, 0, 0);
    case 2: return f3(0, -1, 0);
    case 3: return f3(0, 1, 0);
    case 4: return f3(0, 0, -1);
    default: return f3(0, 0, 1);
  }
}
CT_D F3 random_quad(F3 n, F3 p, float sx, float sy, uint32_t &seed) {
  float u = rnd(seed);
  float v = rnd(seed);
  float qx = u * sx, qy = v * sy;
  F3 t1 = fabsf(n.x) > fabsf(n.y) ? normalize(f3(-n.z, 0.0f, n.x)) : normalize(f3(0.0f, n.z, -n.y));
  F3 t2 = normalize(cross(n, t1));
  return add(add(p, muls(t1, qx)), muls(t2, qy));
}
CT_D F3 random_on_hemisphere(uint32_t &seed, F3 n) {
  F3 p = f3(0, 0, 0);
  for (int k = 0; k < 1024; ++k) {
    float x = rnd_interval(seed, -1.0f, 1.0f);
    float y = rnd_interval(seed, -1.0f, 1.0f);
    float z = rnd_interval(seed, -1.0f, 1.0f);
    p = f3(x, y, z);
    if (x * x + y * y + z * z >= 1.0f) continue;
    break;
  }
  F3 u = normalize(p);
  return dot(u, n) > 0.0f ? u : neg(u);
}

// LCG state after 33 steps: s * a^33 + c * (a^32 + ... + 1) (mod 2^32).
constexpr uint32_t lcg_mul_pow(int k) {
  uint32_t r = 1u;
  for (int i = 0; i < k; ++i) r *= 1664525u;
  return r;
}
constexpr uint32_t lcg_add_sum(int k) {
  uint32_t r = 0u, p = 1u;
  for (int i = 0; i < k; ++i) {
    r += p;
    p *= 1664525u;
  }
  return r * 1013904223u;
}
CT_D uint32_t lcg_skip33(uint32_t s) {
  constexpr uint32_t A = lcg_mul_pow(33), C = lcg_add_sum(33);
  return A * s + C;
}
// True when the light voxel lies wholly on the back side of the surface at
// P (axis-aligned normal n): then dot(normalize(P - x), n) >= 0 for every
// sample point x of the voxel, light.glsl:343 rejects all draws.  The
// 1e-4 margin covers the rounding of the sample position; near-boundary
// cases take the full loop.
CT_D bool light_behind_surface(F3 P, F3 n, F3 c, float half) {
  const float m = half + 1e-4f;
  if (n.y == 0.0f && n.z == 0.0f) {
    if (n.x == 1.0f) return P.x >= c.x + m;
    if (n.x == -1.0f) return P.x <= c.x - m;
  } else if (n.x == 0.0f && n.z == 0.0f) {
    if (n.y == 1.0f) return P.y >= c.y + m;
    if (n.y == -1.0f) return P.y <= c.y - m;
  } else if (n.x == 0.0f && n.y == 0.0f) {
    if (n.z == 1.0f) return P.z >= c.z + m;
    if (n.z == -1.0f) return P.z <= c.z - m;
  }
  return false;
}

struct LS {
  bool ok;      // the sample is valid before the shadow test
  bool ray;     // a shadow ray would be traced
  F3 P, dir;
  float dist;
  uint32_t ti, tp;
  F3 l_pos, Le;
  F3 wi;        // normalize(l_pos - P), the direction calculate_sampled_light shades
  float G;
};

// sample_lights, light.glsl:295-399, without the shadow ray.
CT_D void sample_lights(const DFrame &F, const Hit &hit, const PLight &l, float &pdf, uint32_t &seed, bool calc_pdf,
                        bool visibility, LS &o) {
  o.ok = false; o.ray = false; o.G = 0.0f; o.Le = f3(0, 0, 0); o.l_pos = f3(0, 0, 0);
  if (l.inst == hit.inst && l.prim == hit.prim) return;
  F3 P = hit.P, n = hit.n;
  float distance = -1.0f;
  F3 l_wi = f3(0, 0, 0), l_pos = f3(0, 0, 0), l_nor = f3(0, 0, 0);
  bool vis = true, check = false;
  F3 em = l.emissive;
  if (l.type == CT_LIGHT_CUBE) {
    float half = l.size * 0.5f, qs = l.size;
    if (light_behind_surface(P, n, l.pos, half)) {
      // Every one of the 11 draws fails the "in front of the surface" test,
      // so the result is a failed sample (its G and pdf only ever multiply
      // a zero contribution) after 11 x 3 = 33 LCG steps.
      seed = lcg_skip33(seed);
      return;
    }
    // light.glsl:333-345.  The face normal is an axis vector, so
    // random_quad's tangents are the exact axis vectors t1, t2 (their
    // normalize divides by 1), and dot(normalize(P - x), axis) has the sign
    // of (P - x) on that axis: the draw is tested by signs and only the
    // last one is normalized.  A normal n off the axes takes the full test.
    int na = -1;
    float ns = 0.0f;
    if (n.y == 0.0f && n.z == 0.0f && fabsf(n.x) == 1.0f) { na = 0; ns = n.x; }
    else if (n.x == 0.0f && n.z == 0.0f && fabsf(n.y) == 1.0f) { na = 1; ns = n.y; }
    else if (n.x == 0.0f && n.y == 0.0f && fabsf(n.z) == 1.0f) { na = 2; ns = n.z; }
    const float hq = qs * 0.5f;
    uint32_t i = 0;
    do {
      l_nor = random_cube_normal(seed);
      l_pos = add(l.pos, muls(l_nor, half));
      const float u = rnd(seed);
      const float v = rnd(seed);
      const float qx = u * hq, qy = v * hq;
      const F3 t1 = fabsf(l_nor.x) > fabsf(l_nor.y) ? f3(-l_nor.z, 0.0f, l_nor.x) : f3(0.0f, l_nor.z, -l_nor.y);
      const F3 t2 = cross(l_nor, t1);
      l_pos = add(add(l_pos, muls(t1, qx)), muls(t2, qy));
      const F3 w = sub(P, l_pos);
      const float wn = l_nor.x != 0.0f ? w.x * l_nor.x : (l_nor.y != 0.0f ? w.y * l_nor.y : w.z * l_nor.z);
      vis = wn > 0.0f;
      if (vis) {
        if (na >= 0) {
          vis = (na == 0 ? w.x : (na == 1 ? w.y : w.z)) * ns < 0.0f;
        } else {
          vis = dot(normalize(w), n) < 0.0f;
        }
      }
    } while (!vis && i++ < 10);
    l_wi = normalize(sub(P, l_pos));
    l_pos = compute_ray_origin(l_pos, l_nor);
    l_pos = compute_ray_origin(l_pos, l_nor);
    // length(P - l_pos), geom_fact_sa(P, l_pos, l_nor) and the shading
    // direction share one square root: P - l_pos is exactly -(l_pos - P),
    // and for the axis vector l_nor, dot(l_nor, dir) is the component of
    // dir on that axis up to sign (the other terms are signed zeros).
    const F3 dd = sub(l_pos, P);
    const float d2 = dot(dd, dd);
    distance = sqrtf(d2);
    o.wi = divs(dd, distance);
    if (calc_pdf) {
      float area = qs * qs * 6.0f;
      pdf /= area;
    }
    check = true;
    o.G = fabsf(l_nor.x != 0.0f ? o.wi.x : (l_nor.y != 0.0f ? o.wi.y : o.wi.z)) / d2;
  } else if (l.type == CT_LIGHT_POINT) {
    // As above: normalize(l_pos - P) is exactly -normalize(P - l_pos).
    l_pos = l.pos;
    const F3 v = sub(P, l_pos);
    const float d2 = dot(v, v);
    distance = sqrtf(d2);
    l_nor = divs(v, distance);
    l_wi = l_nor;
    o.wi = neg(l_nor);
    o.G = fabsf(dot(l_nor, o.wi)) / d2;
  } else if (l.type == CT_LIGHT_ENV_MAP) {
    F3 l_dir = random_on_hemisphere(seed, n);
    l_pos = add(P, muls(l_dir, kMaxDistance));
    l_nor = normalize(sub(P, l_pos));
    distance = length(sub(P, l_pos));
    l_wi = l_nor;
    em = mulv(em, sky_color(F.time, F.afternoon != 0, l_dir));
    if (calc_pdf) pdf /= kInv4Pi;
    o.G = 1.0f;
  }
  if (l.type != CT_LIGHT_CUBE && l.type != CT_LIGHT_POINT) o.wi = normalize(sub(l_pos, P));
  o.ray = visibility && vis;
  o.P = P;
  o.dir = neg(l_wi);
  o.dist = distance;
  o.ti = check ? l.inst : kInvalid;
  o.tp = check ? l.prim : kInvalid;
  o.l_pos = l_pos;
  o.Le = em;
  o.ok = vis;
}

// calculate_sampled_light, light.glsl:401-435.  With use_vis the shadow
// ray goes through `v`.
CT_D F3 sampled_light(const DFrame &F, const Shade &sh, Hit &hit, const Mat &mat, const PLight &l, float pdf,
                      float &pdf_out, float &G, uint32_t &seed, bool calc_pdf, bool use_pdf, bool use_vis, Vis *v) {
  pdf_out = pdf;
  G = 0.0f;
  if (l.emissive.x <= 0.0f && l.emissive.y <= 0.0f && l.emissive.z <= 0.0f) return f3(0, 0, 0);
  LS ls;
  sample_lights(F, hit, l, pdf_out, seed, calc_pdf, use_vis, ls);
  G = ls.G;
  bool ok = ls.ok;
  if (ls.ray && occluded(F, *v, ls.P, ls.dir, ls.dist, ls.ti, ls.tp)) ok = false;
  if (!ok) return f3(0, 0, 0);
  const F3 wi = ls.wi;  // normalize(ls.l_pos - hit.P)
  F3 brdf = evaluate_material(mat, sh, wi);
  hit.scatter = wi;
  F3 bl = mulv(brdf, ls.Le);
  return use_pdf ? divs(muls(bl, G), pdf_out) : muls(bl, G);
}

// reservoir_get_radiance, reservoir.glsl:98-112
CT_D F3 res_radiance(const DFrame &F, const Res &r, const Shade &sh, Hit hit, const Mat &mat, Vis &v) {
  if (!res_valid(r)) return f3(0, 0, 0);
  PLight l;
  if (!light_by_type(F, r.type, r.Y, l)) return f3(0, 0, 0);
  float pdf_out, G;
  uint32_t s = r.seed;
  return sampled_light(F, sh, hit, mat, l, 1.0f, pdf_out, G, s, false, false, true, &v);
}
CT_D F3 res_radiance(const DFrame &F, const Res &r, F3 ray_d, Hit hit, const Mat &mat, Vis &v) {
  if (!res_valid(r)) return f3(0, 0, 0);
  return res_radiance(F, r, shade_of(ray_d, hit), hit, mat, v);
}

// reservoir_visibility_pass, reservoir.glsl:76-96, 114-126.  Returns the
// visibility it applied (in an emitting pass: whether the light sample is
// valid, its ray queued).
CT_D bool visibility_pass(const DFrame &F, Res &r, const Hit &hit, Vis &v) {
  bool visibility = false;
  if (res_valid(r)) {
    PLight l;
    if (light_by_type(F, r.type, r.Y, l)) {
      float pdf = 1.0f;
      uint32_t s = r.seed;
      LS ls;
      sample_lights(F, hit, l, pdf, s, false, true, ls);
      visibility = ls.ok;
      if (ls.ray && occluded(F, v, ls.P, ls.dir, ls.dist, ls.ti, ls.tp)) visibility = false;
    }
  }
  r.W_y = visibility ? (r.W_sum / luminance(r.F)) : 0.0f;
  r.F = muls(r.F, visibility ? 1.0f : 0.0f);
  return visibility;
}

// k_ris records the outcome of its visibility pass's light sample in the
// pixel's occlusion mask, above the ray slots' bits: kRisOk (the sample is
// valid) and kRisRay (its ray took slot 0).  The finishing pass then applies
// the ray's bit without drawing the sample again.
constexpr uint32_t kRisOk = 1u << 31, kRisRay = 1u << 30;
static_assert(kRaySlots <= 30, "the RIS outcome bits sit above the ray slots");
CT_D void visibility_finish(Res &r, uint32_t mask) {
  bool visibility = (mask & kRisOk) != 0u;
  if ((mask & kRisRay) && (mask & 1u)) visibility = false;
  r.W_y = visibility ? (r.W_sum / luminance(r.F)) : 0.0f;
  r.F = muls(r.F, visibility ? 1.0f : 0.0f);
}

// RIS candidate streaming of one light type, restir_resampling.glsl:72-178
CT_D void ris_strategy(const DFrame &F, Res &target, uint32_t count, uint32_t nsamp, uint32_t type, const float4 *lights,
                       uint32_t nee, uint32_t n_brdf, const Shade &sh, Hit &hit, const Mat &mat, uint32_t &seed) {
  float pdf = 1.0f / (float)count;
  for (uint32_t k = 0; k < nsamp; ++k) {
    uint32_t idx = min(urnd_interval(seed, 0u, count), count - 1u);
    PLight light = ld_light(lights, idx);
    float light_pdf = 1.0f, G = 0.0f;
    uint32_t current_seed = seed;
    F3 Fv = sampled_light(F, sh, hit, mat, light, pdf, light_pdf, G, seed, true, false, false, nullptr);
    float brdf_pdf = material_pdf(sh, hit.scatter) * G;
    float m_i = 1.0f / gmax((float)nee * light_pdf + (float)n_brdf * brdf_pdf, 1e-6f);
    float w_i = luminance(Fv) * m_i;
    update_reservoir(target, idx, type, current_seed, Fv, w_i, 1.0f, seed);
  }
}

// Candidate counts of RIS (restir_resampling.glsl:50-65) that its BRDF
// strategy's weights use.
struct RisCounts {
  uint32_t nee, n_brdf, cube_lights;
};

// RIS, restir_resampling.glsl:19-255: the point, environment and cube
// strategies (the BRDF strategy follows in ris_brdf).
CT_D Res ris(const DFrame &F, float confidence, F3 ray_d, Hit &hit, const Mat &mat, uint32_t &seed, RisCounts &rc) {
  const ct_restir_params &rp = F.rp;
  confidence = clampf(confidence, 0.0f, 1.0f);
  const uint32_t pc = F.cfg.point_light_count, ec = F.cfg.env_map_count;
  uint32_t n_point = max(f2u_sat(gmin((float)rp.max_ris_point_samples * (1.0f - confidence), (float)pc)), rp.min_ris_point_samples);
  uint32_t n_env = max(f2u_sat(gmin((float)rp.max_ris_env_samples * (1.0f - confidence), (float)ec)), rp.min_ris_env_samples);
  uint32_t cc = F.cfg.cube_light_count < 1 ? 1u : F.cfg.cube_light_count;
  uint32_t n_cube = max(f2u_sat(gmin((float)rp.max_ris_cube_samples * (1.0f - confidence), (float)cc)), rp.min_ris_cube_samples);
  uint32_t n_brdf = F.cfg.brdf_count + n_env;
  uint32_t nee = n_point + n_cube;
  rc.nee = nee;
  rc.n_brdf = n_brdf;
  rc.cube_lights = cc;
  const Shade sh = shade_of(ray_d, hit);
  Res r = res_init();
  if ((F.active & CT_RIS_POINT_LIGHT_BIT) && F.n_point > 0)
    ris_strategy(F, r, pc, n_point, CT_LIGHT_POINT, F.point, nee, n_brdf, sh, hit, mat, seed);
  if ((F.active & CT_RIS_ENV_LIGHT_BIT) && F.n_env > 0) {
    Res er = res_init();
    ris_strategy(F, er, ec, n_env, CT_LIGHT_ENV_MAP, F.env, nee, n_brdf, sh, hit, mat, seed);
    update_reservoir(r, er.Y, er.type, er.seed, er.F, er.W_sum, er.M, seed);
  }
  if ((F.active & CT_RIS_CUBE_LIGHT_BIT) && F.n_cube > 0) {
    Res cr = res_init();
    ris_strategy(F, cr, cc, n_cube, CT_LIGHT_CUBE, F.cube, nee, n_brdf, sh, hit, mat, seed);
    update_reservoir(r, cr.Y, cr.type, cr.seed, cr.F, cr.W_sum, cr.M, seed);
  }
  return r;
}

// TEMPORAL_REUSE, restir_resampling.glsl:311-416.  With CT_REMAP_BIT the
// previous frame's hit primitive and cube light go through the remapping
// tables (:338-345, :380-387): a deleted primitive ends the reuse, a moved
// one or a moved light takes its current index.
// Its two visibility queries take the slots v.slot and v.slot + 1, whether
// made or not, so that temporal_emit can queue them before the RIS sample's
// own visibility is known.  temporal_emit also evaluates both targets; this
// pass takes them from the pixel's record (tpc) and applies the rays' bits.
constexpr uint32_t kTpRayA = 2u, kTpDeleted = 4u, kTpRayB = 16u;
CT_D void temporal_reuse(const DFrame &F, Res &r, Res prev, uint32_t pidx, F3 ray_d, const Hit &hit, const Mat &mat,
                         uint32_t &seed, Vis &v) {
  const float mcap = F.rp.max_influence_from_the_past;
  const bool remap = (F.active & CT_REMAP_BIT) != 0u;
  const uint32_t slot0 = v.slot;
  prev.M = gmin(prev.M, mcap * r.M);
  const float prev_M = prev.M;
  const float new_M = r.M + prev_M;
  // the two targets temporal_emit evaluated (unoccluded) and its flags
  const float4 ta = F.tpc[v.px], tb = F.tpc[F.npix + v.px];
  const uint32_t fl = __float_as_uint(ta.w);
  if (r.W_sum > 0.0f) {
    float tlp = 0.0f;
    if (luminance(r.F) > 1e-6f) {
      if (fl & kTpDeleted) return;  // the previous hit primitive was deleted
      F3 t = xyz(ta);
      v.slot = slot0;
      if ((fl & kTpRayA) && occluded(F, v, f3(0, 0, 0), f3(0, 0, 0), 0.0f, 0u, 0u)) t = f3(0, 0, 0);
      tlp = luminance(t);
    }
    const float p_curr = r.M * luminance(r.F);
    const float m_curr = p_curr / gmax(p_curr + prev_M * tlp, 1e-6f);
    r.W_sum *= m_curr;
  }
  if (remap && prev.type == CT_LIGHT_CUBE && prev.Y < F.remap_light_cap) {
    const uint32_t m = F.remap_light[prev.Y];
    if (m != 0u) prev.Y = m;
  }
  if (prev.Y != kInvalid) {
    v.slot = slot0 + 1u;
    F3 ct = xyz(tb);
    if ((fl & kTpRayB) && occluded(F, v, f3(0, 0, 0), f3(0, 0, 0), 0.0f, 0u, 0u)) ct = f3(0, 0, 0);
    const float lc = luminance(ct);
    if (lc > 0.0f) {
      const float lp = luminance(prev.F);
      const float p_prev = prev.M * lp;
      const float m_prev = p_prev / gmax(p_prev + r.M * lc, 1e-6f);
      const float w_prev = m_prev * lc * prev.W_y;
      update_reservoir(r, prev.Y, prev.type, prev.seed, ct, w_prev, 1.0f, seed);
    }
  }
  float tl = luminance(r.F);
  r.W_y = tl > 0.0f ? r.W_sum / tl : 0.0f;
  r.M = new_M;
}

// Queues every shadow ray temporal_reuse may query, before the RIS sample's
// visibility is known: `r` is the RIS reservoir as if its sample were
// visible. A finish pass that sees the sample occluded makes a subset of
// these queries (the first needs luminance(r.F) > 0, which occlusion only
// clears; the second depends on `prev` alone), in the same slots; the rays
// it does not query are traced and ignored.
CT_D void temporal_emit(const DFrame &F, const Res &r, Res prev, uint32_t pidx, F3 ray_d, const Hit &hit,
                        const Mat &mat, Vis &v) {
  const bool remap = (F.active & CT_REMAP_BIT) != 0u;
  const uint32_t slot0 = v.slot;
  prev.M = gmin(prev.M, F.rp.max_influence_from_the_past * r.M);
  uint32_t fl = 0u;
  F3 ta = f3(0, 0, 0), tb = f3(0, 0, 0);
  if (r.W_sum > 0.0f && luminance(r.F) > 1e-6f) {
    DI dp = ld_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, pidx);
    uint32_t m = 0u;
    if (remap) {
      const uint32_t slot = (dp.inst < F.S.n_inst ? F.prev_first[dp.inst] : 0u) + dp.prim;
      m = slot < F.remap_cap ? F.remap_prim[slot] : 0u;
      if (m != kInvalid && m != 0u) dp.prim = m;
    }
    if (m != kInvalid) {
      Hit ph = hit_from_di(dp);
      Mat pm = ld_mat(F, dp.mat);
      F3 pd = normalize(sub(ph.P, dp.org));
      v.slot = slot0;
      ta = res_radiance(F, r, pd, ph, pm, v);
      if (v.slot != slot0) fl |= kTpRayA;
    } else {
      fl |= kTpDeleted;
    }
  }
  if (remap && prev.type == CT_LIGHT_CUBE && prev.Y < F.remap_light_cap) {
    const uint32_t m = F.remap_light[prev.Y];
    if (m != 0u) prev.Y = m;
  }
  if (prev.Y != kInvalid) {
    v.slot = slot0 + 1u;
    tb = res_radiance(F, prev, ray_d, hit, mat, v);
    if (v.slot != slot0 + 1u) fl |= kTpRayB;
  }
  v.slot = slot0 + 2u;
  F.tpc[v.px] = make_float4(ta.x, ta.y, ta.z, __uint_as_float(fl));
  F.tpc[F.npix + v.px] = make_float4(tb.x, tb.y, tb.z, 0.0f);
}

// The emit pass of spatial reuse keeps, per pixel, what the finish pass needs of each
// neighbour in place of a second gather: header spc[px] = (neighbour count,
// the seed after the neighbour draw); neighbour i's fields j = 0..2 at
// spc[(1 + 3 i + j) * npix + px]: (target of its sample here, luminance of
// this pixel's sample's target there), (M, luminance of its F, W_y, flags),
// (Y, type, seed).  The targets are the unoccluded ones; flag bits 1 and 2
// say whether each took a visibility slot, whose bit the finish pass reads.
CT_D float4 *spc_at(const DFrame &F, uint32_t field, uint32_t px) { return F.spc + (size_t)field * F.npix + px; }

// SPATIAL_REUSE, restir_resampling.glsl:418-540, with pairwise_mis.glsl:8-114.  kEmit:
// the pass that queues the visibility rays (taken as unoccluded) and keeps
// the per-neighbour results; else the pass that finishes from them with the
// rays' bits.
template <bool kEmit>
CT_D void spatial_reuse(const DFrame &F, Res &r, float confidence, uint32_t x, uint32_t y, F3 ray_d, const Hit &hit,
                        const Mat &mat, uint32_t &seed, Vis &v) {
  const uint32_t px = y * F.W + x;
  if (!kEmit) {
    const float4 hd = *spc_at(F, 0, px);
    const uint32_t num = __float_as_uint(hd.x);
    seed = __float_as_uint(hd.y);
    if (num == 0) return;
    Res mres = res_init();
    float m_c = 1.0f;
    float M_s = r.M;
    const float fk = (float)num;
    for (uint32_t i = 0; i < num; ++i) {
      const float4 a = *spc_at(F, 1 + 3 * i, px), b = *spc_at(F, 2 + 3 * i, px), c = *spc_at(F, 3 + 3 * i, px);
      const uint32_t fl = __float_as_uint(b.w);
      const float in_M = b.x;
      bool occ_a = false, occ_b = false;
      if (fl & 2u) occ_a = occluded(F, v, f3(0, 0, 0), f3(0, 0, 0), 0.0f, 0u, 0u);
      if (fl & 4u) occ_b = occluded(F, v, f3(0, 0, 0), f3(0, 0, 0), 0.0f, 0u, 0u);
      const F3 curr_target = occ_a ? f3(0, 0, 0) : xyz(a);
      float m_i = 0.0f;
      if (fl & 1u) {
        const float tl = luminance(curr_target);
        const float mm = in_M * b.y;
        const float den = mm + (r.M / fk) * tl;
        m_i = den > 0.0f ? mm / den : 0.0f;
      }
      {
        const float pic = occ_b ? 0.0f : a.w;
        const float pcc = luminance(r.F);
        const float numer = in_M * pic;
        const bool den_ok = (pcc + numer) > 0.0f;
        m_c += 1.0f;
        m_c -= den_ok ? numer / (numer + (r.M / fk) * pcc) : 0.0f;
      }
      if (fl & 1u) {
        const float w_i = luminance(curr_target) * b.z * m_i;
        update_reservoir(mres, __float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.z), curr_target, w_i,
                         1.0f, seed);
      }
      M_s += in_M;
    }
    uint32_t s2 = seed;
    const float w_c = luminance(r.F) * r.W_y * m_c;
    update_reservoir(mres, r.Y, r.type, r.seed, r.F, w_c, 1.0f, s2);
    mres.M = M_s;
    const float tl = luminance(mres.F);
    mres.W_y = tl > 0.0f ? mres.W_sum / (tl * (float)(1u + num)) : 0.0f;
    r = mres;
    return;
  }
  const ct_restir_params &rp = F.rp;
  if (r.W_y == 0.0f) confidence = 0.0f;
  float radius = mixf(rp.max_neighbor_radius, rp.min_neighbor_radius, clampf(hit.distance / rp.max_distance_to_hit, 0.0f, 1.0f));
  uint32_t k = f2u_sat(mixf((float)rp.max_neighbors, (float)rp.min_neighbors, clampf(confidence, 0.0f, 1.0f)));
  uint32_t idx[CT_MAX_NEIGHBORS];
  uint32_t num = 0;
  for (uint32_t i = 0; i < k; ++i) {
    float ox = 2.0f * rnd(seed) - 1.0f;
    float oy = 2.0f * rnd(seed) - 1.0f;
    int64_t nx = (int64_t)x + f2i_sat(ox * radius);
    int64_t ny = (int64_t)y + f2i_sat(oy * radius);
    if (nx < 0 || nx >= (int64_t)F.W || ny < 0 || ny >= (int64_t)F.H) continue;
    if (nx == (int64_t)x && ny == (int64_t)y) continue;
    uint32_t l = (uint32_t)ny * F.W + (uint32_t)nx;
    if (gather_distance(F.gat, l) <= 0.0f) continue;
    if (num < CT_MAX_NEIGHBORS) idx[num++] = l;
  }
  *spc_at(F, 0, px) = make_float4(__uint_as_float(num), __uint_as_float(seed), 0.0f, 0.0f);
  if (num == 0) return;
  const Shade sh = shade_of(ray_d, hit);
  for (uint32_t i = 0; i < num; ++i) {
    Res in;
    DI nd;
    ld_gather(F.gat, idx[i], nd, in);
    Hit nh = hit_from_di(nd);
    F3 nray_d = normalize(sub(nh.P, nd.org));
    Mat nm = ld_mat(F, nd.mat);
    F3 curr_target = f3(0, 0, 0);
    uint32_t fl = 0u;
    float pii = 0.0f;
    const uint32_t s0 = v.slot;
    if (res_valid(in)) {
      curr_target = res_radiance(F, in, sh, hit, mat, v);
      pii = luminance(in.F);
      fl |= 1u;
    }
    const uint32_t s1 = v.slot;
    F3 input_target = f3(0, 0, 0);
    if (res_valid(r)) input_target = res_radiance(F, r, nray_d, nh, nm, v);
    if (s1 != s0) fl |= 2u;
    if (v.slot != s1) fl |= 4u;
    *spc_at(F, 1 + 3 * i, px) = make_float4(curr_target.x, curr_target.y, curr_target.z, luminance(input_target));
    *spc_at(F, 2 + 3 * i, px) = make_float4(in.M, pii, in.W_y, __uint_as_float(fl));
    *spc_at(F, 3 + 3 * i, px) = make_float4(__uint_as_float(in.Y), __uint_as_float(in.type), __uint_as_float(in.seed), 0.0f);
  }
}

// Camera ray of pixel (x, y): get_ray_from_current_pixel, rgen.glsl:18-42
// (jitter 0).
CT_D void camera_ray(const DFrame &F, uint32_t x, uint32_t y, F3 &o, F3 &d) {
  float pcx = ((float)x + 0.0f) + 0.5f, pcy = ((float)y + 0.0f) + 0.5f;
  float ux = pcx / (float)F.W, uy = pcy / (float)F.H;
  float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
  float ov[4], t4[4], d4[4];
  mat4_vec(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, ov);
  mat4_vec(F.inv_proj, dx, dy, 1.0f, 1.0f, t4);
  F3 nt = normalize(f3(t4[0], t4[1], t4[2]));
  mat4_vec(F.inv_view, nt.x, nt.y, nt.z, 0.0f, d4);
  o = f3(ov[0], ov[1], ov[2]);
  d = f3(d4[0], d4[1], d4[2]);
}

#include "restir_pt.h"

// Sums a per-lane count over the wave; lane 0 adds it to `dst`.
CT_D void wave_add(uint32_t v, uint32_t *dst) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (__lane_id() == 0 && v) atomicAdd(dst, v);
}
CT_D void wave_add_cnt(const TraceCnt &c, unsigned long long *dst) {
  unsigned long long a = c.nodes, b = c.leaves, d = c.insts;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
    d += __shfl_down(d, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(&dst[0], a);
    atomicAdd(&dst[1], b);
    atomicAdd(&dst[2], d);
  }
}

// ----------------------------------------------------------- queue push
// Appends this thread's emitted slots (bit k -> ray pixel*16 + k) to the
// ray queue with two atomics per block and adds their number to counter
// `counter` (none when negative). Slots in `far` (rays of at least kFarRay)
// fill the queue from the front (count kCntQueue), the others from the back
// end down (count kCntQueueNear): the traversal takes far rays first, so
// the long walks start early and waves hold rays of like length. Every
// thread of the block calls it.
CT_D void queue_flush(const DFrame &F, uint32_t px, uint32_t emitted, uint32_t far, int counter) {
  __shared__ uint32_t s_wsum[kBlock / 64];
  __shared__ uint32_t s_base_far, s_base_near;
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  far &= emitted;
  const uint32_t nf = (uint32_t)__popc(far), nn = (uint32_t)__popc(emitted) - nf;
  const uint32_t c = nf | (nn << 16);  // both counts in one word (at most 4096 per block)
  uint32_t incl = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t v = __shfl_up(incl, off);
    if ((int)lane >= off) incl += v;
  }
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int i = 0; i < kBlock / 64; ++i) {
      uint32_t t = s_wsum[i];
      s_wsum[i] = tot;
      tot += t;
    }
    const uint32_t tf = tot & 0xFFFFu, tn = tot >> 16;
    s_base_far = tf ? atomicAdd(&F.counters[kCntQueue], tf) : 0u;
    s_base_near = tn ? atomicAdd(&F.counters[kCntQueueNear], tn) : 0u;
    if ((tf + tn) && counter >= 0) atomicAdd(&F.counters[counter], tf + tn);
  }
  __syncthreads();
  const uint32_t before = s_wsum[w] + incl - c;
  uint32_t pf = s_base_far + (before & 0xFFFFu), pn = s_base_near + (before >> 16);
  while (emitted) {
    const uint32_t k = (uint32_t)(__ffs(emitted) - 1);
    emitted &= emitted - 1u;
    const uint32_t pos = ((far >> k) & 1u) ? pf++ : F.queue_cap - 1u - pn++;
    F.queue[pos] = F.qtag | (px * kRaySlots + k);
  }
}

// ---------------------------------------------------------------- kernels
// Primary rays: raygen rgen.glsl:18-42 + closest hit; one 8x8 tile per wave.
template <bool kCount>
__global__ void __launch_bounds__(kBlock) k_primary(FrameSet FS, unsigned long long *stats) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t tiles_x = (F.W + 15) / 16;
  const uint32_t bx = blk_.y % tiles_x, by = blk_.y / tiles_x;
  const uint32_t x = bx * 16 + (w & 1u) * 8 + (lane & 7u);
  const uint32_t y = by * 16 + (w >> 1) * 8 + (lane >> 3);
  TraceCnt cnt;
  if (x < F.W && y < F.H) {
    F3 o, d;
    camera_ray(F, x, y, o, d);
    float bt = kMaxDistance;
    uint32_t bi = kInvalid, bp = kInvalid;
    bool ovf = false;
    traverse<false, kCount>(F.S, o, d, kMaxDistance, kInvalid, kInvalid, bt, bi, bp, stk + tid, kBlock, cnt, ovf);
    F.hit[y * F.W + x] = make_float4(bi != kInvalid ? bt : -1.0f, __uint_as_float(bi), __uint_as_float(bp), 0.0f);
  }
  if (kCount) {
    unsigned long long a = cnt.nodes, b = cnt.leaves, c = cnt.insts;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      c += __shfl_down(c, off);
    }
    if (lane == 0) {
      atomicAdd(&stats[0], a);
      atomicAdd(&stats[1], b);
      atomicAdd(&stats[2], c);
    }
  }
}

// Pixel of this thread in a pass that walks the image by tiles: a block
// covers a 16x16 tile (entry `tile` of the tile table), each wave an 8x8
// quarter, as the primary rays do.  A wave's pixels then read neighbouring
// previous-frame pixels, neighbour records and voxels that its 2D
// neighbours share in the caches, and the rays it queues start close
// together.  On the 1080p frame this is 4% faster than runs of 256
// consecutive pixels (3.41 against 3.55 ms; k_ris 0.47 against 0.54 ms,
// spatial reuse 0.36 against 0.39 ms).  CT_TILED=0 builds the runs.
#ifndef CT_TILED
#define CT_TILED 1
#endif
constexpr bool kTiled = CT_TILED != 0;
CT_D bool tile_pixel(const DFrame &F, uint32_t tile, uint32_t &px) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t tiles_x = (F.W + 15) / 16;
  const uint32_t x = (tile % tiles_x) * 16 + (w & 1u) * 8 + (lane & 7u);
  const uint32_t y = (tile / tiles_x) * 16 + (w >> 1) * 8 + (lane >> 3);
  px = y * F.W + x;
  return x < F.W && y < F.H;
}
// This thread's pixel: tiled, or the block's run of consecutive pixels.
CT_D bool pass_pixel(const DFrame &F, uint2 blk, uint32_t &px) {
  if (kTiled) return tile_pixel(F, blk.y, px);
  px = blk.y + threadIdx.x;
  return px < F.npix;
}

// The BRDF strategy of RIS, restir_resampling.glsl:190-244: brdf_count
// directions from sample_material (light.glsl:232-270 with the Lambertian
// callable, rcall_scatter.glsl:73-84, which also leaves the direction in
// hit.scatter), each traced to its closest voxel in place (intersect(),
// bounce.glsl:9-130; `stk` is the lane's LDS stack column).  A candidate
// whose voxel emits has F = brdf * Le * G with G = geom_fact_sa at the hit
// (NaN, inf and negative G taken as 0) and m_i = 1 / max(n_brdf * pdf * G +
// nee * p_light * G, 1e-6), p_light = (1 / cube lights) / (6 VOXEL_EXTENT^2)
// (sample_lights_pdf, light.glsl:272-293), and names the voxel's cube light
// (PRIMITIVE.light_index); any other candidate has weight 0 and light -1.
// All are CUBE candidates.  Returns whether the merge into `r` selected the
// strategy's reservoir: then the reservoir skips the visibility pass.
CT_D bool ris_brdf(const DFrame &F, Res &r, F3 ray_o, Hit &hit, const Mat &mat, const RisCounts &rc, uint32_t &seed,
                   uint32_t *stk, uint32_t &rays) {
  Res br = res_init();
  for (uint32_t l = 0; l < F.cfg.brdf_count; ++l) {
    float w_i = 0.0f;
    F3 Fv = f3(0, 0, 0);
    const F3 P = hit.P, n = hit.n;
    const F3 wo = normalize(sub(ray_o, P));
    uint32_t light_index = kInvalid;
    const uint32_t current_seed = seed;
    F3 wi = random_cosine_direction(seed, n);
    if (normal_near_zero(wi)) wi = n;
    const float mat_pdf = material_pdf(n, wo, wi);
    hit.scatter = wi;
    const F3 ro = compute_ray_origin(P, wi);
    float bt = kMaxDistance;
    uint32_t bi = kInvalid, bp = kInvalid;
    bool ovf = false;
    TraceCnt cnt;
    traverse_mode<kTraceBounce, false>(F.S, ro, wi, kMaxDistance, kInvalid, kInvalid, bt, bi, bp, stk, kBlock, cnt,
                                       ovf);
    ++rays;
    if (bi != kInvalid) {
      const uint32_t gp = F.S.inst[bi].first + bp;
      const Mat hm = ld_mat(F, F.prims[gp].material_index);
      if (hm.emission.x > 0.0f || hm.emission.y > 0.0f || hm.emission.z > 0.0f) {
        float t;
        F3 pos, hn;
        hit_from_ray(F, bi, bp, ro, wi, t, pos, hn);
        float G = geom_fact_sa(P, pos, hn);
        if (isnan(G) || isinf(G) || G < 0.0f) G = 0.0f;
        const float light_pdf = (1.0f / (float)rc.cube_lights) / (kVoxelExtent * kVoxelExtent * 6.0f);
        const float m_i = 1.0f / gmax((float)rc.n_brdf * mat_pdf * G + (float)rc.nee * light_pdf * G, 1e-6f);
        const F3 brdf = evaluate_material(mat, n, wo, wi);
        Fv = muls(mulv(brdf, hm.emission), G);
        w_i = luminance(Fv) * m_i;
        light_index = F.prims[gp].light_index;
      }
    }
    update_reservoir(br, light_index, CT_LIGHT_CUBE, current_seed, Fv, w_i, 1.0f, seed);
  }
  return update_reservoir(r, br.Y, br.type, br.seed, br.F, br.W_sum, br.M, seed);
}

// An intermediate reservoir that skips the visibility pass (its BRDF
// candidate was selected) carries these bits as its W_y from k_ris to
// k_temporal_finish, which stores W_y = 0 (initialise_reservoir's value,
// reservoir.glsl:17-25) in their place.  Every other reservoir carries +0
// there until the pass.
constexpr uint32_t kNoVisibilityPass = 0x80000000u;

// Prepass after the primary hit: rchit (rchit.glsl:18-54), miss
// (rmiss.glsl:19-29), motion vector, temporal gather, confidence and RIS.
// EMIT for the RIS visibility ray.  kBrdf: some frame of the set has the
// RIS BRDF strategy on (its rays need an LDS stack).
template <bool kBrdf>
CT_D void ris_pass(const FrameSet &FS) {
  __shared__ uint32_t stk[kBrdf ? (kStackLds > 0 ? kStackLds : 1) * kBlock : 1];
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px;
  const bool mine = pass_pixel(F, blk_, px);
  uint32_t emitted = 0, far = 0, ris_ray = 0, brdf_rays = 0, ris_mask = 0;
  if (mine) {
    const uint32_t x = px % F.W, y = px / F.W;
    float4 h = F.hit[px];
    uint32_t hi = __float_as_uint(h.y), hp = __float_as_uint(h.z);
    // The camera ray again: raygen is cheap and the traversal kernel keeps
    // no per-pixel state besides the hit.
    F3 ro, rd;
    camera_ray(F, x, y, ro, rd);
    DI di;
    di.seed = tea(y * F.W + x, F.frame);
    di.org = ro;
    di.conf = 1.0f;
    F3 hv = f3(1.0f, 1.0f, 1.0f);
    di.pos = f3(0, 0, 0);
    di.nrm = f3(0, 0, 0);
    di.distance = -1.0f;
    di.inst = kMaxInstances;
    di.prim = kMaxPrimitives;
    di.mat = kInvalid;
    if (hi != kInvalid) {
      const DInst &I = F.S.inst[hi];
      di.pos = add(ro, muls(rd, h.x));
      uint32_t gp = I.first + hp;
      const ct_aabb &bb = F.aabbs[gp];
      F3 c = muls(add(f3(bb.minimum[0], bb.minimum[1], bb.minimum[2]), f3(bb.maximum[0], bb.maximum[1], bb.maximum[2])), 0.5f);
      float m[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) m[k] = I.model[k];
      c = xform3(m, c, 1.0f);
      di.nrm = cube_like_normal(normalize(sub(di.pos, c)));
      di.distance = length(sub(di.pos, ro));
      di.mat = F.prims[gp].material_index;
      di.inst = hi;
      di.prim = hp;
    } else {
      hv = mulv(hv, sky_color(F.time, F.afternoon != 0, rd));
    }
    if (!(di.distance > 0.0f)) {
      F.img[px] = make_float4(hv.x, hv.y, hv.z, 1.0f);
      if (F.active & CT_TAA_BIT) F.taa[px] = make_float4(hv.x, hv.y, hv.z, 1.0f);
    } else {
      // motion_vectors.glsl:70-89: P = prev_view_proj * geometry_T
      const float *g = F.geo_t + 16 * (size_t)di.inst;
      float pt[16];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          pt[cc * 4 + r] = F.prev_vp[0 * 4 + r] * g[cc * 4 + 0] + F.prev_vp[1 * 4 + r] * g[cc * 4 + 1] +
                           F.prev_vp[2 * 4 + r] * g[cc * 4 + 2] + F.prev_vp[3 * 4 + r] * g[cc * 4 + 3];
      float xi_x = (float)x + 0.5f, xi_y = (float)y + 0.5f;
      float ps[4];
      mat4_vec(pt, di.pos.x, di.pos.y, di.pos.z, 1.0f, ps);
      float x1 = ps[0] / ps[3], y1 = ps[1] / ps[3];
      x1 = 0.5f * x1 + 0.5f;
      y1 = 0.5f * y1 + 0.5f;
      float mvx = x1 * (float)F.W - xi_x, mvy = y1 * (float)F.H - xi_y;
      F.vel[px] = make_float2(mvx, mvy);
      Mat mat = ld_mat(F, di.mat);
      uint32_t pxx = f2u_sat(xi_x + mvx), pyy = f2u_sat(xi_y + mvy);
      Hit hit = hit_from_di(di);
      Res prev = res_init();
      uint32_t pidx = pxx + pyy * F.W;
      if (pidx < F.W * F.H) prev = ld_res(F.rA_p, F.rB_p, F.rC_p, pidx);
      float predicted = clampf(prev.M / F.rp.max_influence_from_the_past, 0.0f, 1.0f);
      float confidence = (prev.W_y > 0.0f) ? predicted : 0.0f;
      di.conf = confidence;
      RisCounts rc;
      Res r = ris(F, confidence, rd, hit, mat, di.seed, rc);
      bool needs_visibility = true;
      if (kBrdf && (F.active & CT_RIS_BRDF_BIT) &&
          ris_brdf(F, r, ro, hit, mat, rc, di.seed, stk + threadIdx.x, brdf_rays))
        needs_visibility = false;
      Vis v{true, px, 0u, 0u, 0u, 0u};
      Res rv = r;
      if (needs_visibility) {
        if (visibility_pass(F, rv, hit, v)) ris_mask |= kRisOk;
        if (v.emitted & 1u) ris_mask |= kRisRay;
      } else {
        r.W_y = __uint_as_float(kNoVisibilityPass);
      }
      ris_ray = v.emitted;
      if (prev.W_y > 0.0f) {
        // The temporal reuse rays (slots 1, 2) go into the same launch; the
        // reuse takes the prepass camera ray (rgen.glsl:261-262).
        v.slot = 1u;
        temporal_emit(F, rv, prev, pidx, rd, hit_from_di(di), mat, v);
      }
      emitted = v.emitted;
      far = v.far;
      st_res(F.rA_i, F.rB_i, F.rC_i, px, r);
      F.pidx[px] = pidx;
    }
    st_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px, di);
    F.mask[px] = ris_mask;
  }
  wave_add(ris_ray, &F.counters[kCntRis]);
  if (kBrdf) wave_add(brdf_rays, &F.counters[kCntPtBounce]);
  queue_flush(F, px, emitted, far, -1);
}
// Five waves per SIMD (at most 96 VGPRs, a few dwords spilled) run the
// VALU-heavy pass faster than four without spills.  With the BRDF
// strategy's traversal inside, five waves would spill ~150 dwords; four
// (128 VGPRs) spill 32.
#ifndef CT_RIS_BRDF_WAVES
#define CT_RIS_BRDF_WAVES 4
#endif
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) k_ris(FrameSet FS) {
  ris_pass<false>(FS);
}
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CT_RIS_BRDF_WAVES)))
k_ris_brdf(FrameSet FS) {
  ris_pass<true>(FS);
}

// RIS visibility (slot 0) and temporal reuse (slots 1, 2) with the results
// of the one any-hit launch after k_ris.  kCntTemporal counts the temporal
// queries made, not the rays traced for them.
__global__ void __launch_bounds__(kBlock) k_temporal_finish(FrameSet FS) {
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px;
  const bool mine = pass_pixel(F, blk_, px);
  uint32_t queries = 0;
  if (mine) {
    const float4 pd = F.dpos_c[px];
    if (!(pd.w > 0.0f)) {
      F.gat[(size_t)kGather * px] = pd;  // only the distance of a miss is ever read
    } else {
      const float4 pn = F.dnrm_c[px], po = F.dorg_c[px];
      const uint4 ph = F.dhit_c[px];
      const DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
      const Hit hit = hit_from_di(di);
      Res r = ld_res(F.rA_i, F.rB_i, F.rC_i, px);
      Vis v{false, px, 0u, F.mask[px], 0u, 0u};
      if (__float_as_uint(r.W_y) == kNoVisibilityPass) r.W_y = 0.0f;
      else visibility_finish(r, v.bits);
      const uint32_t pidx = F.pidx[px];
      Res prev = res_init();
      if (pidx < F.W * F.H) prev = ld_res(F.rA_p, F.rB_p, F.rC_p, pidx);
      if (prev.W_y > 0.0f) {
        const Mat mat = ld_mat(F, di.mat);
        F3 ro, rd;
        camera_ray(F, px % F.W, px / F.W, ro, rd);
        v.slot = 1u;
        v.queries = 0u;
        uint32_t seed = di.seed;
        temporal_reuse(F, r, prev, pidx, rd, hit, mat, seed, v);
        queries = v.queries;
        F.dorg_c[px].w = __uint_as_float(seed);
      }
      st_res(F.rA_i, F.rB_i, F.rC_i, px, r);
      st_gather(F.gat, px, pd, pn, po, ph.x, ph.y, r);
    }
  }
  wave_add(queries, &F.counters[kCntTemporal]);
}

// Shading pass: spatial reuse (EMIT / FINISH) then shade and store.
// The finish pass streams its pixel's own records (the emit pass's
// per-neighbour results, neighbour-major): it takes runs of consecutive
// pixels, which read them in whole cache lines.
template <bool kEmit>
__global__ void __launch_bounds__(kBlock) k_spatial(FrameSet FS) {
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px = blk_.y + threadIdx.x;
  const bool mine = kEmit ? pass_pixel(F, blk_, px) : px < F.npix;
  uint32_t emitted = 0, far = 0;
  if (mine) {
    const uint32_t x = px % F.W, y = px / F.W;
    float4 pd = F.dpos_c[px];
    if (pd.w > 0.0f) {
      DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
      F3 ray_d = normalize(sub(di.pos, di.org));
      Mat mat = ld_mat(F, di.mat);
      Hit hit = hit_from_di(di);
      Res sr = ld_res(F.rA_i, F.rB_i, F.rC_i, px);
      Vis v{kEmit, px, 0u, kEmit ? 0u : F.mask[px], 0u, 0u};
      uint32_t seed = di.seed;
      spatial_reuse<kEmit>(F, sr, di.conf, x, y, ray_d, hit, mat, seed, v);
      emitted = v.emitted;
      far = v.far;
      if (!kEmit && F.rp.indirect) {
        // The indirect term, emission and the NaN scrub follow in k_pt_spatial.
        F3 hv = f3(0, 0, 0);
        hv = add(hv, muls(sr.F, sr.W_y));
        F.img[px] = make_float4(hv.x, hv.y, hv.z, 1.0f);
        st_res(F.rA_p, F.rB_p, F.rC_p, px, sr);
        di.seed = seed;
        st_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, px, di);
      } else if (!kEmit) {
        F3 hv = f3(0, 0, 0);
        hv = add(hv, muls(sr.F, sr.W_y));
        if (F.rp.direct_emittance) hv = add(hv, mat.emission);
        if (isinf(hv.x) || isnan(hv.x)) hv.x = 0.0f;
        if (isinf(hv.y) || isnan(hv.y)) hv.y = 0.0f;
        if (isinf(hv.z) || isnan(hv.z)) hv.z = 0.0f;
        float4 fp = make_float4(hv.x, hv.y, hv.z, 1.0f);
        F.img[px] = fp;
        if (F.active & CT_TAA_BIT) F.taa[px] = fp;
        st_res(F.rA_p, F.rB_p, F.rC_p, px, sr);
        di.seed = seed;
        st_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, px, di);
      }
    } else if (!kEmit) {
      reinterpret_cast<float *>(&F.dpos_p[px])[3] = 0.0f;
    }
    if (kEmit) F.mask[px] = 0u;
  }
  if (kEmit) queue_flush(F, px, emitted, far, kCntSpatial);
}

// Appends this thread's item (when `active`) to a queue with one atomic per
// block; returns its position.  Every thread of the block calls it.
CT_D uint32_t queue_push(bool active, uint32_t *count, uint32_t *also) {
  __shared__ uint32_t s_wsum[kBlock / 64];
  __shared__ uint32_t s_base;
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  const unsigned long long b = __ballot(active);
  const uint32_t before = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) s_wsum[w] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
      const uint32_t t = s_wsum[k];
      s_wsum[k] = tot;
      tot += t;
    }
    s_base = tot ? atomicAdd(count, tot) : 0u;
    if (tot && also) atomicAdd(also, tot);
  }
  __syncthreads();
  return active ? s_base + s_wsum[w] + before : kInvalid;
}

// ReSTIR PT in wavefront form (rgen.glsl:271-315 for the first pass).
// k_pt_begin starts every hit pixel's path (generate_path + primary
// scatter) and queues its first segment; per depth, k_trace_segments finds
// each queued segment's closest voxel and k_pt_step runs one iteration of
// the path loop, queueing the next segment; k_pt_end finalizes the path
// reservoir and runs temporal path reuse.  Counters: kCntPtQueue +
// (depth & 1) holds the length of the queue of that depth.
__global__ void __launch_bounds__(kBlock) k_pt_begin(FrameSet FS) {
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px;
  const bool mine = pass_pixel(F, blk_, px);
  bool more = false;
  F3 ro, rd;
  if (mine && F.dpos_c[px].w > 0.0f) {
    const DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
    PState p;
    more = pt_begin(F, p, isect_from_di(F, di), di.seed);
    st_pstate(F.pst, F.npix, px, p);
    ro = compute_ray_origin(p.origin, p.dir);
    rd = p.dir;
  }
  const uint32_t pos = queue_push(more, &F.counters[kCntPtQueue], &F.counters[kCntPtBounce]);
  if (more) {
    F.ptq[0][pos] = F.qtag | px;
    F.seg_o[pos] = make_float4(ro.x, ro.y, ro.z, 0.0f);
    F.seg_d[pos] = make_float4(rd.x, rd.y, rd.z, 0.0f);
  }
}

// Batches of 64 queued rays for a persistent wave. The first batch is the
// wave's own index in the grid, taken without an atomic; later batches come
// from the shared head, offset by the rays the first batches cover. With a
// grid of resident blocks every first batch starts at once, so a launch with
// few rays costs about one batch, not one atomic per wave of the grid.
struct RayBatches {
  uint32_t base, count, stride;
  uint32_t *head;
  CT_D RayBatches(uint32_t count_, uint32_t *head_) : count(count_), head(head_) {
    stride = gridDim.x * (kBlock / 64u) * 64u;
    base = (blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6)) * 64u;
  }
  CT_D bool valid() const { return base < count; }
  CT_D void next() {
    uint32_t b = 0xFFFFFFFFu;
    if ((threadIdx.x & 63u) == 0u) b = atomicAdd(head, 64u) + stride;
    base = __shfl(b, 0);
  }
};

// Closest voxel of each queued path segment (intersect() semantics), one
// 64-ray fetch per wave.  kMulti: segments of several frames (scenes).
template <bool kCount, bool kMulti>
__global__ void __launch_bounds__(kBlock) k_trace_segments(FrameSet FS, uint32_t d, unsigned long long *stats) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const DFrame &F0 = FS.fb[0];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t n = F0.counters[kCntPtQueue + (d & 1u)];
  TraceCnt cnt;
  for (RayBatches rb(n, &F0.counters[kCntPtHead]); rb.valid(); rb.next()) {
    const uint32_t q = rb.base + lane;
    if (q < n) {
      const uint32_t view = kMulti ? (F0.ptq[d & 1u][q] >> kViewShift) : 0u;
      const float4 o = F0.seg_o[q], dd = F0.seg_d[q];
      float bt = kMaxDistance;
      uint32_t bi = kInvalid, bp = kInvalid;
      bool ovf = false;
      traverse_mode<kTraceBounce, kCount>(FS.fb[view].S, xyz(o), xyz(dd), kMaxDistance, kInvalid, kInvalid, bt, bi,
                                          bp, stk + tid, kBlock, cnt, ovf);
      F0.seg_hit[q] = make_float4(bt, __uint_as_float(bi), __uint_as_float(bp), 0.0f);
    }
  }
  if (kCount) wave_add_cnt(cnt, stats);
}

// One path-loop iteration for every path of queue `d & 1`.
template <bool kMulti>
__global__ void __launch_bounds__(kBlock) k_pt_step(FrameSet FS, uint32_t d) {
  const DFrame &F0 = FS.fb[0];
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = F0.counters[kCntPtQueue + (d & 1u)];
  bool more = false;
  uint32_t entry = 0;
  F3 ro, rd;
  if (q < n) {
    entry = F0.ptq[d & 1u][q];
    const DFrame &F = FS.fb[kMulti ? (entry >> kViewShift) : 0u];
    const uint32_t px = kMulti ? (entry & kSlotMask) : entry;
    PState p;
    ld_pstate(F.pst, F.npix, px, p, false);
    const float4 h = F0.seg_hit[q];
    const F3 box_o = compute_ray_origin(p.origin, p.dir);
    Isect i = segment_record(F, p.origin, box_o, p.dir, __float_as_uint(h.y), __float_as_uint(h.z));
    // only an escape or an emitting voxel reaches the path reservoir
    // (pt_handle_miss, pt_emissive): load it for those steps
    if (!i.is_hit || i.mat.emission.x > 0.0f || i.mat.emission.y > 0.0f || i.mat.emission.z > 0.0f)
      ld_pstate_res(F.pst, F.npix, px, p);
    more = pt_step(F, p, i);
    st_pstate(F.pst, F.npix, px, p, false);
    ro = compute_ray_origin(p.origin, p.dir);
    rd = p.dir;
  }
  const uint32_t pos = queue_push(more, &F0.counters[kCntPtQueue + ((d + 1u) & 1u)], &F0.counters[kCntPtBounce]);
  if (more) {
    F0.ptq[(d + 1u) & 1u][pos] = entry;
    F0.seg_o[pos] = make_float4(ro.x, ro.y, ro.z, 0.0f);
    F0.seg_d[pos] = make_float4(rd.x, rd.y, rd.z, 0.0f);
  }
}

// path_output and temporal path reuse (indirect_illumination.glsl:454-475)
// in two passes around one any-hit launch: kVisEmit queues the reuse's
// visibility rays (taken as visible), kVisFinish recomputes with their
// results and stores.  No traversal here keeps the kernels' registers low.
// kHybrid: built with the hybrid shift.  Its retraces are found by a third
// pass before them (kVisRequest, which stores nothing else) and traced by
// k_pt_retrace; both passes read the results.
template <int kVis, bool kHybrid>
__global__ void __launch_bounds__(kBlock) k_pt_end(FrameSet FS) {
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px;
  const bool mine = pass_pixel(F, blk_, px);
  PtTrace tr;
  tr.stk = nullptr;
  tr.bounce_rays = 0u;
  tr.shadow_rays = 0u;
  tr.overflow = false;
  tr.px = px;
  tr.slot = 0u;
  tr.bits = 0u;
  tr.emitted = 0u;
  tr.far = 0u;
  tr.rc_slot = 0u;
  tr.rc_prev = 0u;
  tr.rc_req = 0u;
  if (mine) {
    if (F.dpos_c[px].w > 0.0f) {
      const uint32_t x = px % F.W, y = px / F.W;
      if (kVis == kVisFinish) tr.bits = F.mask[px];
      PState p;
      ld_pstate(F.pst, F.npix, px, p);
      F3 ind;
      pt_end(p, ind);
      PRes out = p.res;
      uint32_t seed = p.seed;
      if (F.rp.pt_temporal) {
        // temporal_path_get_reprojected, indirect_illumination.glsl:29-57
        const DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
        const Isect ci = isect_from_di(F, di);
        const float2 v = F.vel[px];
        const float r1 = rnd(seed);
        const float r2 = rnd(seed);
        const uint32_t pcx = f2u_sat(((float)x + v.x) + r1), pcy = f2u_sat(((float)y + v.y) + r2);
        const uint32_t prev = min(F.npix - 1u, pcx + pcy * F.W);
        const PRes temporal = ld_pres(F.pres_t, prev);
        const Isect pi = isect_from_di_raw(F, ld_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, prev));
        tr.rc_prev = prev;
        out = pt_temporal<false, kVis, kHybrid>(F, out, temporal, ci, pi, seed, tr);
        ind = muls(out.F, out.weight);
      }
      if (kVis == kVisFinish) {
        if (has_nan_inf(ind)) ind = f3(0, 0, 0);  // rgen.glsl:293-294
        st_pres(F.pres_o, px, out);
        F.indirect[px] = make_float4(ind.x, ind.y, ind.z, 0.0f);
        F.dorg_c[px].w = __uint_as_float(seed);
      }
    }
    if (kVis == kVisEmit) F.mask[px] = 0u;
  }
  if (kVis == kVisEmit) queue_flush(F, px, tr.emitted, tr.far, kCntPtShadow);
  if (kVis == kVisRequest) {
    // the segment queues are idle after the path loop: slot k's retraces
    // go to queue k
    const bool r0 = (tr.rc_req & 1u) != 0u, r1 = (tr.rc_req & 2u) != 0u;
    const uint32_t p0 = queue_push(r0, &F.counters[kCntPtQueue], nullptr);
    if (r0) F.ptq[0][p0] = F.qtag | px;
    const uint32_t p1 = queue_push(r1, &F.counters[kCntPtQueue + 1], nullptr);
    if (r1) F.ptq[1][p1] = F.qtag | px;
  }
}

// The hybrid retraces that temporal path reuse requested, at full waves:
// queue k holds the pixels whose slot-k shift replays a source path (slot
// 0 from the pixel's own primary hit, slot 1 from the previous pixel's);
// each result replaces the request in the pixel's retrace record.
template <bool kMulti>
__global__ void __launch_bounds__(kBlock) k_pt_retrace(FrameSet FS) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const DFrame &F0 = FS.fb[0];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t n0 = F0.counters[kCntPtQueue], n = n0 + F0.counters[kCntPtQueue + 1];
  PtTrace tr;
  tr.stk = stk + tid;
  tr.bounce_rays = 0u;
  tr.shadow_rays = 0u;
  tr.overflow = false;
  for (RayBatches rb(n, &F0.counters[kCntPtHead]); rb.valid(); rb.next()) {
    const uint32_t q = rb.base + lane;
    if (q < n) {
      const uint32_t slot = q < n0 ? 0u : 1u;
      const uint32_t e = F0.ptq[slot][slot ? q - n0 : q];
      const DFrame &F = FS.fb[kMulti ? (e >> kViewShift) : 0u];
      const uint32_t px = kMulti ? (e & kSlotMask) : e;
      const float4 req = rc_rec(F, px, slot)[0];
      const uint32_t di_index = __float_as_uint(req.x);
      PRes src;
      src.init_seed = __float_as_uint(req.y);
      src.flags = __float_as_uint(req.z);
      const Isect dst = slot == 0u ? isect_from_di(F, ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, di_index))
                                   : isect_from_di_raw(F, ld_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, di_index));
      st_rc(F, px, slot, hybrid_retrace(F, dst, src, tr));
    }
  }
}

// Spatial path reuse (indirect_illumination_spatial_reuse, rgen.glsl:427-452)
// and the rest of the shading pass: indirect term, emission, NaN scrub,
// image store (rgen.glsl:449-489).
template <bool kCount, bool kHybrid>
__global__ void __launch_bounds__(kBlock) k_pt_spatial(FrameSet FS, unsigned long long *stats) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  uint32_t px;
  const bool mine = pass_pixel(F, blk_, px);
  PtTrace tr;
  tr.stk = stk + threadIdx.x;
  tr.bounce_rays = 0u;
  tr.shadow_rays = 0u;
  tr.overflow = false;
  if (mine && F.dpos_c[px].w > 0.0f) {
    const uint32_t x = px % F.W, y = px / F.W;
    const DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
    uint32_t seed = __float_as_uint(F.dorg_p[px].w);  // after the DI spatial reuse
    F3 ind;
    if (F.rp.pt_spatial) {
      const Isect ci = isect_from_di(F, di);
      const F3 cam = f3(F.inv_view[12], F.inv_view[13], F.inv_view[14]);
      const PRes fin = pt_spatial<kCount, kHybrid>(F, x, y, ci, cam, seed, tr);
      st_pres(F.pres_t, px, fin);
      ind = muls(fin.F, fin.weight);
    } else {
      // RESTIR_PT_SPATIAL_ON == 0: the first pass's reservoir is the history
      // (indirect_illumination.glsl:270-272), copied here so no pixel of the
      // first pass reads a history entry another pixel rewrites.
      if (F.rp.pt_temporal) st_pres(F.pres_t, px, ld_pres(F.pres_o, px));
      ind = xyz(F.indirect[px]);
    }
    const float4 h = F.img[px];
    F3 hv = add(f3(h.x, h.y, h.z), ind);
    if (F.rp.direct_emittance) hv = add(hv, ld_mat(F, di.mat).emission);
    if (isinf(hv.x) || isnan(hv.x)) hv.x = 0.0f;
    if (isinf(hv.y) || isnan(hv.y)) hv.y = 0.0f;
    if (isinf(hv.z) || isnan(hv.z)) hv.z = 0.0f;
    const float4 fp = make_float4(hv.x, hv.y, hv.z, 1.0f);
    F.img[px] = fp;
    if (F.active & CT_TAA_BIT) F.taa[px] = fp;
    F.dorg_p[px].w = __uint_as_float(seed);
  }
  wave_add(tr.shadow_rays, &F.counters[kCntPtShadow]);
  if (kCount) wave_add_cnt(tr.cnt, stats + 9);
}

// Path segments on caller rays (ct_trace_bounce).
template <bool kCount>
__global__ void __launch_bounds__(kBlock) k_trace_bounce(DFrame F, const float4 *ro, const float4 *rd, uint32_t n,
                                                        float4 *pos_dist, float4 *nrm_ids, float4 *wo_prim,
                                                        unsigned long long *stats) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  PtTrace tr;
  tr.stk = stk + threadIdx.x;
  tr.bounce_rays = 0u;
  tr.shadow_rays = 0u;
  tr.overflow = false;
  if (i < n) {
    const Isect h = bounce_intersect<kCount>(F, xyz(ro[i]), xyz(rd[i]), tr);
    const uint32_t inst = h.is_hit ? h.inst : kInvalid, prim = h.is_hit ? h.prim : kInvalid;
    pos_dist[i] = make_float4(h.P.x, h.P.y, h.P.z, h.distance);
    nrm_ids[i] = make_float4(h.n.x, h.n.y, h.n.z, __uint_as_float(inst));
    wo_prim[i] = make_float4(h.wo.x, h.wo.y, h.wo.z, __uint_as_float(prim));
  }
  if (kCount) wave_add_cnt(tr.cnt, stats);
}

// taa.glsl:16-108 over the dispatch region.
__global__ void __launch_bounds__(kBlock) k_taa(FrameSet FS) {
  const uint2 blk_ = FS.blk[blockIdx.x];
  const DFrame &F = FS.fb[blk_.x];
  if (!(F.active & CT_TAA_BIT)) return;
  const uint32_t tw = F.taa_w, th = F.taa_h;
  const uint32_t t = blk_.y + threadIdx.x;
  if (t >= tw * th) return;
  const uint32_t x = t % tw, y = t / tw;
  if (x >= F.W || y >= F.H) return;
  const uint32_t W = F.W, H = F.H, idx = x + y * W;
  float4 pd = F.dpos_c[idx];
  if (!(pd.w > 0.0f)) return;
  float4 pn = F.dnrm_c[idx];
  F3 dn = xyz(pn);
  uint32_t dmat = __float_as_uint(pn.w);
  float dist = pd.w;
  float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
  uint32_t count = 0;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      int64_t nx = (int64_t)x + i, ny = (int64_t)y + j;
      if (nx < 0 || nx >= (int64_t)W || ny < 0 || ny >= (int64_t)H) continue;
      uint32_t ni = (uint32_t)nx + (uint32_t)ny * W;
      float4 nn = F.dnrm_c[ni];
      float nd = F.dpos_c[ni].w;
      if (dist <= 0.0f || dmat != __float_as_uint(nn.w) || (dot(dn, xyz(nn)) < 0.906f) ||
          (fabsf(dist - nd) > 0.1f * dist))
        continue;
      ++count;
      float4 tv = F.taa[ni];
      b0 = b0 + tv.x; b1 = b1 + tv.y; b2 = b2 + tv.z; b3 = b3 + tv.w;
    }
  float fc = (float)count;
  b0 = b0 / fc; b1 = b1 / fc; b2 = b2 / fc; b3 = b3 / fc;
  float2 vv = F.vel[idx];
  uint32_t s = __float_as_uint(F.dorg_c[idx].w);
  float jx = rnd(s);
  float jy = rnd(s);
  float x1 = ((float)x + vv.x) + jx, y1 = ((float)y + vv.y) + jy;
  int32_t ix = f2i_sat(x1), iy = f2i_sat(y1);
  if (ix >= 0 && ix < (int32_t)W && iy >= 0 && iy < (int32_t)H) {
    uint32_t pi = (uint32_t)ix + (uint32_t)iy * W;
    float4 pp = F.img_prev[pi];
    float4 pnn = F.dnrm_p[pi];
    bool valid = dot(xyz(pnn), dn) >= 0.99f && __float_as_uint(pnn.w) == dmat;
    if (valid) {
      const float a = F.rp.taa_alpha, oma = 1.0f - F.rp.taa_alpha;
      b0 = a * b0 + oma * pp.x; b1 = a * b1 + oma * pp.y; b2 = a * b2 + oma * pp.z; b3 = a * b3 + oma * pp.w;
    }
  }
  F.img[idx] = make_float4(b0, b1, b2, b3);
}

__global__ void k_reset_counters(uint32_t *c, int first, int n) {
  int i = threadIdx.x;
  if (i < n) c[first + i] = 0u;
}
// Zeroes counter i for every bit i of `mask` (counters 0..31).
__global__ void k_reset_mask(uint32_t *c, uint32_t mask) {
  const uint32_t i = threadIdx.x;
  if (i < 32u && ((mask >> i) & 1u)) c[i] = 0u;
}
__global__ void k_reset_u64(unsigned long long *c, int n) {
  int i = threadIdx.x;
  if (i < n) c[i] = 0ull;
}
// The frame record travels as a kernel argument and is stored for the
// pass kernels (no host staging buffer to keep alive).
__global__ void k_store_frame(DFrame F, DFrame *dst) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&F);
  uint32_t *d = reinterpret_cast<uint32_t *>(dst);
  for (uint32_t i = threadIdx.x; i < sizeof(DFrame) / 4; i += blockDim.x) d[i] = src[i];
}

// Any-hit traversal over a ray queue; each wave fetches 64 rays at a time.
template <bool kCount>
__global__ void __launch_bounds__(kBlock) k_trace_any(DScene S, TraceAnyArgs A) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  uint32_t count = *A.count;
  if (count > A.max_rays) count = A.max_rays;
  TraceCnt cnt;
  for (RayBatches rb(count, A.head); rb.valid(); rb.next()) {
    const uint32_t q = rb.base + lane;
    if (q < count) {
      uint32_t idx = A.queue ? A.queue[q] : q;
      float4 o = A.ray_o[idx], d = A.ray_d[idx];
      uint32_t tp = A.ray_t[idx];
      float bt = 0.0f;
      uint32_t bi = 0, bp = 0;
      bool ovf = false;
      bool occ = traverse<true, kCount>(S, f3(o.x, o.y, o.z), f3(d.x, d.y, d.z), o.w, __float_as_uint(d.w), tp, bt, bi,
                                        bp, stk + tid, kBlock, cnt, ovf);
      if (occ && A.mask) atomicOr(&A.mask[idx / kRaySlots], 1u << (idx % kRaySlots));
      if (A.occ) A.occ[idx] = occ ? 1 : 0;
    }
  }
  if (kCount) {
    unsigned long long a = cnt.nodes, b = cnt.leaves, c = cnt.insts;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      c += __shfl_down(c, off);
    }
    if (lane == 0) {
      atomicAdd(&A.stats[0], a);
      atomicAdd(&A.stats[1], b);
      atomicAdd(&A.stats[2], c);
    }
  }
}

// The frame path's any-hit traversal: the queue holds tagged slots of the
// frames' ray buffers.
// Queries of the launch after k_ris: kSel 0 traces every ray in the frame's
// scene; in a frame that traces the previous set (prev_set), kSel 1 skips
// its slot-1 rays (temporal reuse's target at the previous pixel) and kSel
// 2 traces only those, in the previous set.
constexpr int kSelAll = 0, kSelCurrent = 1, kSelPrevious = 2;
template <bool kCount, bool kMulti, int kSel = kSelAll>
__global__ void __launch_bounds__(kBlock) k_trace_any_frames(FrameSet FS, uint32_t max_rays,
                                                             unsigned long long *stats) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const DFrame &F0 = FS.fb[0];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t nfar = F0.counters[kCntQueue];
  uint32_t count = nfar + F0.counters[kCntQueueNear];
  if (count > max_rays) count = max_rays;
  TraceCnt cnt;
  for (RayBatches rb(count, &F0.counters[kCntHead]); rb.valid(); rb.next()) {
    const uint32_t q = rb.base + lane;
    if (q < count) {
      // far rays from the front, then near rays from the back end down
      const uint32_t e = F0.queue[q < nfar ? q : F0.queue_cap - 1u - (q - nfar)];
      const DFrame &F = FS.fb[kMulti ? (e >> kViewShift) : 0u];
      const uint32_t idx = kMulti ? (e & kSlotMask) : e;
      const bool mine = kSel == kSelAll || ((F.prev_set != 0u && idx % kRaySlots == 1u) == (kSel == kSelPrevious));
      if (mine) {
        const float4 o = F.ray_o[idx], d = F.ray_d[idx];
        const uint32_t tp = F.ray_t[idx];
        float bt = 0.0f;
        uint32_t bi = 0, bp = 0;
        bool ovf = false;
        const bool occ = traverse<true, kCount>(kSel == kSelPrevious ? F.Sp : F.S, xyz(o), xyz(d), o.w,
                                                __float_as_uint(d.w), tp, bt, bi, bp, stk + tid, kBlock, cnt, ovf);
        if (occ) atomicOr(&F.mask[idx / kRaySlots], 1u << (idx % kRaySlots));
      }
    }
  }
  if (kCount) wave_add_cnt(cnt, stats);
}

template <bool kCount>
__global__ void __launch_bounds__(kBlock) k_trace_closest(DScene S, TraceClosestArgs A) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t i = blockIdx.x * blockDim.x + tid;
  TraceCnt cnt;
  if (i < A.n) {
    float4 o = A.ray_o[i], d = A.ray_d[i];
    float bt = A.tmax;
    uint32_t bi = kInvalid, bp = kInvalid;
    bool ovf = false;
    traverse<false, kCount>(S, f3(o.x, o.y, o.z), f3(d.x, d.y, d.z), A.tmax, kInvalid, kInvalid, bt, bi, bp, stk + tid,
                            kBlock, cnt, ovf);
    A.hit[i] = make_float4(bi != kInvalid ? bt : -1.0f, __uint_as_float(bi), __uint_as_float(bp), 0.0f);
  }
  if (kCount) {
    unsigned long long a = cnt.nodes, b = cnt.leaves, c = cnt.insts;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
      c += __shfl_down(c, off);
    }
    if (lane == 0) {
      atomicAdd(&A.stats[0], a);
      atomicAdd(&A.stats[1], b);
      atomicAdd(&A.stats[2], c);
    }
  }
}

inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

int trace_grid() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
  }
  return cus * 5;
}

// Blocks of a persistent kernel that fit on the device at once (VGPRs and the
// LDS stack bound it to 5 per CU on gfx950). A grid of at most this many
// blocks starts every wave together, so each wave's first batch of rays (its
// own index) is taken at once; blocks beyond it would only start when others
// end and hold their first batches back to the end of the launch.
template <class K>
uint32_t resident_grid(K kernel) {
  int dev = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t p;
  const int cus = hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  return (uint32_t)(per_cu * cus);
}
template <bool kCount, bool kMulti, int kSel = kSelAll>
uint32_t any_resident() {
  static const uint32_t n = resident_grid(k_trace_any_frames<kCount, kMulti, kSel>);
  return n;
}
template <bool kCount>
uint32_t any_single_resident() {
  static const uint32_t n = resident_grid(k_trace_any<kCount>);
  return n;
}
template <bool kCount, bool kMulti>
uint32_t segments_resident() {
  static const uint32_t n = resident_grid(k_trace_segments<kCount, kMulti>);
  return n;
}
template <bool kMulti>
uint32_t retrace_resident() {
  static const uint32_t n = resident_grid(k_pt_retrace<kMulti>);
  return n;
}

}  // namespace

int launch_trace_any(const DScene &s, const TraceAnyArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, a.head, 0, 1);
  const uint32_t grid = std::min<uint32_t>(std::max<uint32_t>(nblocks(a.max_rays), 1),
                                           a.stats ? any_single_resident<true>() : any_single_resident<false>());
  if (a.stats)
    hipLaunchKernelGGL(k_trace_any<true>, dim3(grid), dim3(kBlock), 0, st, s, a);
  else
    hipLaunchKernelGGL(k_trace_any<false>, dim3(grid), dim3(kBlock), 0, st, s, a);
  CT_HIP(hipGetLastError());
  return CT_OK;
}

int launch_trace_closest(const DScene &s, const TraceClosestArgs &a, hipStream_t st) {
  if (a.n == 0) return CT_OK;
  if (a.stats)
    hipLaunchKernelGGL(k_trace_closest<true>, dim3(nblocks(a.n)), dim3(kBlock), 0, st, s, a);
  else
    hipLaunchKernelGGL(k_trace_closest<false>, dim3(nblocks(a.n)), dim3(kBlock), 0, st, s, a);
  CT_HIP(hipGetLastError());
  return CT_OK;
}

}  // namespace ct

using namespace ct;

namespace {

int alloc_frame(Ctx &c) {
  FrameBufs &b = c.fb;
  const size_t np = (size_t)c.max_w * c.max_h;
  int rc;
  for (int k = 0; k < 2; ++k) {
    if ((rc = b.di_pos[k].reserve(16 * np)) || (rc = b.di_nrm[k].reserve(16 * np)) || (rc = b.di_org[k].reserve(16 * np)) ||
        (rc = b.di_hit[k].reserve(16 * np)) || (rc = b.resA[k].reserve(16 * np)) || (rc = b.resB[k].reserve(16 * np)) ||
        (rc = b.resC[k].reserve(4 * np)) || (rc = b.img[k].reserve(16 * np)))
      return rc;
  }
  if ((rc = b.velocity.reserve(8 * np)) || (rc = b.taa.reserve(16 * np)) || (rc = b.hit.reserve(16 * np)) ||
      (rc = b.pidx.reserve(4 * np)) || (rc = b.mask.reserve(4 * np)) || (rc = b.ray_o.reserve(16 * np * kRaySlots)) ||
      (rc = b.ray_d.reserve(16 * np * kRaySlots)) || (rc = b.ray_t.reserve(4 * np * kRaySlots)) ||
      (rc = b.queue.reserve(4 * np * kRaySlots)) || (rc = b.counters.reserve(64 * sizeof(uint32_t))) ||
      (rc = b.gather.reserve(16 * kGather * np)) ||
      (rc = b.prev_res_copy.reserve(64 * sizeof(unsigned long long))))
    return rc;
  return CT_OK;
}

int zero_frame(Ctx &c, hipStream_t st) {
  FrameBufs &b = c.fb;
  DevBuf *bufs[] = {&b.di_pos[0], &b.di_pos[1], &b.di_nrm[0], &b.di_nrm[1], &b.di_org[0], &b.di_org[1],
                    &b.di_hit[0], &b.di_hit[1], &b.resA[0],   &b.resA[1],   &b.resB[0],   &b.resB[1],
                    &b.resC[0],   &b.resC[1],   &b.img[0],    &b.img[1],    &b.velocity,  &b.taa,
                    &b.hit,       &b.pidx,      &b.mask,      &b.counters};
  for (DevBuf *d : bufs) CT_HIP(hipMemsetAsync(d->p, 0, d->bytes, st));
  if (b.pt_ready)
    for (DevBuf *d : {&b.pres[0], &b.pres[1], &b.indirect}) CT_HIP(hipMemsetAsync(d->p, 0, d->bytes, st));
  CT_HIP(hipStreamSynchronize(st));
  return CT_OK;
}

// Path reservoirs exist only for contexts that render indirect light; they
// start zeroed like every other frame buffer (main.cpp:234-318).
int alloc_pt(Ctx &c, hipStream_t st) {
  FrameBufs &b = c.fb;
  if (b.pt_ready) return CT_OK;
  const size_t np = (size_t)c.max_w * c.max_h;
  int rc;
  if ((rc = b.pres[0].reserve(96 * np)) || (rc = b.pres[1].reserve(96 * np)) || (rc = b.indirect.reserve(16 * np)) ||
      (rc = b.pstate.reserve(16 * kPathStateVec * np)) || (rc = b.ptq[0].reserve(4 * np)) ||
      (rc = b.ptq[1].reserve(4 * np)))
    return rc;
  for (DevBuf *d : {&b.pres[0], &b.pres[1], &b.indirect}) CT_HIP(hipMemsetAsync(d->p, 0, d->bytes, st));
  b.pt_ready = true;
  return CT_OK;
}

void pack_lights(const ct_light *L, uint32_t n, std::vector<float4> &out) {
  out.resize(3 * (size_t)std::max<uint32_t>(n, 1));
  for (uint32_t i = 0; i < n; ++i) {
    const ct_light &l = L[i];
    float tb, ib, pb;
    std::memcpy(&tb, &l.type, 4);
    std::memcpy(&ib, &l.instance_info.instance_id, 4);
    std::memcpy(&pb, &l.instance_info.primitive_id, 4);
    out[3 * i] = make_float4(l.position[0], l.position[1], l.position[2], l.size);
    out[3 * i + 1] = make_float4(l.emissive[0], l.emissive[1], l.emissive[2], tb);
    out[3 * i + 2] = make_float4(ib, pb, 0.0f, 0.0f);
  }
}

int upload_lights(Ctx &c, DevBuf &dst, const std::vector<ct_light> &src) {
  std::vector<float4> pk;
  pack_lights(src.data(), (uint32_t)src.size(), pk);
  int rc;
  if ((rc = dst.reserve(16 * pk.size()))) return rc;
  CT_HIP(hipMemcpy(dst.p, pk.data(), 16 * pk.size(), hipMemcpyHostToDevice));
  (void)c;
  return CT_OK;
}

// ------------------------------------------------------------ frame adjoint
// The frame's shading output with its final reservoirs held fixed, as a
// function of the materials' diffuse albedo d and emission e: a shaded
// pixel shows F_y W_y + I_ind + e[m_px] (e with direct emittance), where
// F_y = brdf(d[m_px]) * Le * G replays the final reservoir's light sample
// from its seed (reservoir_get_radiance, reservoir.glsl:98-112) with the
// visibility the frame found, and Le is the emission of the light voxel's
// material (a cube light; other lights keep their own).  W_y and the
// indirect term I_ind are held.  F_y is linear in d and in Le, so its
// partials are F_y re-evaluated with a unit albedo or a unit Le.  A channel
// whose F_y W_y + e is not finite (the frame's NaN scrub) has no gradient.
// Adds g . dI/dd into gd[m_px] and g . dI/de into ge[m_light] and ge[m_px].
__global__ void __launch_bounds__(kBlock) k_frame_adjoint(const DFrame *Fp, const float4 *grad_img, float *gd,
                                                          float *ge) {
  const DFrame &F = *Fp;
  const uint32_t px = blockIdx.x * kBlock + threadIdx.x;
  if (px >= F.npix) return;
  if (!(F.dpos_p[px].w > 0.0f)) return;
  const DI di = ld_di(F.dpos_p, F.dnrm_p, F.dorg_p, F.dhit_p, px);
  if (di.mat >= F.n_mats) return;
  const Res r = ld_res(F.rA_p, F.rB_p, F.rC_p, px);
  const Mat mat = ld_mat(F, di.mat);
  const float4 g = grad_img[px];
  F3 Fv = f3(0, 0, 0), Fd = f3(0, 0, 0), Fe = f3(0, 0, 0);
  uint32_t ml = kInvalid;
  bool lit = false;
  PLight l;
  if (r.W_y != 0.0f && !(r.F.x == 0.0f && r.F.y == 0.0f && r.F.z == 0.0f) && res_valid(r) &&
      light_by_type(F, r.type, r.Y, l)) {
    if (l.type == CT_LIGHT_CUBE && l.inst < F.S.n_inst) {
      const DInst &I = F.S.inst[l.inst];
      if (l.prim < I.count) ml = F.prims[I.first + l.prim].material_index;
      if (ml >= F.n_mats) ml = kInvalid;
    }
    const F3 ray_d = normalize(sub(di.pos, di.org));
    auto eval = [&](F3 d, F3 le) {
      Mat m = mat;
      m.diffuse = d;
      PLight L = l;
      L.emissive = le;
      Hit hit = hit_from_di(di);
      const Shade sh = shade_of(ray_d, hit);
      float pdf_out, G;
      uint32_t s = r.seed;
      return sampled_light(F, sh, hit, m, L, 1.0f, pdf_out, G, s, false, false, false, nullptr);
    };
    Fv = eval(mat.diffuse, l.emissive);
    Fd = eval(f3(1.0f, 1.0f, 1.0f), l.emissive);
    if (ml != kInvalid) Fe = eval(mat.diffuse, f3(1.0f, 1.0f, 1.0f));
    lit = true;
  }
  const bool emit = F.rp.direct_emittance != 0u;
  const float gc[3] = {g.x, g.y, g.z}, fv[3] = {Fv.x, Fv.y, Fv.z}, fd[3] = {Fd.x, Fd.y, Fd.z},
              fe[3] = {Fe.x, Fe.y, Fe.z}, em[3] = {mat.emission.x, mat.emission.y, mat.emission.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float v = (lit ? fv[k] * r.W_y : 0.0f) + (emit ? em[k] : 0.0f);
    if (isinf(v) || isnan(v)) continue;
    if (lit) {
      atomicAdd(&gd[3 * (size_t)di.mat + k], gc[k] * (fd[k] * r.W_y));
      if (ml != kInvalid) atomicAdd(&ge[3 * (size_t)ml + k], gc[k] * (fe[k] * r.W_y));
    }
    if (emit) atomicAdd(&ge[3 * (size_t)di.mat + k], gc[k]);
  }
}

// ------------------------------------------------------------ frame plans
int check_frame(const Ctx *c, const Accel *a, const ct_frame_params *p) {
  if (!c || !a || !p) return fail(CT_E_INVALID, "ct_frame_render: null argument");
  if (p->width == 0 || p->height == 0 || p->width > c->max_w || p->height > c->max_h)
    return fail(CT_E_INVALID, "ct_frame_render: resolution exceeds the context");
  if (p->restir.max_neighbors > CT_MAX_NEIGHBORS || p->restir.min_neighbors > CT_MAX_NEIGHBORS)
    return fail(CT_E_INVALID, "ct_frame_render: neighbour count above CT_MAX_NEIGHBORS");
  if ((p->is_active & CT_TAA_BIT) && !p->has_previous)
    return fail(CT_E_INVALID, "ct_frame_render: TAA needs a previous frame");
  if (a->set.tlas_count == 0) return fail(CT_E_STATE, "ct_frame_render: empty scene (call update_scene first)");
  if (c->device != a->device) return fail(CT_E_INVALID, "ct_frame_render: context and accel on different devices");
  const bool pt = p->restir.indirect != 0;
  if (pt && (p->restir.max_depth == 0 || p->restir.max_depth > CT_MAX_PT_DEPTH))
    return fail(CT_E_INVALID, "ct_frame_render: max_depth must be in [1, CT_MAX_PT_DEPTH]");
  if (pt && p->restir.pt_neighbor_count > CT_MAX_PT_NEIGHBORS)
    return fail(CT_E_INVALID, "ct_frame_render: pt_neighbor_count above CT_MAX_PT_NEIGHBORS");
  if (pt && p->restir.pt_shift_mapping != CT_SHIFT_RECONNECTION && p->restir.pt_shift_mapping != CT_SHIFT_HYBRID)
    return fail(CT_E_UNSUPPORTED, "ct_frame_render: only the reconnection and hybrid shifts exist");
  return CT_OK;
}

// Uploads the frame's light and motion tables and fills its record with
// the context's own buffers (a batch later points the shared ones at its
// arena).  Flips the image ping-pong.
int prepare_frame(Ctx *c, Accel *a, const ct_frame_params *p, hipStream_t st, DFrame &F) {
  int rc;
  const bool pt = p->restir.indirect != 0;
  if (pt && (rc = alloc_pt(*c, st))) return rc;
  // retrace records exist for contexts that render the hybrid shift; each
  // frame writes every record it reads, so they need no clearing
  const bool hybrid = pt && p->restir.pt_temporal && p->restir.pt_shift_mapping == CT_SHIFT_HYBRID;
  if (hybrid && (rc = c->fb.retrace.reserve(16 * 2 * kRcVec * (size_t)c->max_w * c->max_h))) return rc;
  // spatial reuse's per-neighbour results: a header and 3 float4 per neighbour
  const size_t nfield =
      1 + 3 * (size_t)std::max<uint32_t>(std::max(p->restir.max_neighbors, p->restir.min_neighbors), 1u);
  if ((rc = c->fb.spatial.reserve(16 * nfield * (size_t)c->max_w * c->max_h))) return rc;
  if ((rc = c->fb.temporal.reserve(32 * (size_t)c->max_w * c->max_h))) return rc;
  // Cube lights from the manager (packed), uploaded after each build.
  const uint32_t ncube = a->set.cube_light_count;
  DevBuf &cube = c->inst_pt;
  if (c->lights_accel != a || c->lights_generation != a->set.generation) {
    CT_HIP(hipStreamSynchronize(st));  // earlier frames may still read the old table
    std::vector<float4> pk;
    pack_lights(a->cube_lights_host.data(), ncube, pk);
    if ((rc = cube.reserve(16 * pk.size()))) return rc;
    CT_HIP(hipMemcpy(cube.p, pk.data(), 16 * pk.size(), hipMemcpyHostToDevice));
    c->lights_accel = a;
    c->lights_generation = a->set.generation;
  }
  c->cur ^= 1;
  FrameBufs &b = c->fb;
  std::memset(&F, 0, sizeof(F));
  F.W = p->width; F.H = p->height; F.npix = p->width * p->height;
  F.frame = p->frame_number; F.active = p->is_active; F.afternoon = p->is_afternoon; F.time = p->time;
  F.n_cube = ncube; F.n_point = (uint32_t)c->point_host.size(); F.n_env = (uint32_t)c->env_host.size();
  F.n_mats = (uint32_t)c->mats_host.size();
  std::memcpy(F.inv_view, p->inv_view, sizeof(F.inv_view));
  std::memcpy(F.inv_proj, p->inv_proj, sizeof(F.inv_proj));
  F.rp = p->restir;
  F.cfg = c->cfg;
  F.cfg.cube_light_count = ncube;  // the manager owns this count (accel_struct_mngr.cpp:2580)
  F.S = scene_view(*a);
  F.prev_set = ((p->is_active & CT_REMAP_BIT) && previous_distinct(*a)) ? 1u : 0u;
  F.Sp = F.prev_set ? previous_scene_view(*a) : F.S;
  F.aabbs = a->set.aabbs.as<ct_aabb>();
  F.prims = a->set.prims.as<ct_primitive>();
  F.mats = c->materials.as<float4>();
  F.cube = cube.as<float4>();
  F.point = c->point_lights.as<float4>();
  F.env = c->env_lights.as<float4>();
  F.geo_t = a->set.geo_t.as<float>();
  F.remap_prim = a->set.remap_prim.as<uint32_t>();
  F.remap_light = a->set.remap_light.as<uint32_t>();
  F.prev_first = a->set.prev_first.as<uint32_t>();
  F.remap_cap = a->set.remap_prim_cap;
  F.remap_light_cap = a->set.remap_light_cap;
  std::memcpy(F.prev_vp, p->prev_view_proj, sizeof(F.prev_vp));
  F.dpos_c = b.di_pos[0].as<float4>(); F.dnrm_c = b.di_nrm[0].as<float4>(); F.dorg_c = b.di_org[0].as<float4>();
  F.dhit_c = b.di_hit[0].as<uint4>();
  F.dpos_p = b.di_pos[1].as<float4>(); F.dnrm_p = b.di_nrm[1].as<float4>(); F.dorg_p = b.di_org[1].as<float4>();
  F.dhit_p = b.di_hit[1].as<uint4>();
  F.rA_i = b.resA[0].as<uint4>(); F.rB_i = b.resB[0].as<float4>(); F.rC_i = b.resC[0].as<float>();
  F.rA_p = b.resA[1].as<uint4>(); F.rB_p = b.resB[1].as<float4>(); F.rC_p = b.resC[1].as<float>();
  F.vel = b.velocity.as<float2>();
  F.img = b.img[c->cur].as<float4>();
  F.img_prev = b.img[c->cur ^ 1].as<float4>();
  F.taa = b.taa.as<float4>();
  F.hit = b.hit.as<float4>();
  F.pidx = b.pidx.as<uint32_t>();
  F.mask = b.mask.as<uint32_t>();
  F.ray_o = b.ray_o.as<float4>(); F.ray_d = b.ray_d.as<float4>(); F.ray_t = b.ray_t.as<uint32_t>();
  F.queue = b.queue.as<uint32_t>();
  F.queue_cap = (uint32_t)(b.queue.bytes / 4);
  F.counters = b.counters.as<uint32_t>();
  F.gat = b.gather.as<float4>();
  F.spc = b.spatial.as<float4>();
  F.tpc = b.temporal.as<float4>();
  F.pres_o = b.pres[0].as<float4>();
  F.pres_t = b.pres[1].as<float4>();
  F.indirect = b.indirect.as<float4>();
  F.pst = b.pstate.as<float4>();
  F.rc = b.retrace.as<float4>();
  F.ptq[0] = b.ptq[0].as<uint32_t>();
  F.ptq[1] = b.ptq[1].as<uint32_t>();
  // path segments reuse the primary-hit and shadow-ray buffers, which are
  // idle while paths are traced
  F.seg_o = b.ray_o.as<float4>();
  F.seg_d = b.ray_d.as<float4>();
  F.seg_hit = b.hit.as<float4>();
  F.qtag = 0u;
  F.taa_w = p->restir.taa_truncate_dispatch ? (F.W / 8) * 8 : F.W;
  F.taa_h = p->restir.taa_truncate_dispatch ? (F.H / 8) * 8 : F.H;
  F.far_ray = kFarRay;
  c->last_frame.assign(reinterpret_cast<const unsigned char *>(&F), reinterpret_cast<const unsigned char *>(&F + 1));
  c->last_frame_accel = a;
  c->last_frame_generation = a->set.generation;
  return CT_OK;
}

// Block tables of a set of frames, one after another: primary-ray tiles
// (16x16 pixels per block), pixel blocks, TAA blocks.
void build_tables(const DFrame *fr, uint32_t n, std::vector<uint2> &out, uint32_t counts[3]) {
  out.clear();
  for (uint32_t v = 0; v < n; ++v) {
    const uint32_t tiles = ((fr[v].W + 15) / 16) * ((fr[v].H + 15) / 16);
    for (uint32_t t = 0; t < tiles; ++t) out.push_back(make_uint2(v, t));
  }
  counts[0] = (uint32_t)out.size();
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t q = 0; q < fr[v].npix; q += kBlock) out.push_back(make_uint2(v, q));
  counts[1] = (uint32_t)out.size() - counts[0];
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t q = 0; q < fr[v].taa_w * fr[v].taa_h; q += kBlock) out.push_back(make_uint2(v, q));
  counts[2] = (uint32_t)out.size() - counts[0] - counts[1];
}

// Everything the launch sequence of a set of frames needs.
struct Launch {
  const DFrame *fb = nullptr;  // device frame records
  const uint2 *blk = nullptr;  // [tiles | pixels | taa] block tables
  uint32_t blocks[3] = {0, 0, 0};
  uint32_t total_px = 0;       // pixels over all frames
  uint32_t pt_depth = 0;       // deepest path of the set (0: no indirect light)
  uint32_t max_neighbors = 0;  // most spatial neighbours of the set
  uint32_t trace_blocks = 0;   // grid of the persistent traversal kernels
  bool multi = false;          // queue entries carry the frame index
  bool prev_trace = false;     // some frame traces the previous set (DFrame::prev_set)
  bool hybrid = false;         // some frame's path reuse uses the hybrid shift
  bool ris_brdf = false;       // some frame's RIS has the BRDF strategy on
  uint32_t *counters = nullptr;
  unsigned long long *stats = nullptr;
};

template <bool kCount, bool kMulti, int kSel>
void launch_any_frames(const Launch &L, FrameSet fs, uint32_t max_rays, hipStream_t st) {
  const uint32_t grid = std::min<uint32_t>(std::min<uint32_t>(std::max<uint32_t>(nblocks(max_rays), 1), L.trace_blocks),
                                           any_resident<kCount, kMulti, kSel>());
  hipLaunchKernelGGL((k_trace_any_frames<kCount, kMulti, kSel>), dim3(grid), dim3(kBlock), 0, st, fs, max_rays,
                     L.stats);
}
// The queue length and the fetch head (kCntHead) are zeroed by the caller
// before the pass that fills the queue.
template <int kSel = kSelAll>
void launch_any(const Launch &L, FrameSet fs, uint32_t max_rays, hipStream_t st) {
  if (L.stats)
    L.multi ? launch_any_frames<true, true, kSel>(L, fs, max_rays, st)
            : launch_any_frames<true, false, kSel>(L, fs, max_rays, st);
  else
    L.multi ? launch_any_frames<false, true, kSel>(L, fs, max_rays, st)
            : launch_any_frames<false, false, kSel>(L, fs, max_rays, st);
}

// The command list of App::draw (main.cpp:1436-1703) for every frame of the
// set: the same kernels, each launched once over all frames.
// `phases` selects parts of the sequence: 1 = primary rays to the path's
// temporal reuse (every read of the previous frame's buffers), 2 = spatial
// reuse and spatial path reuse (which rewrite those buffers and read
// neighbours' phase-1 results), 4 = TAA. Run in order on one stream they
// are the whole frame; ct_frame_render also runs phases 1 and 2 of two
// halves of the image on two streams, joined in between.
constexpr unsigned kPhaseFirst = 1u, kPhaseSpatial = 2u, kPhaseTaa = 4u, kPhasesAll = 7u;
template <class Mark, class SegMark>
int enqueue_frames(const Launch &L, hipStream_t st, Mark mark, SegMark seg_mark, unsigned phases = kPhasesAll) {
  int rc;
  const FrameSet tiles{L.fb, L.blk}, runs{L.fb, L.blk + L.blocks[0]}, taa{L.fb, L.blk + L.blocks[0] + L.blocks[1]};
  // the per-pixel passes walk the image by tiles (kTiled) or by runs of
  // consecutive pixels; `pix` and `nb` are their table and block count
  const FrameSet pix = kTiled ? tiles : runs;
  const uint32_t np = L.total_px, nb = kTiled ? L.blocks[0] : L.blocks[1];
  if (phases & kPhaseFirst) {
  // queue, fetch head and the ray statistics
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, L.counters, 0, 8);
  if (L.stats) hipLaunchKernelGGL(k_reset_u64, dim3(1), dim3(64), 0, st, L.stats, 12);
  if ((rc = mark())) return rc;  // kMarkStart
  if (L.stats)
    hipLaunchKernelGGL(k_primary<true>, dim3(L.blocks[0]), dim3(kBlock), 0, st, tiles, L.stats + 3);
  else
    hipLaunchKernelGGL(k_primary<false>, dim3(L.blocks[0]), dim3(kBlock), 0, st, tiles, L.stats);
  if ((rc = mark())) return rc;  // kMarkPrimary
  if (L.ris_brdf)
    hipLaunchKernelGGL(k_ris_brdf, dim3(nb), dim3(kBlock), 0, st, pix);
  else
    hipLaunchKernelGGL(k_ris, dim3(nb), dim3(kBlock), 0, st, pix);
  if ((rc = mark())) return rc;  // kMarkRis
  if (L.prev_trace) {
    // RIS and temporal reuse visibility; the target-at-previous queries of
    // frames that trace the previous set in a second pass over the queue
    launch_any<kSelCurrent>(L, pix, 3 * np, st);
    hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, st, L.counters, 1u << kCntHead);
    launch_any<kSelPrevious>(L, pix, 3 * np, st);
  } else {
    launch_any(L, pix, 3 * np, st);  // RIS and temporal reuse visibility
  }
  if ((rc = mark())) return rc;  // kMarkTraceRis
  hipLaunchKernelGGL(k_temporal_finish, dim3(nb), dim3(kBlock), 0, st, pix);
  if ((rc = mark())) return rc;  // kMarkTemporalFinish
  if (L.pt_depth) {
    // both segment queues and their head, and the ray queue and head that
    // path-reuse visibility uses after the path loop
    hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, st, L.counters,
                       (1u << kCntPtQueue) | (1u << (kCntPtQueue + 1)) | (1u << kCntPtHead) | (1u << kCntQueue) |
                           (1u << kCntQueueNear) | (1u << kCntHead));
    hipLaunchKernelGGL(k_pt_begin, dim3(nb), dim3(kBlock), 0, st, pix);
    const uint32_t grid = std::min<uint32_t>(std::min<uint32_t>(std::max<uint32_t>(nblocks(np), 1), L.trace_blocks),
                                             L.stats ? (L.multi ? segments_resident<true, true>() : segments_resident<true, false>())
                                                     : (L.multi ? segments_resident<false, true>() : segments_resident<false, false>()));
    for (uint32_t d = 0; d < L.pt_depth; ++d) {
      // the next queue and the fetch counter start at zero (already so at depth 0)
      if (d > 0)
        hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, st, L.counters,
                           (1u << (kCntPtQueue + ((d + 1) & 1u))) | (1u << kCntPtHead));
      if ((rc = seg_mark())) return rc;
      if (L.stats && L.multi)
        hipLaunchKernelGGL((k_trace_segments<true, true>), dim3(grid), dim3(kBlock), 0, st, pix, d, L.stats + 6);
      else if (L.stats)
        hipLaunchKernelGGL((k_trace_segments<true, false>), dim3(grid), dim3(kBlock), 0, st, pix, d, L.stats + 6);
      else if (L.multi)
        hipLaunchKernelGGL((k_trace_segments<false, true>), dim3(grid), dim3(kBlock), 0, st, pix, d, L.stats);
      else
        hipLaunchKernelGGL((k_trace_segments<false, false>), dim3(grid), dim3(kBlock), 0, st, pix, d, L.stats);
      if ((rc = seg_mark())) return rc;
      if (L.multi)
        hipLaunchKernelGGL(k_pt_step<true>, dim3(nblocks(np)), dim3(kBlock), 0, st, pix, d);
      else
        hipLaunchKernelGGL(k_pt_step<false>, dim3(nblocks(np)), dim3(kBlock), 0, st, pix, d);
    }
    if (L.hybrid) {
      // the retraces temporal path reuse needs: requested into the idle
      // segment queues, traced at full waves, read by both passes below
      hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, st, L.counters,
                         (1u << kCntPtQueue) | (1u << (kCntPtQueue + 1)) | (1u << kCntPtHead));
      hipLaunchKernelGGL((k_pt_end<kVisRequest, true>), dim3(nb), dim3(kBlock), 0, st, pix);
      const uint32_t rgrid =
          std::min<uint32_t>(std::min<uint32_t>(std::max<uint32_t>(nblocks(2 * np), 1), L.trace_blocks),
                             L.multi ? retrace_resident<true>() : retrace_resident<false>());
      if (L.multi)
        hipLaunchKernelGGL(k_pt_retrace<true>, dim3(rgrid), dim3(kBlock), 0, st, pix);
      else
        hipLaunchKernelGGL(k_pt_retrace<false>, dim3(rgrid), dim3(kBlock), 0, st, pix);
      hipLaunchKernelGGL((k_pt_end<kVisEmit, true>), dim3(nb), dim3(kBlock), 0, st, pix);
    } else {
      hipLaunchKernelGGL((k_pt_end<kVisEmit, false>), dim3(nb), dim3(kBlock), 0, st, pix);
    }
  }
  if ((rc = mark())) return rc;  // kMarkPtEmit
  if (L.pt_depth) launch_any(L, pix, 2 * np, st);
  if ((rc = mark())) return rc;  // kMarkTracePt
  if (L.pt_depth && L.hybrid) hipLaunchKernelGGL((k_pt_end<kVisFinish, true>), dim3(nb), dim3(kBlock), 0, st, pix);
  if (L.pt_depth && !L.hybrid) hipLaunchKernelGGL((k_pt_end<kVisFinish, false>), dim3(nb), dim3(kBlock), 0, st, pix);
  if ((rc = mark())) return rc;  // kMarkPt
  }
  if (phases & kPhaseSpatial) {
  hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, st, L.counters,
                     (1u << kCntQueue) | (1u << kCntQueueNear) | (1u << kCntHead));
  hipLaunchKernelGGL(k_spatial<true>, dim3(nb), dim3(kBlock), 0, st, pix);
  if ((rc = mark())) return rc;  // kMarkSpatialEmit
  launch_any(L, pix, 2 * L.max_neighbors * np, st);
  if ((rc = mark())) return rc;  // kMarkTraceSpatial
  hipLaunchKernelGGL(k_spatial<false>, dim3(L.blocks[1]), dim3(kBlock), 0, st, runs);
  if ((rc = mark())) return rc;  // kMarkSpatialFinish
  if (L.pt_depth) {
    if (L.stats && L.hybrid)
      hipLaunchKernelGGL((k_pt_spatial<true, true>), dim3(nb), dim3(kBlock), 0, st, pix, L.stats);
    else if (L.stats)
      hipLaunchKernelGGL((k_pt_spatial<true, false>), dim3(nb), dim3(kBlock), 0, st, pix, L.stats);
    else if (L.hybrid)
      hipLaunchKernelGGL((k_pt_spatial<false, true>), dim3(nb), dim3(kBlock), 0, st, pix, L.stats);
    else
      hipLaunchKernelGGL((k_pt_spatial<false, false>), dim3(nb), dim3(kBlock), 0, st, pix, L.stats);
  }
  if ((rc = mark())) return rc;  // kMarkPtSpatial
  }
  if (phases & kPhaseTaa) {
    if (L.blocks[2]) hipLaunchKernelGGL(k_taa, dim3(L.blocks[2]), dim3(kBlock), 0, st, taa);
    if ((rc = mark())) return rc;  // kMarkTaa
  }
  CT_HIP(hipGetLastError());
  return CT_OK;
}

// A frame of at least kSplitMinPixels runs as kSplitParts parts on as many
// streams: tile row r of 16-pixel tiles goes to part r % kSplitParts, so all
// parts hold the same mix of the image. Each tile row must be a whole number
// of 256-pixel blocks (width a multiple of 16) so that no pixel block spans
// two rows.
constexpr uint32_t kSplitMinPixels = 1u << 20;
#ifndef CT_SPLIT_PARTS
#define CT_SPLIT_PARTS 2
#endif
constexpr uint32_t kSplitParts = CT_SPLIT_PARTS;
static_assert(kSplitParts >= 2 && kSplitParts <= (uint32_t)Ctx::kMaxParts, "split parts");
bool can_split(const DFrame &F) {
  return F.npix >= kSplitMinPixels && F.W % 16u == 0u && (F.H + 15) / 16 >= kSplitParts;
}

// Block tables of the parts: each part's tiles and pixels, and for part 0
// also the whole frame's TAA blocks.
void split_tables(const DFrame &F, uint32_t parts, std::vector<uint2> *tbl, uint32_t (*cnt)[3], uint32_t *px) {
  const uint32_t tx = F.W / 16, ty = (F.H + 15) / 16, row_px = 16u * F.W;
  for (uint32_t k = 0; k < parts; ++k) {
    tbl[k].clear();
    px[k] = 0u;
  }
  for (uint32_t r = 0; r < ty; ++r)
    for (uint32_t t = r * tx; t < (r + 1) * tx; ++t) tbl[r % parts].push_back(make_uint2(0u, t));
  for (uint32_t k = 0; k < parts; ++k) cnt[k][0] = (uint32_t)tbl[k].size();
  for (uint32_t r = 0; r < ty; ++r) {
    const uint32_t q1 = std::min(F.npix, (r + 1) * row_px);
    for (uint32_t q = r * row_px; q < q1; q += kBlock) tbl[r % parts].push_back(make_uint2(0u, q));
    px[r % parts] += q1 - r * row_px;
  }
  for (uint32_t k = 0; k < parts; ++k) {
    cnt[k][1] = (uint32_t)tbl[k].size() - cnt[k][0];
    cnt[k][2] = 0u;
  }
  for (uint32_t q = 0; q < F.taa_w * F.taa_h; q += kBlock) tbl[0].push_back(make_uint2(0u, q));
  cnt[0][2] = (uint32_t)tbl[0].size() - cnt[0][0] - cnt[0][1];
}

}  // namespace

// Random replay over a frame's pixels: pixel (x, y) replays the path
// reservoir of pixel (x + dx, y + dy) (its init seed and reconnection
// length) from its own primary hit.  out: 12 u32 per pixel -- ok, length
// reached, the reservoir's path length and reconnection length, the vertex
// at that length (instance, primitive, hit xyz bits), the vertex before it
// (instance, primitive), and 0.
__global__ void __launch_bounds__(kBlock) k_pt_replay(DFrame F, int32_t dx, int32_t dy, uint32_t *out) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlock];
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= F.npix) return;
  uint32_t *o = out + 12 * (size_t)px;
  for (int k = 0; k < 12; ++k) o[k] = 0u;
  const int32_t x = (int32_t)(px % F.W) + dx, y = (int32_t)(px / F.W) + dy;
  if (!(F.dpos_c[px].w > 0.0f) || x < 0 || y < 0 || x >= (int32_t)F.W || y >= (int32_t)F.H) return;
  const uint32_t src = (uint32_t)y * F.W + (uint32_t)x;
  const PRes r = ld_pres(F.pres_o, src);
  const uint32_t rc_len = pf_rc_length(r.flags);
  o[2] = pf_path_length(r.flags);
  o[3] = rc_len;
  if (!(r.M > 0.0f) || rc_len == 0u || rc_len > F.rp.max_depth) return;
  PtTrace tr;
  tr.stk = stk + threadIdx.x;
  tr.bounce_rays = 0u;
  tr.shadow_rays = 0u;
  tr.overflow = false;
  tr.px = px;
  tr.slot = 0u;
  tr.bits = 0u;
  tr.emitted = 0u;
  tr.far = 0u;
  const DI di = ld_di(F.dpos_c, F.dnrm_c, F.dorg_c, F.dhit_c, px);
  const Replay rp = pt_replay(F, isect_from_di(F, di), r.init_seed, rc_len, tr);
  o[0] = rp.ok ? 1u : 0u;
  o[1] = rp.length;
  o[4] = rp.rc.inst;
  o[5] = rp.rc.prim;
  o[6] = __float_as_uint(rp.rc.hit.x);
  o[7] = __float_as_uint(rp.rc.hit.y);
  o[8] = __float_as_uint(rp.rc.hit.z);
  o[9] = rp.rc_prev.inst;
  o[10] = rp.rc_prev.prim;
}

extern "C" {

int ct_ctx_create(const ct_ctx_desc *d, ct_ctx **out) {
  if (!d || !out) return fail(CT_E_INVALID, "ct_ctx_create: null argument");
  *out = nullptr;
  if (d->max_width == 0 || d->max_height == 0 || (uint64_t)d->max_width * d->max_height * kRaySlots >= (1ull << 32))
    return fail(CT_E_INVALID, "ct_ctx_create: bad size");
  if (device_count_gfx950() == 0) return fail(CT_E_NODEVICE, "ct_ctx_create: no gfx950 device");
  CT_HIP(hipSetDevice(d->device));
  ct_ctx *c = new ct_ctx();
  c->device = d->device;
  c->max_w = d->max_width;
  c->max_h = d->max_height;
  int rc;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(CT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if ((rc = alloc_frame(*c)) || (rc = zero_frame(*c, c->stream))) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  for (int k = 0; k < 48; ++k) (void)hipEventCreate(&c->ev[k]);
  void *pin = nullptr;
  if (hipHostMalloc(&pin, 64 * sizeof(uint64_t), hipHostMallocDefault) == hipSuccess)
    c->counters_pinned = static_cast<uint64_t *>(pin);
  std::memset(&c->cfg, 0, sizeof(c->cfg));
  *out = c;
  return CT_OK;
}

int ct_ctx_destroy(ct_ctx *c) {
  if (!c) return fail(CT_E_INVALID, "ct_ctx_destroy: null handle");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 48; ++k)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  if (c->counters_pinned) (void)hipHostFree(c->counters_pinned);
  for (hipStream_t ps : c->part_stream)
    if (ps) {
      (void)hipStreamSynchronize(ps);
      (void)hipStreamDestroy(ps);
    }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  for (int k = 0; k < Ctx::kMaxParts; ++k) {
    if (c->ev_p1[k]) (void)hipEventDestroy(c->ev_p1[k]);
    if (c->ev_p2[k]) (void)hipEventDestroy(c->ev_p2[k]);
  }
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CT_OK;
}

int ct_ctx_append_materials(ct_ctx *c, const ct_material *m, uint32_t count, uint32_t *base) {
  if (!c || (!m && count)) return fail(CT_E_INVALID, "append_materials: null argument");
  for (uint32_t i = 0; i < count; ++i)
    if ((m[i].type & CT_MATERIAL_TYPE_MASK) != CT_MATERIAL_LAMBERTIAN)
      return fail(CT_E_UNSUPPORTED, "only Lambertian materials are supported");
  if (base) *base = (uint32_t)c->mats_host.size();
  c->mats_host.insert(c->mats_host.end(), m, m + count);
  std::vector<float4> pk(2 * std::max<size_t>(c->mats_host.size(), 1));
  for (size_t i = 0; i < c->mats_host.size(); ++i) {
    const ct_material &q = c->mats_host[i];
    float tb;
    std::memcpy(&tb, &q.type, 4);
    pk[2 * i] = make_float4(q.diffuse[0], q.diffuse[1], q.diffuse[2], tb);
    pk[2 * i + 1] = make_float4(q.emission[0], q.emission[1], q.emission[2], q.roughness);
  }
  CT_HIP(hipSetDevice(c->device));
  int rc;
  if ((rc = c->materials.reserve(16 * pk.size()))) return rc;
  CT_HIP(hipMemcpy(c->materials.p, pk.data(), 16 * pk.size(), hipMemcpyHostToDevice));
  return CT_OK;
}

int ct_ctx_set_materials(ct_ctx *c, const ct_material *m, uint32_t count) {
  if (!c) return fail(CT_E_INVALID, "set_materials: null handle");
  c->mats_host.clear();
  return ct_ctx_append_materials(c, m, count, nullptr);
}

int ct_ctx_set_point_lights(ct_ctx *c, const ct_light *l, uint32_t n) {
  if (!c || (!l && n)) return fail(CT_E_INVALID, "set_point_lights: null argument");
  c->point_host.assign(l, l + n);
  CT_HIP(hipSetDevice(c->device));
  return upload_lights(*c, c->point_lights, c->point_host);
}

int ct_ctx_set_env_lights(ct_ctx *c, const ct_light *l, uint32_t n) {
  if (!c || (!l && n)) return fail(CT_E_INVALID, "set_env_lights: null argument");
  c->env_host.assign(l, l + n);
  CT_HIP(hipSetDevice(c->device));
  return upload_lights(*c, c->env_lights, c->env_host);
}

int ct_ctx_set_light_config(ct_ctx *c, const ct_light_config *cfg) {
  if (!c || !cfg) return fail(CT_E_INVALID, "set_light_config: null argument");
  c->cfg = *cfg;
  return CT_OK;
}

int ct_ctx_get_light_config(ct_ctx *c, ct_light_config *out) {
  if (!c || !out) return fail(CT_E_INVALID, "get_light_config: null argument");
  *out = c->cfg;
  return CT_OK;
}

int ct_ctx_reset(ct_ctx *c, void *stream) {
  if (!c) return fail(CT_E_INVALID, "ct_ctx_reset: null handle");
  CT_HIP(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  c->cur = 1;
  return zero_frame(*c, st);
}

// A batch renders its views in up to kMaxGroups groups, each a contiguous
// range of views with its own ray and segment queues and counters, on its
// own stream. Groups are independent, so one group's traversal tail runs
// beside another group's shading passes.
constexpr uint32_t kMaxGroups = 4;
struct BatchGroup {
  uint32_t first = 0, count = 0;  // views [first, first + count)
  uint64_t cap_px = 0;
  DevBuf queue, counters, seg_o, seg_d, seg_hit, ptq[2], blk;
  uint32_t blocks[3] = {0, 0, 0};
  hipStream_t stream = nullptr;   // group 0 runs on the render stream
  hipEvent_t done = nullptr;
};
struct ct_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<ct_ctx *> ctxs;
  std::vector<ct_accel *> accels;
  std::vector<uint64_t> view_cap;   // pixels each view can hold
  DevBuf frames;
  BatchGroup groups[kMaxGroups];
  uint32_t ngroups = 1;
  std::vector<uint32_t> tables_key, graph_key;
  DFrame *pinned = nullptr;  // [2][n] frame records being uploaded
  hipEvent_t slot_free[2] = {};
  bool slot_used[2] = {false, false};
  int slot = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t blocks_per_cu = 5;
  float far_ray = kFarRay;              // the queue's far/near split (world units)
  uint32_t *counters_pinned = nullptr;  // [kMaxGroups][8]
  hipEvent_t t0 = nullptr, t1 = nullptr, fork = nullptr;
  hipStream_t last_stream = nullptr;
  bool pending = false;
  uint64_t primary_rays = 0;
};

namespace {
int cu_count(int device) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) == hipSuccess && p.multiProcessorCount > 0) return p.multiProcessorCount;
  return 256;
}
void drop_graph(ct_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}
}  // namespace

namespace {
// Splits the views into g contiguous groups of about equal pixel capacity
// and sizes each group's queues.
int set_groups(ct_batch *b, uint32_t g) {
  const uint32_t n = (uint32_t)b->ctxs.size();
  g = std::max(1u, std::min(std::min(g, kMaxGroups), n));
  uint64_t total = 0;
  for (uint64_t c : b->view_cap) total += c;
  uint32_t v = 0;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < g; ++k) {
    BatchGroup &G = b->groups[k];
    G.first = v;
    G.cap_px = 0;
    const uint64_t target = total * (k + 1) / g;
    // at least one view per group, and one left for each later group
    while (v < n && (v == G.first || (acc < target && n - v > g - 1 - k))) {
      acc += b->view_cap[v];
      G.cap_px += b->view_cap[v];
      ++v;
    }
    if (k == g - 1)
      while (v < n) {
        G.cap_px += b->view_cap[v];
        ++v;
      }
    G.count = v - G.first;
    int rc;
    const uint64_t cap = G.cap_px;
    if ((rc = G.queue.reserve(4 * cap * kRaySlots)) || (rc = G.counters.reserve(64 * sizeof(uint32_t))) ||
        (rc = G.seg_o.reserve(16 * cap)) || (rc = G.seg_d.reserve(16 * cap)) || (rc = G.seg_hit.reserve(16 * cap)) ||
        (rc = G.ptq[0].reserve(4 * cap)) || (rc = G.ptq[1].reserve(4 * cap)))
      return rc;
    if (k > 0 && !G.stream) {
      if (hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&G.done, hipEventDisableTiming) != hipSuccess)
        return fail(CT_E_HIP, "ct_batch: stream creation failed");
    }
  }
  if (g != b->ngroups) {
    b->ngroups = g;
    b->tables_key.clear();
    drop_graph(b);
  }
  return CT_OK;
}
}  // namespace

int ct_batch_create(uint32_t n, ct_ctx *const *ctxs, ct_accel *const *accels, ct_batch **out) {
  if (!ctxs || !accels || !out || n == 0) return fail(CT_E_INVALID, "ct_batch_create: null argument or no views");
  *out = nullptr;
  if (n > (1u << (32 - kViewShift))) return fail(CT_E_INVALID, "ct_batch_create: at most 64 views");
  uint64_t cap = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!ctxs[i] || !accels[i]) return fail(CT_E_INVALID, "ct_batch_create: null view");
    if (ctxs[i]->device != ctxs[0]->device || accels[i]->device != ctxs[0]->device)
      return fail(CT_E_INVALID, "ct_batch_create: views on different devices");
    for (uint32_t j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(CT_E_INVALID, "ct_batch_create: a context appears twice");
    const uint64_t np = (uint64_t)ctxs[i]->max_w * ctxs[i]->max_h;
    if (np * kRaySlots > kSlotMask) return fail(CT_E_INVALID, "ct_batch_create: a view above 2^22 pixels");
    cap += np;
  }
  if (cap * kRaySlots >= (1ull << 32)) return fail(CT_E_INVALID, "ct_batch_create: too many pixels");
  CT_HIP(hipSetDevice(ctxs[0]->device));
  ct_batch *b = new ct_batch();
  b->device = ctxs[0]->device;
  b->ctxs.assign(ctxs, ctxs + n);
  b->accels.assign(accels, accels + n);
  for (uint32_t i = 0; i < n; ++i) b->view_cap.push_back((uint64_t)ctxs[i]->max_w * ctxs[i]->max_h);
  int rc = CT_OK;
  void *pin = nullptr, *cpin = nullptr;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      (rc = b->frames.reserve(sizeof(DFrame) * n)) ||
      hipHostMalloc(&pin, 2 * sizeof(DFrame) * n, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&cpin, kMaxGroups * 8 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&b->slot_free[0]) != hipSuccess || hipEventCreate(&b->slot_free[1]) != hipSuccess ||
      hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess || (rc = set_groups(b, 2))) {
    b->pinned = static_cast<DFrame *>(pin);
    b->counters_pinned = static_cast<uint32_t *>(cpin);
    ct_batch_destroy(b);
    return rc ? rc : fail(CT_E_HIP, "ct_batch_create: HIP allocation failed");
  }
  b->pinned = static_cast<DFrame *>(pin);
  b->counters_pinned = static_cast<uint32_t *>(cpin);
  b->blocks_per_cu = 5;
  *out = b;
  return CT_OK;
}

int ct_batch_destroy(ct_batch *b) {
  if (!b) return fail(CT_E_INVALID, "ct_batch_destroy: null handle");
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->last_stream) (void)hipStreamSynchronize(b->last_stream);
  for (BatchGroup &G : b->groups)
    if (G.stream) (void)hipStreamSynchronize(G.stream);
  drop_graph(b);
  for (hipEvent_t e : {b->slot_free[0], b->slot_free[1], b->t0, b->t1, b->fork})
    if (e) (void)hipEventDestroy(e);
  for (BatchGroup &G : b->groups) {
    if (G.done) (void)hipEventDestroy(G.done);
    if (G.stream) (void)hipStreamDestroy(G.stream);
  }
  if (b->pinned) (void)hipHostFree(b->pinned);
  if (b->counters_pinned) (void)hipHostFree(b->counters_pinned);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CT_OK;
}

int ct_batch_render(ct_batch *b, const ct_frame_params *params, int use_graph, void *stream) {
  if (!b || !params) return fail(CT_E_INVALID, "ct_batch_render: null argument");
  const uint32_t n = (uint32_t)b->ctxs.size();
  int rc;
  for (uint32_t i = 0; i < n; ++i)
    if ((rc = check_frame(b->ctxs[i], b->accels[i], &params[i]))) return rc;
  CT_HIP(hipSetDevice(b->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : b->stream;
  // the host records of this slot were uploaded two renders ago
  if (b->slot_used[b->slot]) CT_HIP(hipEventSynchronize(b->slot_free[b->slot]));
  DFrame *rec = b->pinned + (size_t)b->slot * n;
  const uint32_t ng = b->ngroups;
  Launch L[kMaxGroups];
  for (uint32_t g = 0; g < ng; ++g) {
    BatchGroup &G = b->groups[g];
    uint32_t pt_depth = 0, max_nb = 0, px = 0;
    bool prev_trace = false, hybrid = false, ris_brdf = false;
    for (uint32_t i = G.first; i < G.first + G.count; ++i) {
      DFrame &F = rec[i];
      if ((rc = prepare_frame(b->ctxs[i], b->accels[i], &params[i], st, F))) return rc;
      F.far_ray = b->far_ray;
      F.queue = G.queue.as<uint32_t>();
      F.queue_cap = (uint32_t)(G.queue.bytes / 4);
      F.counters = G.counters.as<uint32_t>();
      F.seg_o = G.seg_o.as<float4>();
      F.seg_d = G.seg_d.as<float4>();
      F.seg_hit = G.seg_hit.as<float4>();
      F.ptq[0] = G.ptq[0].as<uint32_t>();
      F.ptq[1] = G.ptq[1].as<uint32_t>();
      F.qtag = G.count > 1 ? ((i - G.first) << kViewShift) : 0u;
      if (F.rp.indirect) pt_depth = std::max(pt_depth, F.rp.max_depth);
      max_nb = std::max(max_nb, F.rp.max_neighbors);
      px += F.npix;
      prev_trace = prev_trace || F.prev_set != 0u;
      hybrid = hybrid || (F.rp.indirect && F.rp.pt_shift_mapping == CT_SHIFT_HYBRID);
      ris_brdf = ris_brdf || (F.active & CT_RIS_BRDF_BIT) != 0u;
    }
    L[g].fb = b->frames.as<DFrame>() + G.first;
    L[g].total_px = px;
    L[g].pt_depth = pt_depth;
    L[g].max_neighbors = max_nb;
    L[g].trace_blocks = (uint32_t)cu_count(b->device) * b->blocks_per_cu;
    L[g].multi = G.count > 1;
    L[g].prev_trace = prev_trace;
    L[g].hybrid = hybrid;
    L[g].ris_brdf = ris_brdf;
    L[g].counters = G.counters.as<uint32_t>();
    L[g].stats = nullptr;
  }
  std::vector<uint32_t> tkey{ng};
  for (uint32_t i = 0; i < n; ++i) {
    tkey.push_back(rec[i].W);
    tkey.push_back(rec[i].H);
    tkey.push_back(rec[i].taa_w);
    tkey.push_back(rec[i].taa_h);
  }
  if (tkey != b->tables_key) {
    CT_HIP(hipStreamSynchronize(st));
    for (uint32_t g = 0; g < ng; ++g) {
      BatchGroup &G = b->groups[g];
      std::vector<uint2> tbl;
      build_tables(rec + G.first, G.count, tbl, G.blocks);
      if ((rc = G.blk.reserve(sizeof(uint2) * tbl.size()))) return rc;
      CT_HIP(hipMemcpy(G.blk.p, tbl.data(), sizeof(uint2) * tbl.size(), hipMemcpyHostToDevice));
    }
    b->tables_key = tkey;
    drop_graph(b);
  }
  for (uint32_t g = 0; g < ng; ++g) {
    L[g].blk = b->groups[g].blk.as<uint2>();
    std::memcpy(L[g].blocks, b->groups[g].blocks, sizeof(L[g].blocks));
  }
  CT_HIP(hipMemcpyAsync(b->frames.p, rec, sizeof(DFrame) * n, hipMemcpyHostToDevice, st));
  CT_HIP(hipEventRecord(b->slot_free[b->slot], st));
  b->slot_used[b->slot] = true;
  b->slot ^= 1;
  auto none = []() -> int { return CT_OK; };
  // Group 0 on `st`, the others on their streams between a fork and a join
  // on `st` (also inside a graph capture, where they become branches).
  auto enqueue_all = [&]() -> int {
    if (ng > 1) CT_HIP(hipEventRecord(b->fork, st));
    for (uint32_t g = 1; g < ng; ++g) CT_HIP(hipStreamWaitEvent(b->groups[g].stream, b->fork, 0));
    for (uint32_t g = 0; g < ng; ++g) {
      int e = enqueue_frames(L[g], g == 0 ? st : b->groups[g].stream, none, none);
      if (e) return e;
    }
    for (uint32_t g = 1; g < ng; ++g) {
      CT_HIP(hipEventRecord(b->groups[g].done, b->groups[g].stream));
      CT_HIP(hipStreamWaitEvent(st, b->groups[g].done, 0));
    }
    return CT_OK;
  };
  CT_HIP(hipEventRecord(b->t0, st));
  if (use_graph) {
    std::vector<uint32_t> gkey = tkey;
    for (uint32_t g = 0; g < ng; ++g)
      gkey.insert(gkey.end(), {L[g].total_px, L[g].pt_depth, L[g].max_neighbors, L[g].trace_blocks, (uint32_t)L[g].multi,
                               (uint32_t)L[g].prev_trace, (uint32_t)L[g].hybrid, (uint32_t)L[g].ris_brdf});
    if (!b->exec || gkey != b->graph_key) {
      drop_graph(b);
      CT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = enqueue_all();
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return fail(CT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      b->graph = gr;
      CT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_key = gkey;
    }
    CT_HIP(hipGraphLaunch(b->exec, st));
  } else if ((rc = enqueue_all())) {
    return rc;
  }
  CT_HIP(hipEventRecord(b->t1, st));
  for (uint32_t g = 0; g < ng; ++g)
    CT_HIP(hipMemcpyAsync(b->counters_pinned + 8 * g, b->groups[g].counters.p, 8 * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, st));
  b->pending = true;
  b->last_stream = st;
  b->primary_rays = 0;
  for (uint32_t g = 0; g < ng; ++g) b->primary_rays += L[g].total_px;
  for (uint32_t i = 0; i < n; ++i) {
    ct_ctx *c = b->ctxs[i];
    c->stats_pending = false;
    c->stats = ct_frame_stats{};
    c->stream_last = st;
    c->last_w = rec[i].W;
    c->last_h = rec[i].H;
  }
  return CT_OK;
}

int ct_batch_get_stats(ct_batch *b, ct_frame_stats *out) {
  if (!b || !out) return fail(CT_E_INVALID, "ct_batch_get_stats: null argument");
  *out = ct_frame_stats{};
  if (!b->pending) return CT_OK;
  CT_HIP(hipSetDevice(b->device));
  CT_HIP(hipStreamSynchronize(b->last_stream));
  out->primary_rays = b->primary_rays;
  for (uint32_t g = 0; g < b->ngroups; ++g) {
    const uint32_t *cnt = b->counters_pinned + 8 * g;
    out->shadow_rays_ris += cnt[kCntRis];
    out->shadow_rays_temporal += cnt[kCntTemporal];
    out->shadow_rays_spatial += cnt[kCntSpatial];
    out->shadow_rays_pt += cnt[kCntPtShadow];
    out->bounce_rays += cnt[kCntPtBounce];
  }
  out->shadow_rays = out->shadow_rays_ris + out->shadow_rays_temporal + out->shadow_rays_spatial + out->shadow_rays_pt;
  float ms = 0.0f;
  CT_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
  out->ms_total = ms;
  return CT_OK;
}

// The tuner is switched off (CT_BATCH_AUTOTUNE 0): one test run of it ended
// in a segmentation fault of the host process, inside this call, whose
// cause reading the code has not found.  Built with CT_BATCH_AUTOTUNE=1 it
// runs as described in ct_abi.h.
#ifndef CT_BATCH_AUTOTUNE
#define CT_BATCH_AUTOTUNE 0
#endif
int ct_batch_autotune(ct_batch *b, const ct_frame_params *params, uint32_t repeats, ct_batch_tuning *out) {
  if (!b || !params || repeats == 0) return fail(CT_E_INVALID, "ct_batch_autotune: bad argument");
#if !CT_BATCH_AUTOTUNE
  (void)out;
  return fail(CT_E_UNSUPPORTED, "ct_batch_autotune: switched off (a host segmentation fault inside it is unexplained)");
#else
  // the split of the views into concurrent groups, the persistent traversal
  // grid (blocks per CU; its waves per CU are 4x that), and the far/near
  // split of the ray queue (the order the traversal takes its rays in)
  const uint32_t groups[] = {1, 2, 3, 4}, per_cu[] = {3, 4, 5};
  const float far[] = {1.0f, 2.0f, 4.0f, 1e30f};
  float best = 0.0f;
  uint32_t best_g = b->ngroups, best_c = b->blocks_per_cu, tried = 0;
  float best_far = b->far_ray;
  int rc;
  for (uint32_t g : groups) {
    if (g > b->ctxs.size()) break;
    for (uint32_t c : per_cu) {
      for (float fr : far) {
        if ((rc = set_groups(b, g))) return rc;
        if (c != b->blocks_per_cu) drop_graph(b);
        b->blocks_per_cu = c;
        b->far_ray = fr;
        float tot = 0.0f;
        for (uint32_t r = 0; r < repeats + 1; ++r) {  // the first render captures the graph
          if ((rc = ct_batch_render(b, params, 1, nullptr))) return rc;
          ct_frame_stats s;
          if ((rc = ct_batch_get_stats(b, &s))) return rc;
          if (r > 0) tot += s.ms_total;
        }
        ++tried;
        const float ms = tot / (float)repeats;
        if (best == 0.0f || ms < best) {
          best = ms;
          best_g = b->ngroups;
          best_c = c;
          best_far = fr;
        }
      }
    }
  }
  if ((rc = set_groups(b, best_g))) return rc;
  b->blocks_per_cu = best_c;
  b->far_ray = best_far;
  drop_graph(b);
  if (out) {
    out->blocks_per_cu = best_c;
    out->ms = best;
    out->candidates = tried;
    out->groups = best_g;
    out->far_ray = best_far;
  }
  return CT_OK;
#endif
}

int ct_ctx_set_instrumentation(ct_ctx *c, uint32_t flags) {
  if (!c) return fail(CT_E_INVALID, "set_instrumentation: null handle");
  c->instrument = flags;
  return CT_OK;
}

int ct_frame_render(ct_ctx *c, ct_accel *a, const ct_frame_params *p, void *stream) {
  int rc;
  if ((rc = check_frame(c, a, p))) return rc;
  CT_HIP(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  DFrame F;
  if ((rc = prepare_frame(c, a, p, st, F))) return rc;
  FrameBufs &b = c->fb;
  unsigned long long *stats = (c->instrument & 2u) ? b.prev_res_copy.as<unsigned long long>() : nullptr;
  if ((rc = c->frame_dev.reserve(sizeof(DFrame)))) return rc;
  auto none = []() -> int { return CT_OK; };
  // Per-pass events and traversal counters describe one sequence: those
  // frames run whole on one stream.
  const uint32_t parts = (c->instrument & 3u) == 0u && can_split(F) ? kSplitParts : 0u;
  c->nparts = parts;
  if (parts) {
    // The parts run phase 1 at once on their streams, join, run phase 2 at
    // once, join, then TAA runs on `st`. Phase 2 reads neighbours' phase-1
    // results across parts and rewrites the previous-frame buffers that
    // phase 1 reads at reprojected pixels, hence the joins. Each part has
    // its own segment buffers: the whole-frame sequence reuses the ray and
    // hit buffers for segments, which parts in flight together cannot.
    const size_t np = (size_t)c->max_w * c->max_h;
    if (!c->ev_fork) {  // only contexts that split hold extra streams
      if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
        return fail(CT_E_HIP, "ct_frame_render: event creation failed");
      for (int k = 0; k < Ctx::kMaxParts; ++k)
        if ((k > 0 && hipStreamCreateWithFlags(&c->part_stream[k], hipStreamNonBlocking) != hipSuccess) ||
            hipEventCreateWithFlags(&c->ev_p1[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_p2[k], hipEventDisableTiming) != hipSuccess)
          return fail(CT_E_HIP, "ct_frame_render: stream creation failed");
    }
    DFrame FP[Ctx::kMaxParts];
    for (uint32_t k = 0; k < parts; ++k) {
      Ctx::PartBufs &P = c->part[k];
      if ((rc = P.queue.reserve(4 * np * kRaySlots)) || (rc = P.counters.reserve(64 * sizeof(uint32_t))) ||
          (rc = P.ptq[0].reserve(4 * np)) || (rc = P.ptq[1].reserve(4 * np)) || (rc = P.seg_o.reserve(16 * np)) ||
          (rc = P.seg_d.reserve(16 * np)) || (rc = P.seg_hit.reserve(16 * np)) ||
          (rc = P.frame.reserve(sizeof(DFrame))))
        return rc;
      FP[k] = F;
      FP[k].queue = P.queue.as<uint32_t>();
      FP[k].queue_cap = (uint32_t)(P.queue.bytes / 4);
      FP[k].counters = P.counters.as<uint32_t>();
      FP[k].ptq[0] = P.ptq[0].as<uint32_t>();
      FP[k].ptq[1] = P.ptq[1].as<uint32_t>();
      FP[k].seg_o = P.seg_o.as<float4>();
      FP[k].seg_d = P.seg_d.as<float4>();
      FP[k].seg_hit = P.seg_hit.as<float4>();
    }
    const uint32_t key[6] = {F.W, F.H, F.taa_w, F.taa_h, parts, 1u};
    if (std::memcmp(key, c->split_key, sizeof(key)) != 0) {
      std::vector<uint2> tbl[Ctx::kMaxParts];
      split_tables(F, parts, tbl, c->split_counts, c->split_px);
      for (uint32_t k = 0; k < parts; ++k)
        if ((rc = c->part[k].blk.reserve(sizeof(uint2) * tbl[k].size()))) return rc;
      CT_HIP(hipDeviceSynchronize());
      for (uint32_t k = 0; k < parts; ++k)
        CT_HIP(hipMemcpy(c->part[k].blk.p, tbl[k].data(), sizeof(uint2) * tbl[k].size(), hipMemcpyHostToDevice));
      std::memcpy(c->split_key, key, sizeof(key));
    }
    Launch L[Ctx::kMaxParts];
    for (uint32_t k = 0; k < parts; ++k) {
      hipLaunchKernelGGL(k_store_frame, dim3(1), dim3(64), 0, st, FP[k], c->part[k].frame.as<DFrame>());
      L[k].fb = c->part[k].frame.as<DFrame>();
      L[k].blk = c->part[k].blk.as<uint2>();
      std::memcpy(L[k].blocks, c->split_counts[k], sizeof(L[k].blocks));
      L[k].total_px = c->split_px[k];
      L[k].pt_depth = F.rp.indirect ? F.rp.max_depth : 0u;
      L[k].max_neighbors = F.rp.max_neighbors;
      L[k].trace_blocks = (uint32_t)trace_grid();
      L[k].multi = false;
      L[k].prev_trace = F.prev_set != 0u;
      L[k].hybrid = F.rp.indirect && F.rp.pt_shift_mapping == CT_SHIFT_HYBRID;
      L[k].ris_brdf = (F.active & CT_RIS_BRDF_BIT) != 0u;
      L[k].counters = FP[k].counters;
      L[k].stats = nullptr;
    }
    auto stream_of = [&](uint32_t k) { return k ? c->part_stream[k] : st; };
    if (c->instrument & 4u) CT_HIP(hipEventRecord(c->ev[kMarkStart], st));
    CT_HIP(hipEventRecord(c->ev_fork, st));
    for (uint32_t k = 1; k < parts; ++k) CT_HIP(hipStreamWaitEvent(stream_of(k), c->ev_fork, 0));
    for (uint32_t k = 0; k < parts; ++k)
      if ((rc = enqueue_frames(L[k], stream_of(k), none, none, kPhaseFirst))) return rc;
    for (uint32_t k = 1; k < parts; ++k) {
      CT_HIP(hipEventRecord(c->ev_p1[k], stream_of(k)));
      CT_HIP(hipStreamWaitEvent(st, c->ev_p1[k], 0));
    }
    CT_HIP(hipEventRecord(c->ev_p1[0], st));
    for (uint32_t k = 1; k < parts; ++k) CT_HIP(hipStreamWaitEvent(stream_of(k), c->ev_p1[0], 0));
    for (uint32_t k = 0; k < parts; ++k)
      if ((rc = enqueue_frames(L[k], stream_of(k), none, none, kPhaseSpatial))) return rc;
    for (uint32_t k = 1; k < parts; ++k) {
      CT_HIP(hipEventRecord(c->ev_p2[k], stream_of(k)));
      CT_HIP(hipStreamWaitEvent(st, c->ev_p2[k], 0));
    }
    if ((rc = enqueue_frames(L[0], st, none, none, kPhaseTaa))) return rc;
    if (c->instrument & 4u) CT_HIP(hipEventRecord(c->ev[kMarkTaa], st));
    c->n_events = 0;
    c->n_seg = 0;
    for (uint32_t k = 0; k < parts; ++k)  // part k's counters 128 B apart
      CT_HIP(hipMemcpyAsync(c->counters_pinned + 16 * k, FP[k].counters, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            st));
  } else {
    // the frame record and its block tables on the device
    const uint32_t key[4] = {F.W, F.H, F.taa_w, F.taa_h};
    if (std::memcmp(key, c->blk_key, sizeof(key)) != 0) {
      std::vector<uint2> tbl;
      build_tables(&F, 1, tbl, c->blk_counts);
      if ((rc = c->blk_dev.reserve(sizeof(uint2) * tbl.size()))) return rc;
      CT_HIP(hipStreamSynchronize(st));
      CT_HIP(hipMemcpy(c->blk_dev.p, tbl.data(), sizeof(uint2) * tbl.size(), hipMemcpyHostToDevice));
      std::memcpy(c->blk_key, key, sizeof(key));
    }
    hipLaunchKernelGGL(k_store_frame, dim3(1), dim3(64), 0, st, F, c->frame_dev.as<DFrame>());
    Launch L;
    L.fb = c->frame_dev.as<DFrame>();
    L.blk = c->blk_dev.as<uint2>();
    std::memcpy(L.blocks, c->blk_counts, sizeof(L.blocks));
    L.total_px = F.npix;
    L.pt_depth = F.rp.indirect ? F.rp.max_depth : 0u;
    L.max_neighbors = F.rp.max_neighbors;
    L.trace_blocks = (uint32_t)trace_grid();
    L.multi = false;
    L.prev_trace = F.prev_set != 0u;
    L.hybrid = F.rp.indirect && F.rp.pt_shift_mapping == CT_SHIFT_HYBRID;
    L.ris_brdf = (F.active & CT_RIS_BRDF_BIT) != 0u;
    L.counters = F.counters;
    L.stats = stats;
    const bool ev = (c->instrument & 1u) != 0;
    int ne = 0, nseg = kMarkCount;
    c->n_seg = 0;
    auto mark = [&]() -> int {
      if (ev) CT_HIP(hipEventRecord(c->ev[ne++], st));
      return CT_OK;
    };
    auto seg_mark = [&]() -> int {
      if (ev && nseg < 48 && c->n_seg < 32) {
        CT_HIP(hipEventRecord(c->ev[nseg], st));
        c->ev_seg[c->n_seg++] = nseg++;
      }
      return CT_OK;
    };
    if (!ev && (c->instrument & 4u)) CT_HIP(hipEventRecord(c->ev[kMarkStart], st));
    if ((rc = enqueue_frames(L, st, mark, seg_mark))) return rc;
    if (!ev && (c->instrument & 4u)) CT_HIP(hipEventRecord(c->ev[kMarkTaa], st));
    c->n_events = ne;
    // Counters for the statistics (copied once the frame is done).
    CT_HIP(hipMemcpyAsync(c->counters_pinned, F.counters, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (stats)
      CT_HIP(hipMemcpyAsync(c->counters_pinned + 8, stats, 12 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  }
  c->stats_pending = true;
  c->stats = ct_frame_stats{};
  c->stats.primary_rays = F.npix;
  c->stream_last = st;
  c->last_w = F.W;
  c->last_h = F.H;
  return CT_OK;
}

int ct_trace_bounce(ct_accel *a, uint32_t n, const float *origins, const float *directions, uint32_t *ids_out,
                    float *pos_out, float *nrm_out, float *dist_out, float *wo_out, ct_trace_stats *stats) {
  if (!a || (n && (!origins || !directions || !ids_out || !pos_out || !nrm_out || !dist_out || !wo_out)))
    return fail(CT_E_INVALID, "ct_trace_bounce: null argument");
  if (a->set.tlas_count == 0) return fail(CT_E_STATE, "ct_trace_bounce: empty scene");
  if (n == 0) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  DevBuf o, d, r0, r1, r2, st;
  int rc;
  if ((rc = o.reserve(16 * (size_t)n)) || (rc = d.reserve(16 * (size_t)n)) || (rc = r0.reserve(16 * (size_t)n)) ||
      (rc = r1.reserve(16 * (size_t)n)) || (rc = r2.reserve(16 * (size_t)n)) ||
      (rc = st.reserve(3 * sizeof(unsigned long long))))
    return rc;
  std::vector<float4> ho(n), hd(n);
  for (uint32_t i = 0; i < n; ++i) {
    ho[i] = make_float4(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2], 0.0f);
    hd[i] = make_float4(directions[3 * i], directions[3 * i + 1], directions[3 * i + 2], 0.0f);
  }
  CT_HIP(hipMemcpy(o.p, ho.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(d.p, hd.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
  CT_HIP(hipMemset(st.p, 0, 3 * sizeof(unsigned long long)));
  DFrame F;
  std::memset(&F, 0, sizeof(F));
  F.S = scene_view(*a);
  F.Sp = F.S;
  F.aabbs = a->set.aabbs.as<ct_aabb>();
  F.prims = a->set.prims.as<ct_primitive>();
  F.n_mats = 0;  // materials live in a frame context; the records here carry indices only
  if (stats)
    hipLaunchKernelGGL(k_trace_bounce<true>, dim3(nblocks(n)), dim3(kBlock), 0, a->stream, F, o.as<float4>(),
                       d.as<float4>(), n, r0.as<float4>(), r1.as<float4>(), r2.as<float4>(),
                       st.as<unsigned long long>());
  else
    hipLaunchKernelGGL(k_trace_bounce<false>, dim3(nblocks(n)), dim3(kBlock), 0, a->stream, F, o.as<float4>(),
                       d.as<float4>(), n, r0.as<float4>(), r1.as<float4>(), r2.as<float4>(), nullptr);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(a->stream));
  std::vector<float4> a0(n), a1(n), a2(n);
  CT_HIP(hipMemcpy(a0.data(), r0.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(a1.data(), r1.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(a2.data(), r2.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    pos_out[3 * i] = a0[i].x; pos_out[3 * i + 1] = a0[i].y; pos_out[3 * i + 2] = a0[i].z;
    dist_out[i] = a0[i].w;
    nrm_out[3 * i] = a1[i].x; nrm_out[3 * i + 1] = a1[i].y; nrm_out[3 * i + 2] = a1[i].z;
    std::memcpy(&ids_out[2 * i], &a1[i].w, 4);
    wo_out[3 * i] = a2[i].x; wo_out[3 * i + 1] = a2[i].y; wo_out[3 * i + 2] = a2[i].z;
    std::memcpy(&ids_out[2 * i + 1], &a2[i].w, 4);
  }
  if (stats) {
    unsigned long long v[3];
    CT_HIP(hipMemcpy(v, st.p, sizeof(v), hipMemcpyDeviceToHost));
    stats->rays = n;
    stats->nodes = v[0];
    stats->leaves = v[1];
    stats->instances = v[2];
  }
  return CT_OK;
}

int ct_pt_random_replay(ct_ctx *c, ct_accel *a, const ct_frame_params *p, int32_t dx, int32_t dy, uint32_t *out,
                        void *stream) {
  if (!c || !a || !p || !out) return fail(CT_E_INVALID, "ct_pt_random_replay: null argument");
  if (!p->restir.indirect || !c->fb.pres[0].p) return fail(CT_E_STATE, "ct_pt_random_replay: no path reservoirs");
  if (int rc0 = check_frame(c, a, p)) return rc0;
  if (p->width != c->last_w || p->height != c->last_h)
    return fail(CT_E_INVALID, "ct_pt_random_replay: parameters of another size than the last frame");
  CT_HIP(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  CT_HIP(hipDeviceSynchronize());
  DFrame F;
  int rc = prepare_frame(c, a, p, st, F);
  if (rc) return rc;
  c->cur ^= 1;  // prepare_frame advanced the image ping-pong; a replay renders nothing
  hipLaunchKernelGGL(k_pt_replay, dim3(nblocks(F.npix)), dim3(kBlock), 0, st, F, dx, dy, out);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));
  return CT_OK;
}

int ct_frame_adjoint(ct_ctx *c, ct_accel *a, const ct_frame_params *p, const float *grad_image, float *grad_diffuse,
                     float *grad_emission, void *stream) {
  if (!c || !a || !p || !grad_image || !grad_diffuse || !grad_emission)
    return fail(CT_E_INVALID, "ct_frame_adjoint: null argument");
  if (c->last_frame.size() != sizeof(DFrame) || c->last_frame_accel != a ||
      c->last_frame_generation != a->set.generation)
    return fail(CT_E_STATE, "ct_frame_adjoint: no frame rendered with this accel's current build");
  DFrame F;
  std::memcpy(&F, c->last_frame.data(), sizeof(F));
  if (F.W != p->width || F.H != p->height || F.n_mats != (uint32_t)c->mats_host.size())
    return fail(CT_E_INVALID, "ct_frame_adjoint: parameters differ from the last frame's");
  CT_HIP(hipSetDevice(c->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  int rc;
  if ((rc = c->frame_dev.reserve(sizeof(DFrame)))) return rc;
  // the record goes through a kernel argument copy, ordered on the stream
  hipLaunchKernelGGL(k_store_frame, dim3(1), dim3(64), 0, st, F, c->frame_dev.as<DFrame>());
  hipLaunchKernelGGL(k_frame_adjoint, dim3(nblocks(F.npix)), dim3(kBlock), 0, st, c->frame_dev.as<DFrame>(),
                     reinterpret_cast<const float4 *>(grad_image), grad_diffuse, grad_emission);
  CT_HIP(hipGetLastError());
  return CT_OK;
}

int ct_frame_get_stats(ct_ctx *c, ct_frame_stats *out) {
  if (!c || !out) return fail(CT_E_INVALID, "get_stats: null argument");
  CT_HIP(hipSetDevice(c->device));
  if (c->stats_pending) {
    CT_HIP(hipStreamSynchronize(c->stream_last));
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(c->counters_pinned);
    for (uint32_t h = 0; h < std::max(c->nparts, 1u); ++h, cnt += 32) {  // parts' counters 128 B apart
      c->stats.shadow_rays_ris += cnt[kCntRis];
      c->stats.shadow_rays_temporal += cnt[kCntTemporal];
      c->stats.shadow_rays_spatial += cnt[kCntSpatial];
      c->stats.shadow_rays_pt += cnt[kCntPtShadow];
      c->stats.bounce_rays += cnt[kCntPtBounce];
    }
    c->stats.shadow_rays = c->stats.shadow_rays_ris + c->stats.shadow_rays_temporal + c->stats.shadow_rays_spatial +
                           c->stats.shadow_rays_pt;
    if ((c->instrument & 5u) == 4u) {
      float v = 0.0f;
      (void)hipEventElapsedTime(&v, c->ev[kMarkStart], c->ev[kMarkTaa]);
      c->stats.ms_total = v;
    }
    if (c->instrument & 2u) {
      const unsigned long long *s = reinterpret_cast<const unsigned long long *>(c->counters_pinned + 8);
      c->stats.nodes = s[0];
      c->stats.leaves = s[1];
      c->stats.instances = s[2];
      c->stats.closest_nodes = s[3];
      c->stats.closest_leaves = s[4];
      c->stats.closest_instances = s[5];
      c->stats.pt_nodes = s[6];
      c->stats.pt_leaves = s[7];
      c->stats.pt_instances = s[8];
    }
    if ((c->instrument & 1u) && c->n_events >= kMarkCount) {
      auto ms = [&](int a, int b) -> float {
        float v = 0.0f;
        (void)hipEventElapsedTime(&v, c->ev[a], c->ev[b]);
        return v;
      };
      c->stats.ms_primary = ms(kMarkStart, kMarkPrimary);
      c->stats.ms_ris = ms(kMarkPrimary, kMarkRis);
      c->stats.ms_temporal = ms(kMarkTraceRis, kMarkTemporalFinish);
      c->stats.ms_pt = ms(kMarkTemporalFinish, kMarkPt);
      c->stats.ms_spatial = ms(kMarkPt, kMarkSpatialEmit) + ms(kMarkTraceSpatial, kMarkSpatialFinish);
      c->stats.ms_pt_spatial = ms(kMarkSpatialFinish, kMarkPtSpatial);
      c->stats.ms_taa = ms(kMarkPtSpatial, kMarkTaa);
      c->stats.ms_trace_any =
          ms(kMarkRis, kMarkTraceRis) + ms(kMarkPtEmit, kMarkTracePt) + ms(kMarkSpatialEmit, kMarkTraceSpatial);
      c->stats.ms_trace_closest = c->stats.ms_primary;
      c->stats.ms_total = ms(kMarkStart, kMarkTaa);
      c->stats.ms_pt_segments = 0.0f;
      for (int k = 0; k + 1 < c->n_seg; k += 2) c->stats.ms_pt_segments += ms(c->ev_seg[k], c->ev_seg[k + 1]);
    }
    c->stats_pending = false;
  }
  *out = c->stats;
  return CT_OK;
}

int ct_frame_read(ct_ctx *c, uint32_t which, void *dst, size_t bytes) {
  if (!c || !dst) return fail(CT_E_INVALID, "ct_frame_read: null argument");
  CT_HIP(hipSetDevice(c->device));
  CT_HIP(hipDeviceSynchronize());
  const size_t np = (size_t)c->last_w * c->last_h;
  FrameBufs &b = c->fb;
  auto need = [&](size_t s) -> bool { return bytes >= s; };
  switch (which) {
    case CT_BUF_IMAGE:
    case CT_BUF_PREV_IMAGE:
    case CT_BUF_TAA_FRAME: {
      if (!need(16 * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      const DevBuf &src = which == CT_BUF_TAA_FRAME ? b.taa : b.img[which == CT_BUF_IMAGE ? c->cur : c->cur ^ 1];
      CT_HIP(hipMemcpy(dst, src.p, 16 * np, hipMemcpyDeviceToHost));
      return CT_OK;
    }
    case CT_BUF_VELOCITY:
      if (!need(8 * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      CT_HIP(hipMemcpy(dst, b.velocity.p, 8 * np, hipMemcpyDeviceToHost));
      return CT_OK;
    case CT_BUF_DI:
    case CT_BUF_PREV_DI: {
      if (!need(sizeof(ct_di_info) * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      int k = which == CT_BUF_DI ? 0 : 1;
      std::vector<float4> P(np), N(np), O(np);
      std::vector<uint4> H(np);
      CT_HIP(hipMemcpy(P.data(), b.di_pos[k].p, 16 * np, hipMemcpyDeviceToHost));
      CT_HIP(hipMemcpy(N.data(), b.di_nrm[k].p, 16 * np, hipMemcpyDeviceToHost));
      CT_HIP(hipMemcpy(O.data(), b.di_org[k].p, 16 * np, hipMemcpyDeviceToHost));
      CT_HIP(hipMemcpy(H.data(), b.di_hit[k].p, 16 * np, hipMemcpyDeviceToHost));
      ct_di_info *o = static_cast<ct_di_info *>(dst);
      for (size_t i = 0; i < np; ++i) {
        o[i].position[0] = P[i].x; o[i].position[1] = P[i].y; o[i].position[2] = P[i].z; o[i].distance = P[i].w;
        o[i].normal[0] = N[i].x; o[i].normal[1] = N[i].y; o[i].normal[2] = N[i].z;
        std::memcpy(&o[i].mat_index, &N[i].w, 4);
        o[i].ray_origin[0] = O[i].x; o[i].ray_origin[1] = O[i].y; o[i].ray_origin[2] = O[i].z;
        std::memcpy(&o[i].seed, &O[i].w, 4);
        o[i].instance_hit.instance_id = H[i].x;
        o[i].instance_hit.primitive_id = H[i].y;
        std::memcpy(&o[i].confidence, &H[i].z, 4);
      }
      return CT_OK;
    }
    case CT_BUF_INT_RESERVOIR:
    case CT_BUF_PREV_RESERVOIR: {
      if (!need(sizeof(ct_reservoir) * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      int k = which == CT_BUF_INT_RESERVOIR ? 0 : 1;
      std::vector<uint4> A(np);
      std::vector<float4> B(np);
      std::vector<float> C(np);
      CT_HIP(hipMemcpy(A.data(), b.resA[k].p, 16 * np, hipMemcpyDeviceToHost));
      CT_HIP(hipMemcpy(B.data(), b.resB[k].p, 16 * np, hipMemcpyDeviceToHost));
      CT_HIP(hipMemcpy(C.data(), b.resC[k].p, 4 * np, hipMemcpyDeviceToHost));
      ct_reservoir *o = static_cast<ct_reservoir *>(dst);
      for (size_t i = 0; i < np; ++i) {
        o[i].Y = A[i].x; o[i].l_type = A[i].y; o[i].seed = A[i].z;
        std::memcpy(&o[i].W_y, &A[i].w, 4);
        o[i].W_sum = B[i].x; o[i].M = B[i].y; o[i].F[0] = B[i].z; o[i].F[1] = B[i].w; o[i].F[2] = C[i];
      }
      return CT_OK;
    }
    case CT_BUF_PATH_RESERVOIR:
    case CT_BUF_TEMPORAL_PATH_RESERVOIR: {
      if (!need(sizeof(ct_path_reservoir) * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      ct_path_reservoir *o = static_cast<ct_path_reservoir *>(dst);
      if (!b.pt_ready) {
        std::memset(o, 0, sizeof(ct_path_reservoir) * np);
        return CT_OK;
      }
      std::vector<float4> Q(6 * np);
      CT_HIP(hipMemcpy(Q.data(), b.pres[which == CT_BUF_PATH_RESERVOIR ? 0 : 1].p, 96 * np, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < np; ++i) {
        const float4 *q = &Q[6 * i];
        ct_path_reservoir &r = o[i];
        r.M = q[0].x; r.weight = q[0].y;
        std::memcpy(&r.path_flags, &q[0].z, 4);
        std::memcpy(&r.rc_random_seed, &q[0].w, 4);
        r.F[0] = q[1].x; r.F[1] = q[1].y; r.F[2] = q[1].z; r.light_pdf = q[1].w;
        r.cached_jacobian[0] = q[2].x; r.cached_jacobian[1] = q[2].y; r.cached_jacobian[2] = q[2].z;
        std::memcpy(&r.init_random_seed, &q[2].w, 4);
        std::memcpy(&r.rc_vertex_hit.object.instance_id, &q[3].x, 4);
        std::memcpy(&r.rc_vertex_hit.object.primitive_id, &q[3].y, 4);
        r.rc_vertex_hit.hit[0] = q[3].z; r.rc_vertex_hit.hit[1] = q[3].w; r.rc_vertex_hit.hit[2] = q[4].x;
        r.rc_vertex_wi[0] = q[4].y; r.rc_vertex_wi[1] = q[4].z; r.rc_vertex_wi[2] = q[4].w;
        r.rc_vertex_irradiance[0] = q[5].x; r.rc_vertex_irradiance[1] = q[5].y; r.rc_vertex_irradiance[2] = q[5].z;
      }
      return CT_OK;
    }
    case CT_BUF_INDIRECT: {
      if (!need(12 * np)) return fail(CT_E_INVALID, "ct_frame_read: buffer too small");
      float *o = static_cast<float *>(dst);
      if (!b.pt_ready) {
        std::memset(o, 0, 12 * np);
        return CT_OK;
      }
      std::vector<float4> Q(np);
      CT_HIP(hipMemcpy(Q.data(), b.indirect.p, 16 * np, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < np; ++i) {
        o[3 * i] = Q[i].x; o[3 * i + 1] = Q[i].y; o[3 * i + 2] = Q[i].z;
      }
      return CT_OK;
    }
    default:
      return fail(CT_E_INVALID, "ct_frame_read: unknown buffer id");
  }
}

}  // extern "C"
