// adjoint.hip -- differentiable voxel path tracer and its path-replay
// adjoint (radiance with respect to per-voxel albedo and emission).
//
// One thread per pixel of the batch of views; each loops over its
// samples in order, so the forward image is deterministic.  Path segments
// use intersect() semantics (closest voxel, traverse_mode<kTraceBounce>),
// scattering is cosine-weighted (weight = albedo).  The adjoint re-traces
// every path from its seed and keeps the path's vertices in registers
// while it walks them backwards:
//   S_{K+1} = 0,  S_k = e_k + a_k * S_{k+1}
//   dL/de_k = beta_k,  dL/da_k = beta_k * S_{k+1}.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ct_device.h"
#include "ct_internal.h"

namespace ct {
using namespace dev;

namespace {

constexpr int kBlockA = 256;
constexpr float kFar = 1e9f;  // MAX_DISTANCE, defines.glsl:30

struct AView {
  uint32_t width, height, base;  // base: global index of the view's first pixel
  uint32_t pad;
  float inv_view[16], inv_proj[16];
};

struct AArgs {
  DScene S;
  const ct_aabb *aabbs;
  const AView *views;
  const uint2 *blocks;  // per block: (view, first pixel of the view)
  uint32_t spp, max_depth, seed;
  const float *albedo;
  const float *emission;
};

struct Vertex {
  uint32_t g;  // primitive slot
  F3 beta;     // throughput arriving at the vertex
};

// Camera ray through (x + u, y + v) (get_ray_from_current_pixel,
// rgen.glsl:18-42, with the jitter drawn from the sample's stream).
CT_D void view_ray(const AView &V, uint32_t x, uint32_t y, uint32_t &seed, F3 &o, F3 &d) {
  const float jx = rnd(seed);
  const float jy = rnd(seed);
  const float pcx = (float)x + jx, pcy = (float)y + jy;
  const float ux = pcx / (float)V.width, uy = pcy / (float)V.height;
  const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
  float ov[4], t4[4], d4[4];
  mat4_vec(V.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, ov);
  mat4_vec(V.inv_proj, dx, dy, 1.0f, 1.0f, t4);
  const F3 nt = normalize(f3(t4[0], t4[1], t4[2]));
  mat4_vec(V.inv_view, nt.x, nt.y, nt.z, 0.0f, d4);
  o = f3(ov[0], ov[1], ov[2]);
  d = f3(d4[0], d4[1], d4[2]);
}

// Closest voxel of the segment from `origin` along `dir` (intersect()):
// returns the slot and the world hit position and normal.
CT_D bool segment(const AArgs &A, F3 origin, F3 dir, uint32_t *stk, uint32_t &g, F3 &pos, F3 &nrm) {
  const F3 ro = compute_ray_origin(origin, dir);
  float bt = kFar;
  uint32_t bi = kInvalid, bp = kInvalid;
  bool ovf = false;
  TraceCnt cnt;
  traverse_mode<kTraceBounce, false>(A.S, ro, dir, kFar, kInvalid, kInvalid, bt, bi, bp, stk, kBlockA, cnt, ovf);
  if (bi == kInvalid) return false;
  const DInst &I = A.S.inst[bi];
  float w[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) w[k] = I.w2o[k];
  const F3 oo = xform3(w, ro, 1.0f), od = xform3(w, dir, 0.0f);
  const F3 sinv = f3(safe_inverse(od.x), safe_inverse(od.y), safe_inverse(od.z));
  const F3 sg = f3(-gsign(od.x), -gsign(od.y), -gsign(od.z));
  g = I.first + bp;
  const ct_aabb &bb = A.aabbs[g];
  const F3 c = muls(add(f3(bb.minimum[0], bb.minimum[1], bb.minimum[2]), f3(bb.maximum[0], bb.maximum[1], bb.maximum[2])),
                    0.5f);
  float t;
  F3 on;
  intersect_box<true, false>(c, oo, od, sg, sinv, t, &on);
  pos = add(ro, muls(dir, t));
  nrm = f3(((w[0] * on.x + w[1] * on.y) + w[2] * on.z) + 0.0f * 0.0f,
           ((w[3] * on.x + w[4] * on.y) + w[5] * on.z) + 0.0f * 0.0f,
           ((w[6] * on.x + w[7] * on.y) + w[8] * on.z) + 0.0f * 0.0f);
  return true;
}

CT_D F3 albedo_of(const AArgs &A, uint32_t g) {
  return f3(A.albedo[3 * (size_t)g], A.albedo[3 * (size_t)g + 1], A.albedo[3 * (size_t)g + 2]);
}

// Traces one sample; fills the vertices (when `path` is given) and returns
// the sample's radiance.
CT_D F3 trace_sample(const AArgs &A, const AView &V, uint32_t x, uint32_t y, uint32_t seed, uint32_t *stk,
                     Vertex *path, uint32_t &nv) {
  F3 o, d;
  view_ray(V, x, y, seed, o, d);
  F3 beta = f3(1.0f, 1.0f, 1.0f), L = f3(0.0f, 0.0f, 0.0f);
  nv = 0;
  for (uint32_t k = 0; k <= A.max_depth; ++k) {
    uint32_t g;
    F3 pos, nrm;
    if (!segment(A, o, d, stk, g, pos, nrm)) break;
    if (path) {
      path[k].g = g;
      path[k].beta = beta;
    }
    nv = k + 1;
    L = add(L, muls(beta, A.emission[g]));
    if (k == A.max_depth) break;
    beta = mulv(beta, albedo_of(A, g));
    F3 sd = random_cosine_direction(seed, nrm);
    if (normal_near_zero(sd)) sd = nrm;
    d = sd;
    o = compute_ray_origin(compute_ray_origin(pos, nrm), nrm);
  }
  return L;
}

CT_D uint32_t sample_seed(uint32_t gpix, uint32_t seed, uint32_t s) { return tea(tea(gpix, seed), s); }

__global__ void __launch_bounds__(kBlockA) k_adj_forward(AArgs A, float *image) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlockA];
  const uint2 b = A.blocks[blockIdx.x];
  const AView &V = A.views[b.x];
  const uint32_t p = b.y + threadIdx.x;
  if (p >= V.width * V.height) return;
  const uint32_t x = p % V.width, y = p / V.width, gpix = V.base + p;
  F3 sum = f3(0, 0, 0);
  uint32_t nv;
  for (uint32_t s = 0; s < A.spp; ++s)
    sum = add(sum, trace_sample(A, V, x, y, sample_seed(gpix, A.seed, s), stk + threadIdx.x, nullptr, nv));
  const F3 I = divs(sum, (float)A.spp);
  image[3 * (size_t)gpix] = I.x;
  image[3 * (size_t)gpix + 1] = I.y;
  image[3 * (size_t)gpix + 2] = I.z;
}

__global__ void __launch_bounds__(kBlockA) k_adj_backward(AArgs A, const float *grad_image, float *grad_albedo,
                                                          float *grad_emission) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlockA];
  const uint2 b = A.blocks[blockIdx.x];
  const AView &V = A.views[b.x];
  const uint32_t p = b.y + threadIdx.x;
  if (p >= V.width * V.height) return;
  const uint32_t x = p % V.width, y = p / V.width, gpix = V.base + p;
  const float inv = 1.0f / (float)A.spp;
  const F3 G = f3(grad_image[3 * (size_t)gpix] * inv, grad_image[3 * (size_t)gpix + 1] * inv,
                  grad_image[3 * (size_t)gpix + 2] * inv);
  Vertex path[CT_ADJOINT_MAX_DEPTH + 1];
  for (uint32_t s = 0; s < A.spp; ++s) {
    uint32_t nv;
    trace_sample(A, V, x, y, sample_seed(gpix, A.seed, s), stk + threadIdx.x, path, nv);
    F3 S = f3(0, 0, 0);  // radiance from the vertices after k, relative to vertex k+1
    for (int k = (int)nv - 1; k >= 0; --k) {
      const uint32_t g = path[k].g;
      const F3 bk = path[k].beta;
      const float e = A.emission[g];
      atomicAdd(&grad_emission[g], dot(G, bk));
      if ((uint32_t)k < A.max_depth) {
        atomicAdd(&grad_albedo[3 * (size_t)g], G.x * bk.x * S.x);
        atomicAdd(&grad_albedo[3 * (size_t)g + 1], G.y * bk.y * S.y);
        atomicAdd(&grad_albedo[3 * (size_t)g + 2], G.z * bk.z * S.z);
        S = add(f3(e, e, e), mulv(albedo_of(A, g), S));
      } else {
        S = f3(e, e, e);
      }
    }
  }
}

struct AdjPlan {
  DevBuf views, blocks;
  uint32_t nblocks = 0;
};

// ----------------------------------------------------------------------
// Plain (non-ReSTIR) path tracing with the material callables
// (indirect_illumination_path_tracing, indirect_illumination.glsl:454-539;
// scatter by type, rcall_scatter.glsl:11-88 via sample_material,
// light.glsl:231-269).  The path is the adjoint's forward model with a
// material per voxel (PRIMITIVE.material_index): emission is collected at
// every vertex and the scattered direction comes from the material's
// callable; the throughput is multiplied by the callable's attenuation
// (Lambertian and constant medium: diffuse; metal: specular; dielectric: 1).

struct PMat {  // a MATERIAL packed for the kernel
  float4 diffuse_type;    // diffuse rgb, type bits
  float4 specular_rough;  // specular rgb, roughness
  float4 emission_ior;    // emission rgb, ior
};

struct PArgs {
  AArgs A;
  const PMat *mats;
  const ct_primitive *prims;
  uint32_t n_mats;
};

CT_D F3 reflection(F3 v, F3 n) { return sub(v, muls(n, 2.0f * dot(v, n))); }  // prng.glsl:209-211

CT_D float pow5(float x) {  // pow(x, 5.0) of reflectance, as two squarings and a product
  const float x2 = x * x;
  return x2 * x2 * x;
}

CT_D float reflectance(float cosine, float ref_idx) {  // prng.glsl:214-218 (Schlick)
  float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
  r0 = r0 * r0;
  return r0 + (1.0f - r0) * pow5(1.0f - cosine);
}

CT_D F3 refraction(F3 uv, F3 n, float eta) {  // prng.glsl:220-225
  const float cos_theta = gmin(dot(neg(uv), n), 1.0f);
  const F3 perp = muls(add(uv, muls(n, cos_theta)), eta);
  const float ls = perp.x * perp.x + perp.y * perp.y + perp.z * perp.z;
  const F3 par = muls(n, -sqrtf(fabsf(1.0f - ls)));
  return add(par, perp);
}

// random_unit_vector (prng.glsl:63-74): rejection sampling in the unit
// ball, bounded at 64 draws (then the normal; the chance of reaching the
// bound is below 1e-20).
CT_D F3 random_unit_vector(uint32_t &seed, F3 fallback) {
  for (int k = 0; k < 64; ++k) {
    const float x = -1.0f + rnd(seed) * 2.0f;
    const float y = -1.0f + rnd(seed) * 2.0f;
    const float z = -1.0f + rnd(seed) * 2.0f;
    const float l2 = x * x + y * y + z * z;
    if (l2 >= 1.0f) continue;
    return normalize(f3(x, y, z));
  }
  return fallback;
}

// The callable of the material's type: the new direction, the attenuation,
// and false when the path is absorbed (a metal ray below the surface).
CT_D bool scatter_material(const PMat &m, F3 dir, F3 nrm, uint32_t &seed, F3 &sd, F3 &att) {
  const uint32_t type = __float_as_uint(m.diffuse_type.w) & CT_MATERIAL_TYPE_MASK;
  if (type == 1u) {  // METAL
    const float r = m.specular_rough.w;
    const F3 refl = reflection(dir, nrm);
    sd = add(muls(random_cosine_direction(seed, refl), gmin(r, 1.0f)), muls(refl, 1.0f - r));
    att = f3(m.specular_rough.x, m.specular_rough.y, m.specular_rough.z);
    return dot(sd, nrm) > 0.0f;
  }
  if (type == 2u) {  // DIELECTRIC
    F3 n = nrm;
    float eta = m.emission_ior.w;
    if (dot(dir, n) > 0.0f) {
      n = neg(n);
      eta = 1.0f / eta;
    }
    const float cos_theta = gmin(dot(neg(dir), n), 1.0f);
    const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    const bool cannot_refract = eta * sin_theta > 1.0f;
    if (cannot_refract || reflectance(cos_theta, eta) > rnd(seed)) sd = reflection(dir, n);
    else sd = refraction(dir, n, eta);
    att = f3(1.0f, 1.0f, 1.0f);
    return true;
  }
  if (type == 3u) {  // CONSTANT_MEDIUM
    sd = random_unit_vector(seed, nrm);
    if (normal_near_zero(sd)) sd = nrm;
  } else {  // LAMBERTIAN
    sd = random_cosine_direction(seed, nrm);
    if (normal_near_zero(sd)) sd = nrm;
  }
  att = f3(m.diffuse_type.x, m.diffuse_type.y, m.diffuse_type.z);
  return true;
}

CT_D F3 path_sample(const PArgs &P, const AView &V, uint32_t x, uint32_t y, uint32_t seed, uint32_t *stk) {
  const AArgs &A = P.A;
  F3 o, d;
  view_ray(V, x, y, seed, o, d);
  F3 beta = f3(1.0f, 1.0f, 1.0f), L = f3(0.0f, 0.0f, 0.0f);
  for (uint32_t k = 0; k <= A.max_depth; ++k) {
    uint32_t g;
    F3 pos, nrm;
    if (!segment(A, o, d, stk, g, pos, nrm)) break;
    const uint32_t mi = P.prims[g].material_index;
    if (mi >= P.n_mats) break;  // no material: absorbs
    const PMat m = P.mats[mi];
    L = add(L, mulv(beta, f3(m.emission_ior.x, m.emission_ior.y, m.emission_ior.z)));
    if (k == A.max_depth) break;
    F3 sd, att;
    if (!scatter_material(m, d, nrm, seed, sd, att)) break;
    beta = mulv(beta, att);
    const uint32_t type = __float_as_uint(m.diffuse_type.w) & CT_MATERIAL_TYPE_MASK;
    const F3 side = (type == 0u || dot(sd, nrm) >= 0.0f) ? nrm : neg(nrm);
    d = sd;
    o = compute_ray_origin(compute_ray_origin(pos, side), side);
  }
  return L;
}

__global__ void __launch_bounds__(kBlockA) k_path_trace(PArgs P, float *image) {
  __shared__ uint32_t stk[(kStackLds > 0 ? kStackLds : 1) * kBlockA];
  const AArgs &A = P.A;
  const uint2 b = A.blocks[blockIdx.x];
  const AView &V = A.views[b.x];
  const uint32_t p = b.y + threadIdx.x;
  if (p >= V.width * V.height) return;
  const uint32_t x = p % V.width, y = p / V.width, gpix = V.base + p;
  F3 sum = f3(0, 0, 0);
  for (uint32_t s = 0; s < A.spp; ++s) sum = add(sum, path_sample(P, V, x, y, sample_seed(gpix, A.seed, s), stk + threadIdx.x));
  const F3 I = divs(sum, (float)A.spp);
  image[3 * (size_t)gpix] = I.x;
  image[3 * (size_t)gpix + 1] = I.y;
  image[3 * (size_t)gpix + 2] = I.z;
}

int plan(Accel &a, const ct_adjoint_desc *d, const ct_adjoint_view *views, hipStream_t st, AdjPlan &P, AArgs &A) {
  if (!d || !views || d->n_views == 0) return fail(CT_E_INVALID, "adjoint: no views");
  if (d->spp == 0 || d->max_depth > CT_ADJOINT_MAX_DEPTH) return fail(CT_E_INVALID, "adjoint: spp 0 or max_depth too large");
  if (a.set.tlas_count == 0) return fail(CT_E_STATE, "adjoint: empty scene");
  std::vector<AView> hv(d->n_views);
  std::vector<uint2> hb;
  uint64_t base = 0;
  for (uint32_t v = 0; v < d->n_views; ++v) {
    const ct_adjoint_view &cv = views[v];
    if (cv.width == 0 || cv.height == 0) return fail(CT_E_INVALID, "adjoint: empty view");
    AView &V = hv[v];
    V.width = cv.width;
    V.height = cv.height;
    V.base = (uint32_t)base;
    V.pad = 0;
    std::memcpy(V.inv_view, cv.inv_view, sizeof(V.inv_view));
    std::memcpy(V.inv_proj, cv.inv_proj, sizeof(V.inv_proj));
    const uint64_t np = (uint64_t)cv.width * cv.height;
    for (uint64_t p = 0; p < np; p += kBlockA) hb.push_back(make_uint2(v, (uint32_t)p));
    base += np;
    if (base >= (1ull << 31)) return fail(CT_E_INVALID, "adjoint: too many pixels");
  }
  int rc;
  if ((rc = P.views.reserve(sizeof(AView) * hv.size())) || (rc = P.blocks.reserve(sizeof(uint2) * hb.size()))) return rc;
  CT_HIP(hipMemcpyAsync(P.views.p, hv.data(), sizeof(AView) * hv.size(), hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(P.blocks.p, hb.data(), sizeof(uint2) * hb.size(), hipMemcpyHostToDevice, st));
  P.nblocks = (uint32_t)hb.size();
  A.S = scene_view(a);
  A.aabbs = a.set.aabbs.as<ct_aabb>();
  A.views = P.views.as<AView>();
  A.blocks = P.blocks.as<uint2>();
  A.spp = d->spp;
  A.max_depth = d->max_depth;
  A.seed = d->seed;
  return CT_OK;
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_adjoint_render(ct_accel *a, const ct_adjoint_desc *d, const ct_adjoint_view *views, const float *albedo,
                      const float *emission, float *image, void *stream) {
  if (!a || !albedo || !emission || !image) return fail(CT_E_INVALID, "ct_adjoint_render: null argument");
  CT_HIP(hipSetDevice(a->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : a->stream;
  AdjPlan P;
  AArgs A;
  int rc;
  if ((rc = plan(*a, d, views, st, P, A))) return rc;
  A.albedo = albedo;
  A.emission = emission;
  hipLaunchKernelGGL(k_adj_forward, dim3(P.nblocks), dim3(kBlockA), 0, st, A, image);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));  // the plan's buffers are released on return
  return CT_OK;
}

int ct_path_trace(ct_accel *a, const ct_adjoint_desc *d, const ct_adjoint_view *views, const ct_material *materials,
                  uint32_t material_count, float *image, void *stream) {
  if (!a || !image) return fail(CT_E_INVALID, "ct_path_trace: null argument");
  if (!materials && material_count) return fail(CT_E_INVALID, "ct_path_trace: null materials");
  if (!materials && !a->pt_mats.p) return fail(CT_E_STATE, "ct_path_trace: no materials uploaded yet");
  CT_HIP(hipSetDevice(a->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : a->stream;
  AdjPlan P;
  PArgs PA;
  int rc;
  if ((rc = plan(*a, d, views, st, P, PA.A))) return rc;
  if (materials) {  // pack and keep them for later calls with materials == NULL
    std::vector<PMat> pm(std::max<uint32_t>(material_count, 1));
    for (uint32_t i = 0; i < material_count; ++i) {
      const ct_material &m = materials[i];
      if ((m.type & CT_MATERIAL_TYPE_MASK) > 3u) return fail(CT_E_INVALID, "ct_path_trace: unknown material type");
      float tb;
      std::memcpy(&tb, &m.type, 4);
      pm[i].diffuse_type = make_float4(m.diffuse[0], m.diffuse[1], m.diffuse[2], tb);
      pm[i].specular_rough = make_float4(m.specular[0], m.specular[1], m.specular[2], m.roughness);
      pm[i].emission_ior = make_float4(m.emission[0], m.emission[1], m.emission[2], m.ior);
    }
    if ((rc = a->pt_mats.reserve(sizeof(PMat) * pm.size()))) return rc;
    CT_HIP(hipMemcpyAsync(a->pt_mats.p, pm.data(), sizeof(PMat) * pm.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipStreamSynchronize(st));  // `pm` is freed on return
    a->pt_mat_count = material_count;
  }
  PA.mats = a->pt_mats.as<PMat>();
  PA.prims = a->set.prims.as<ct_primitive>();
  PA.n_mats = a->pt_mat_count;
  PA.A.albedo = nullptr;
  PA.A.emission = nullptr;
  hipLaunchKernelGGL(k_path_trace, dim3(P.nblocks), dim3(kBlockA), 0, st, PA, image);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));  // the plan's buffers are released on return
  return CT_OK;
}

int ct_adjoint_backward(ct_accel *a, const ct_adjoint_desc *d, const ct_adjoint_view *views, const float *albedo,
                        const float *emission, const float *grad_image, float *grad_albedo, float *grad_emission,
                        void *stream) {
  if (!a || !albedo || !emission || !grad_image || !grad_albedo || !grad_emission)
    return fail(CT_E_INVALID, "ct_adjoint_backward: null argument");
  CT_HIP(hipSetDevice(a->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : a->stream;
  AdjPlan P;
  AArgs A;
  int rc;
  if ((rc = plan(*a, d, views, st, P, A))) return rc;
  A.albedo = albedo;
  A.emission = emission;
  hipLaunchKernelGGL(k_adj_backward, dim3(P.nblocks), dim3(kBlockA), 0, st, A, grad_image, grad_albedo,
                     grad_emission);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));
  return CT_OK;
}

}  // extern "C"
