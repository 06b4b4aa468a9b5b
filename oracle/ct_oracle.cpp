// ct_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU oracle.
//
// A plain C++ restatement of cube-tracing's per-frame ray-tracing path,
// used as the checker for the HIP library and as the CPU baseline.  Built
// with -ffp-contract=off and IEEE division/sqrt so every float operation is
// a single correctly rounded IEEE operation in the order written here.
//
// Each function cites the reference file:line it restates.  Where the
// reference leaves a result to the Vulkan driver (BVH layout, traversal
// order, equal-t ties) the rule used is stated at the function.
//
// Pinning (see tests/): RNG against the TEA/LCG published constants and
// known-answer vectors; the slab test against brute force on analytic
// boxes; BVH traversal against brute force over every primitive; LBVH
// arrays against invariants (every leaf once, parent contains child,
// Karras split = highest differing bit); ReSTIR DI radiance against the
// closed-form expectation for one unoccluded light and against invariants
// (k = 0 spatial reuse leaves the reservoir unchanged, static camera
// reprojects to the own pixel).

#include "ct_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kHalfVoxel = 0.015625f;        // HALF_VOXEL_EXTENT, shared.inl:46
constexpr float kVoxelExtent = 0.03125f;       // VOXEL_EXTENT, shared.inl:45
constexpr float kLeafPad = 0.000244140625f;    // 2^-12 conservative node padding
constexpr float kInvPi = 0.31830988618379067153776752674503f;   // shared.inl:52
constexpr float kInv4Pi = 0.079577471545947667884441881686255f; // shared.inl:54
constexpr float kMaxDistance = 1e9f;           // MAX_DISTANCE, defines.glsl:30
constexpr float kMinCosTheta = 1e-6f;          // MIN_COS_THETA, defines.glsl:31
constexpr int kStack = 96;

// ------------------------------------------------------------------ math
struct V3 { float x, y, z; };
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 muls(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
// GLSL dot/length/normalize/cross, evaluated left to right.
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) { return divs(a, length(a)); }
static inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline V3 vabs(V3 a) { return v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
static inline V3 load3(const float *p) { return v3(p[0], p[1], p[2]); }
static inline void store3(float *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
// GLSL min/max/clamp/mix/sign.
static inline float gmin(float a, float b) { return std::fmin(a, b); }
static inline float gmax(float a, float b) { return std::fmax(a, b); }
static inline float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline V3 mix3(V3 x, V3 y, float a) {
  return v3(mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a));
}
static inline float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static inline bool finite_or_nan_bad(float w) { return w < 0.0f || std::isnan(w) || std::isinf(w); }

// Column-major 4x4 (m[c*4 + r]); mat * vec sums columns left to right.
static inline void mat4_mul_vec(const float *m, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r)
    out[r] = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1] + m[2 * 4 + r] * v[2] +
             m[3 * 4 + r] * v[3];
}
static inline void mat4_mul(const float *a, const float *b, float *out) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] = a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1] +
                       a[2 * 4 + r] * b[c * 4 + 2] + a[3 * 4 + r] * b[c * 4 + 3];
}
static inline V3 xform_point(const float *m, V3 p) {
  float v[4] = {p.x, p.y, p.z, 1.0f}, o[4];
  mat4_mul_vec(m, v, o);
  return v3(o[0], o[1], o[2]);
}
static inline V3 xform_dir(const float *m, V3 p) {
  float v[4] = {p.x, p.y, p.z, 0.0f}, o[4];
  mat4_mul_vec(m, v, o);
  return v3(o[0], o[1], o[2]);
}

// Saturating float->int conversions (NaN -> 0), the GPU conversion rule.
static inline uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
static inline int32_t f2i_sat(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483648.0f) return 0x7FFFFFFF;
  if (x <= -2147483648.0f) return (int32_t)0x80000000;
  return (int32_t)x;
}

// ------------------------------------------------------------------- RNG
// random.glsl:26-40 (TEA, 16 rounds)
static inline uint32_t tea(uint32_t val0, uint32_t val1) {
  uint32_t v0 = val0, v1 = val1, s0 = 0;
  for (uint32_t n = 0; n < 16; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
// random.glsl:44-50 (Numerical Recipes LCG, 24-bit output)
static inline uint32_t lcg(uint32_t &prev) {
  prev = 1664525u * prev + 1013904223u;
  return prev & 0x00FFFFFFu;
}
// random.glsl:53-56
static inline float rnd(uint32_t &prev) { return (float)lcg(prev) / (float)0x01000000; }
// random.glsl:58-61
static inline float rnd_interval(uint32_t &prev, float mn, float mx) {
  return mn + rnd(prev) * (mx - mn);
}
// random.glsl:69-77
static inline uint32_t urnd_interval(uint32_t &prev, uint32_t mn, uint32_t mx) {
  if (mn == mx) return mn;
  if (mn > mx) return f2u_sat(rnd_interval(prev, (float)mx, (float)mn));
  return f2u_sat(rnd_interval(prev, (float)mn, (float)mx));
}

// ------------------------------------------------------------- geometry
// Box.glsl:22-23
static inline float safe_inverse(float x) { return (x == 0.0f) ? 1e12f : (1.0f / x); }

// primitives.glsl:233-257
static V3 compute_ray_origin(V3 pos, V3 normal) {
  const float origin = 1.f / 32.f, f_scale = 1.f / 65536.f, i_scale = 256.f;
  int32_t ix = f2i_sat(normal.x * i_scale), iy = f2i_sat(normal.y * i_scale),
          iz = f2i_sat(normal.z * i_scale);
  auto bump = [](float p, int32_t off) {
    uint32_t bits;
    std::memcpy(&bits, &p, 4);
    bits += (uint32_t)(p < 0.0f ? -off : off);
    float r;
    std::memcpy(&r, &bits, 4);
    return r;
  };
  V3 i_pos = v3(bump(pos.x, ix), bump(pos.y, iy), bump(pos.z, iz));
  V3 f_off = muls(normal, f_scale);
  return v3(std::fabs(pos.x) < origin ? pos.x + f_off.x : i_pos.x,
            std::fabs(pos.y) < origin ? pos.y + f_off.y : i_pos.y,
            std::fabs(pos.z) < origin ? pos.z + f_off.z : i_pos.z);
}

// Multiplication by Box.rotation = mat3(1): rotation * v and v * rotation
// are evaluated as the matrix products the shader writes.
static inline V3 rot_mul(V3 v) {
  return v3(1.0f * v.x + 0.0f * v.y + 0.0f * v.z, 0.0f * v.x + 1.0f * v.y + 0.0f * v.z,
            0.0f * v.x + 0.0f * v.y + 1.0f * v.z);
}
static inline V3 mul_rot(V3 v) {
  return v3(v.x * 1.0f + v.y * 0.0f + v.z * 0.0f, v.x * 0.0f + v.y * 1.0f + v.z * 0.0f,
            v.x * 0.0f + v.y * 0.0f + v.z * 1.0f);
}

// primitives.glsl:273-355 (jcgt 2018 ray-box), rotation = identity.
static bool intersect_box(V3 center, V3 radius, V3 inv_radius, V3 ro, V3 rd,
                          float &distance, V3 &normal, bool can_start,
                          bool oriented, V3 inv_dir) {
  V3 o = rot_mul(sub(ro, center));
  V3 d = oriented ? mul_rot(rd) : rd;
  float winding = 1.0f;
  if (can_start) {
    V3 rb = mulv(vabs(o), inv_radius);
    winding = (gmax(rb.x, gmax(rb.y, rb.z)) < 1.0f) ? -1.0f : 1.0f;
  }
  V3 sgn = v3(-gsign(d.x), -gsign(d.y), -gsign(d.z));
  V3 dtp = sub(mulv(muls(radius, winding), sgn), o);
  if (oriented)
    dtp = v3(dtp.x / d.x, dtp.y / d.y, dtp.z / d.z);
  else
    dtp = mulv(dtp, inv_dir);
  bool tx = (dtp.x >= 0.0f) && (std::fabs(o.y + d.y * dtp.x) < radius.y) &&
            (std::fabs(o.z + d.z * dtp.x) < radius.z);
  bool ty = (dtp.y >= 0.0f) && (std::fabs(o.z + d.z * dtp.y) < radius.z) &&
            (std::fabs(o.x + d.x * dtp.y) < radius.x);
  bool tz = (dtp.z >= 0.0f) && (std::fabs(o.x + d.x * dtp.z) < radius.x) &&
            (std::fabs(o.y + d.y * dtp.z) < radius.y);
  sgn = tx ? v3(sgn.x, 0.0f, 0.0f)
           : (ty ? v3(0.0f, sgn.y, 0.0f) : v3(0.0f, 0.0f, tz ? sgn.z : 0.0f));
  distance = (sgn.x != 0.0f) ? dtp.x : ((sgn.y != 0.0f) ? dtp.y : dtp.z);
  normal = oriented ? rot_mul(sgn) : sgn;
  return (sgn.x != 0.0f) || (sgn.y != 0.0f) || (sgn.z != 0.0f);
}

// primitives.glsl:499-507
static V3 cube_like_normal(V3 n) {
  V3 a = vabs(n);
  float mc = gmax(gmax(a.x, a.y), a.z);
  return (mc == a.x) ? v3(gsign(n.x), 0, 0)
                     : (mc == a.y) ? v3(0, gsign(n.y), 0) : v3(0, 0, gsign(n.z));
}

// mat.glsl:111-135
static V3 sky_grad(V3 dir, V3 a, V3 b) {
  V3 u = normalize(dir);
  float t = 0.5f * (u.y + 1.0f);
  return mix3(a, b, t);
}
static V3 day_color(V3 d) { return sky_grad(d, v3(1.0f, 1.0f, 1.0f), v3(0.5f, 0.7f, 1.0f)); }
static V3 night_color(V3 d) { return sky_grad(d, v3(0.2f, 0.2f, 0.3f), v3(0.2f, 0.2f, 0.5f)); }
static V3 sunset_color(V3 d) { return sky_grad(d, v3(0.8f, 0.4f, 0.0f), v3(1.0f, 0.4f, 0.0f)); }
static V3 sunrise_color(V3 d) { return sky_grad(d, v3(0.7f, 0.35f, 0.0f), v3(0.8f, 0.4f, 0.0f)); }
// mat.glsl:137-202 (the blends use `t`, the time, as written).
static V3 sky_color(float t, bool afternoon, V3 dir) {
  const float night_duration = 0.5f, sunrise_duration = 0.10f,
              full_daylight_transition = 0.10f, full_night_transition = 0.10f;
  float sunrise_end = night_duration + sunrise_duration;
  float sunrise_to_daylight = sunrise_end + full_daylight_transition;
  float zenith_transition = 1.0f - sunrise_to_daylight;
  float daylight_end = 1.0f - zenith_transition;
  float sunset_start = 1.0f - zenith_transition - sunrise_duration;
  float sunset_end = 1.0f - zenith_transition - sunrise_duration - full_night_transition;
  if (afternoon) {
    if (t < night_duration) return night_color(dir);
    if (t < sunrise_end) return mix3(night_color(dir), sunrise_color(dir), t);
    if (t < sunrise_to_daylight) return mix3(sunrise_color(dir), sunset_color(dir), t);
    return mix3(sunset_color(dir), day_color(dir), t);
  }
  if (t > daylight_end) return mix3(sunrise_color(dir), day_color(dir), t);
  if (t > sunset_start) return mix3(sunset_color(dir), sunrise_color(dir), t);
  if (t > sunset_end) return mix3(night_color(dir), sunset_color(dir), t);
  return night_color(dir);
}

// light.glsl:78-80 (Rec.709)
static inline float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// ------------------------------------------------------------------ LBVH
static inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
static inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (expand_bits(x) << 2) | (expand_bits(y) << 1) | expand_bits(z);
}
static inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

struct Bvh {
  uint32_t n = 0;
  std::vector<uint32_t> keys, ids, child, parent;
  std::vector<float> box;          // per internal node: child0 lo,hi, child1 lo,hi
  std::vector<float> leaf_center;  // sorted order (BLAS)
  uint32_t root_ref = kLeaf;
  uint32_t depth = 0;
  float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
};

// Morton keys: 10 bits per axis of the centre, quantised over the centres'
// bounds: q = min(max((c - lo) * (1024 / (hi - lo)), 0), 1023).
static void morton_keys(uint32_t n, const float *centers, std::vector<uint32_t> &keys) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = gmin(lo[a], centers[3 * i + a]);
      hi[a] = gmax(hi[a], centers[3 * i + a]);
    }
  float scale[3];
  for (int a = 0; a < 3; ++a) {
    float ext = hi[a] - lo[a];
    scale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
  }
  keys.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      float f = (centers[3 * i + a] - lo[a]) * scale[a];
      f = gmin(gmax(f, 0.0f), 1023.0f);
      q[a] = (uint32_t)f;
    }
    keys[i] = morton3(q[0], q[1], q[2]);
  }
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs",
// Fig. 4, on keys sorted by (key, original index); equal keys split on
// the index (delta = 32 + clz(i ^ j)).
static void karras(uint32_t n, const uint32_t *keys, uint32_t *child) {
  auto delta = [&](int64_t i, int64_t j) -> int {
    if (j < 0 || j > (int64_t)n - 1) return -1;
    uint32_t ki = keys[i], kj = keys[j];
    if (ki == kj) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(ki ^ kj);
  };
  for (int64_t i = 0; i + 1 < (int64_t)n; ++i) {
    int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(i, j);
    int64_t s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (delta(i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + std::min(d, 0);
    int64_t lo = std::min(i, j), hi = std::max(i, j);
    child[2 * i + 0] = (lo == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
    child[2 * i + 1] = (hi == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  }
}

static bool build_lbvh(uint32_t n, const float *centers, const float *lo, const float *hi,
                       Bvh &b) {
  b = Bvh();
  b.n = n;
  if (n == 0) return false;
  std::vector<uint32_t> raw;
  morton_keys(n, centers, raw);
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t c) { return raw[a] < raw[c]; });
  b.keys.resize(n);
  b.ids.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    b.keys[i] = raw[order[i]];
    b.ids[i] = order[i];
  }
  b.parent.assign(2 * (size_t)n - 1, kInvalid);
  b.child.assign(n > 1 ? 2 * (size_t)(n - 1) : 0, 0);
  b.box.assign(n > 1 ? 12 * (size_t)(n - 1) : 0, 0.0f);
  if (n == 1) {
    b.root_ref = kLeaf | 0u;
    b.depth = 0;
    for (int a = 0; a < 3; ++a) {
      b.root_lo[a] = lo[3 * b.ids[0] + a];
      b.root_hi[a] = hi[3 * b.ids[0] + a];
    }
    return true;
  }
  karras(n, b.keys.data(), b.child.data());
  for (uint32_t i = 0; i + 1 < n; ++i)
    for (int c = 0; c < 2; ++c) {
      uint32_t r = b.child[2 * i + c];
      if (r & kLeaf)
        b.parent[(n - 1) + (r & ~kLeaf)] = i;
      else
        b.parent[r] = i;
    }
  b.root_ref = 0;
  // Refit: node box = union of its children (min/max are exact).
  std::vector<float> nb(6 * (size_t)(n - 1));
  std::vector<uint32_t> stack;
  std::vector<uint8_t> state(n - 1, 0);
  stack.push_back(0);
  auto ref_box = [&](uint32_t r, float *out) {
    if (r & kLeaf) {
      uint32_t id = b.ids[r & ~kLeaf];
      for (int a = 0; a < 3; ++a) {
        out[a] = lo[3 * id + a];
        out[3 + a] = hi[3 * id + a];
      }
    } else {
      std::memcpy(out, &nb[6 * (size_t)r], 6 * sizeof(float));
    }
  };
  while (!stack.empty()) {
    uint32_t i = stack.back();
    if (state[i] == 0) {
      state[i] = 1;
      for (int c = 0; c < 2; ++c)
        if (!(b.child[2 * i + c] & kLeaf)) stack.push_back(b.child[2 * i + c]);
      continue;
    }
    stack.pop_back();
    float c0[6], c1[6];
    ref_box(b.child[2 * i + 0], c0);
    ref_box(b.child[2 * i + 1], c1);
    std::memcpy(&b.box[12 * (size_t)i], c0, 6 * sizeof(float));
    std::memcpy(&b.box[12 * (size_t)i + 6], c1, 6 * sizeof(float));
    for (int a = 0; a < 3; ++a) {
      nb[6 * (size_t)i + a] = gmin(c0[a], c1[a]);
      nb[6 * (size_t)i + 3 + a] = gmax(c0[3 + a], c1[3 + a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    b.root_lo[a] = nb[a];
    b.root_hi[a] = nb[3 + a];
  }
  // Depth: internal nodes on the longest root-to-leaf path.
  std::vector<uint32_t> nd(n - 1, 0);
  std::vector<uint32_t> order2;
  order2.push_back(0);
  nd[0] = 1;
  uint32_t depth = 1;
  for (size_t k = 0; k < order2.size(); ++k) {
    uint32_t i = order2[k];
    for (int c = 0; c < 2; ++c) {
      uint32_t r = b.child[2 * i + c];
      if (!(r & kLeaf)) {
        nd[r] = nd[i] + 1;
        depth = std::max(depth, nd[r]);
        order2.push_back(r);
      }
    }
  }
  b.depth = depth;
  return true;
}

// -------------------------------------------------------------- scene
struct Scene {
  std::vector<ct_instance> inst;
  std::vector<float> w2o;       // [n][16]
  std::vector<ct_aabb> aabbs;
  std::vector<ct_primitive> prims;
  std::vector<ct_material> mats;
  std::vector<ct_light> cube, point, env;
  ct_light_config cfg{};
  std::vector<Bvh> blas;
  Bvh tlas;
  std::vector<uint32_t> tlas_inst;  // TLAS item -> instance id (instances with primitives)
  uint32_t prim_capacity = 0;
  // The previous frame's scene as the frame after an update sees it:
  // instance transforms and first indices, and the remapping tables
  // (0 = unchanged, kInvalid = deleted).  Empty = same as the current.
  std::vector<float> prev_transform;
  std::vector<uint32_t> prev_first, remap_prim, remap_light;
};

struct TraceCount {
  uint64_t rays = 0, nodes = 0, leaves = 0, instances = 0;
  uint64_t bounce = 0;  // closest-hit path segments (counted apart from `rays`)
};

// Conservative slab test of a BVH child box (the build's own rule, used
// identically by the HIP traversal).
static inline bool slab(const float *bx, V3 o, V3 inv, float limit, float &tn) {
  float t1x = (bx[0] - o.x) * inv.x, t2x = (bx[3] - o.x) * inv.x;
  float t1y = (bx[1] - o.y) * inv.y, t2y = (bx[4] - o.y) * inv.y;
  float t1z = (bx[2] - o.z) * inv.z, t2z = (bx[5] - o.z) * inv.z;
  float n0 = gmax(gmax(gmin(t1x, t2x), gmin(t1y, t2y)), gmin(t1z, t2z));
  float f0 = gmin(gmin(gmax(t1x, t2x), gmax(t1y, t2y)), gmax(t1z, t2z));
  tn = n0;
  return (n0 <= f0) && (f0 >= 0.0f) && (n0 <= limit);
}

struct ClosestState {
  float t;
  uint32_t inst, prim;
};

// traceRayEXT(tmin 0, tmax) with rint.glsl:10-43 as the intersection
// shader: a primitive is a candidate when intersect_box (rayCanStartInBox
// = true, oriented = true) hits with t > 0 and t <= the current tmax.
// The closest candidate wins; equal t goes to the smaller (instance,
// primitive) pair (the reference leaves ties to the driver).
static inline bool better(float t, uint32_t i, uint32_t p, const ClosestState &s) {
  if (t < s.t) return true;
  if (t == s.t) return (i < s.inst) || (i == s.inst && p < s.prim);
  return false;
}

// kBounce: the box test of intersect() (bounce.glsl:86-98, is_hit_from_ray
// with rayCanStartInBox = true, non-oriented) accepting t >= 0.  The
// reference commits the first generated hit; the closest one is taken.
template <bool kBounce = false>
static void closest_leaf(const Scene &sc, uint32_t ii, uint32_t pid, V3 oo, V3 od,
                         ClosestState &st) {
  const ct_instance &in = sc.inst[ii];
  const ct_aabb &bb = sc.aabbs[in.first_primitive_index + pid];
  V3 c = muls(add(load3(bb.minimum), load3(bb.maximum)), 0.5f);
  V3 rad = v3(kHalfVoxel, kHalfVoxel, kHalfVoxel);
  V3 irad = v3(safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel));
  float t;
  V3 nrm;
  V3 inv = v3(safe_inverse(od.x), safe_inverse(od.y), safe_inverse(od.z));
  bool hit = intersect_box(c, rad, irad, oo, od, t, nrm, true, !kBounce, inv);
  if (hit && (kBounce ? t >= 0.0f : t > 0.0f) && better(t, ii, pid, st)) {
    st.t = t;
    st.inst = ii;
    st.prim = pid;
  }
}

// is_vertex_visible, light.glsl:82-142: a primitive occludes when
// is_hit_from_ray (rayCanStartInBox = false, non-oriented) hits with
// 0 <= t <= distance and it is not the target voxel.
static bool any_leaf(const Scene &sc, uint32_t ii, uint32_t pid, V3 oo, V3 od,
                     float dist, uint32_t ti, uint32_t tp) {
  const ct_instance &in = sc.inst[ii];
  const ct_aabb &bb = sc.aabbs[in.first_primitive_index + pid];
  V3 c = muls(add(load3(bb.minimum), load3(bb.maximum)), 0.5f);
  V3 rad = v3(kHalfVoxel, kHalfVoxel, kHalfVoxel);
  V3 irad = v3(safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel));
  V3 inv = v3(safe_inverse(od.x), safe_inverse(od.y), safe_inverse(od.z));
  float t;
  V3 nrm;
  bool hit = intersect_box(c, rad, irad, oo, od, t, nrm, false, false, inv);
  if (!hit || !(t >= 0.0f) || !(t <= dist)) return false;
  if (ii == ti && pid == tp) return false;
  return true;
}

// Two-level traversal: TLAS in world space, each BLAS in object space.
// Closest-hit walks visit the near child first (ties go to child 0) and push
// the far one; any-hit walks visit the far child first (ties go to child 1)
// and push the near one, the build's order for shadow rays (the outcome of
// an any-hit query does not depend on it; the node counts do).
template <bool kAny, bool kBounce = false>
static bool traverse(const Scene &sc, V3 o, V3 d, float limit, uint32_t ti, uint32_t tp,
                     ClosestState &st, TraceCount &cnt) {
  uint32_t stack[kStack];
  int sp = 0;
  V3 winv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  uint32_t ref = sc.tlas.root_ref;
  if (sc.tlas.n == 0) return false;
  for (;;) {
    if (ref & kLeaf) {
      uint32_t ii = sc.tlas_inst[sc.tlas.ids[ref & ~kLeaf]];
      cnt.instances++;
      const float *m = &sc.w2o[16 * (size_t)ii];
      V3 oo = xform_point(m, o), od = xform_dir(m, d);
      V3 oinv = v3(1.0f / od.x, 1.0f / od.y, 1.0f / od.z);
      const Bvh &b = sc.blas[ii];
      int base = sp;
      uint32_t r = b.root_ref;
      for (;;) {
        if (r & kLeaf) {
          cnt.leaves++;
          uint32_t pid = b.ids[r & ~kLeaf];
          if (kAny) {
            if (any_leaf(sc, ii, pid, oo, od, limit, ti, tp)) return true;
          } else {
            closest_leaf<kBounce>(sc, ii, pid, oo, od, st);
          }
        } else {
          cnt.nodes++;
          const float *bx = &b.box[12 * (size_t)r];
          float lim = kAny ? limit : st.t;
          float tn0, tn1;
          bool h0 = slab(bx, oo, oinv, lim, tn0), h1 = slab(bx + 6, oo, oinv, lim, tn1);
          uint32_t c0 = b.child[2 * r], c1 = b.child[2 * r + 1];
          if (h0 && h1) {
            const bool swap = kAny ? !(tn1 < tn0) : (tn1 < tn0);
            stack[sp++] = swap ? c0 : c1;
            if (sp >= kStack) { std::fprintf(stderr, "oracle stack overflow\n"); std::abort(); }
            r = swap ? c1 : c0;
            continue;
          }
          if (h0) { r = c0; continue; }
          if (h1) { r = c1; continue; }
        }
        if (sp == base) break;
        r = stack[--sp];
      }
    } else {
      cnt.nodes++;
      const float *bx = &sc.tlas.box[12 * (size_t)ref];
      float lim = kAny ? limit : st.t;
      float tn0, tn1;
      bool h0 = slab(bx, o, winv, lim, tn0), h1 = slab(bx + 6, o, winv, lim, tn1);
      uint32_t c0 = sc.tlas.child[2 * ref], c1 = sc.tlas.child[2 * ref + 1];
      if (h0 && h1) {
        const bool swap = kAny ? !(tn1 < tn0) : (tn1 < tn0);
        stack[sp++] = swap ? c0 : c1;
        ref = swap ? c1 : c0;
        continue;
      }
      if (h0) { ref = c0; continue; }
      if (h1) { ref = c1; continue; }
    }
    if (sp == 0) break;
    ref = stack[--sp];
  }
  return false;
}

static bool trace_any(const Scene &sc, V3 o, V3 d, float dist, uint32_t ti, uint32_t tp,
                      TraceCount &cnt) {
  ClosestState st{0, 0, 0};
  cnt.rays++;
  return traverse<true>(sc, o, d, dist, ti, tp, st, cnt);
}

static ClosestState trace_closest(const Scene &sc, V3 o, V3 d, float tmax, TraceCount &cnt) {
  ClosestState st{tmax, kInvalid, kInvalid};
  cnt.rays++;
  traverse<false>(sc, o, d, tmax, kInvalid, kInvalid, st, cnt);
  return st;
}

static ClosestState trace_bounce(const Scene &sc, V3 o, V3 d, TraceCount &cnt) {
  ClosestState st{kMaxDistance, kInvalid, kInvalid};
  cnt.bounce++;
  traverse<false, true>(sc, o, d, kMaxDistance, kInvalid, kInvalid, st, cnt);
  return st;
}

static void build_scene_bvh(Scene &sc) {
  const uint32_t ni = (uint32_t)sc.inst.size();
  sc.blas.assign(ni, Bvh{});
  std::vector<float> tl_c(3 * (size_t)ni), tl_lo(3 * (size_t)ni), tl_hi(3 * (size_t)ni);
  const float r = kHalfVoxel + kLeafPad;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ii = 0; ii < (int64_t)ni; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const ct_instance &in = sc.inst[i];
    uint32_t n = in.primitive_count;
    if (n == 0) continue;  // an unused instance id
    std::vector<float> c(3 * (size_t)n), lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) {
      const ct_aabb &bb = sc.aabbs[in.first_primitive_index + k];
      V3 cc = muls(add(load3(bb.minimum), load3(bb.maximum)), 0.5f);
      store3(&c[3 * k], cc);
      store3(&lo[3 * k], v3(cc.x - r, cc.y - r, cc.z - r));
      store3(&hi[3 * k], v3(cc.x + r, cc.y + r, cc.z + r));
    }
    build_lbvh(n, c.data(), lo.data(), hi.data(), sc.blas[i]);
    // Instance world box: the 8 transformed corners of the BLAS root box,
    // padded once more.
    const Bvh &b = sc.blas[i];
    float wl[3] = {INFINITY, INFINITY, INFINITY}, wh[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
      V3 p = v3((k & 1) ? b.root_hi[0] : b.root_lo[0], (k & 2) ? b.root_hi[1] : b.root_lo[1],
                (k & 4) ? b.root_hi[2] : b.root_lo[2]);
      V3 w = xform_point(in.transform, p);
      wl[0] = gmin(wl[0], w.x); wl[1] = gmin(wl[1], w.y); wl[2] = gmin(wl[2], w.z);
      wh[0] = gmax(wh[0], w.x); wh[1] = gmax(wh[1], w.y); wh[2] = gmax(wh[2], w.z);
    }
    for (int a = 0; a < 3; ++a) {
      tl_lo[3 * i + a] = wl[a] - kLeafPad;
      tl_hi[3 * i + a] = wh[a] + kLeafPad;
      tl_c[3 * i + a] = (tl_lo[3 * i + a] + tl_hi[3 * i + a]) * 0.5f;
    }
  }
  // TLAS over the instances that hold primitives, in id order.
  sc.tlas_inst.clear();
  std::vector<float> uc, ul, uh;
  for (uint32_t i = 0; i < ni; ++i) {
    if (sc.inst[i].primitive_count == 0) continue;
    sc.tlas_inst.push_back(i);
    for (int a = 0; a < 3; ++a) {
      uc.push_back(tl_c[3 * i + a]);
      ul.push_back(tl_lo[3 * i + a]);
      uh.push_back(tl_hi[3 * i + a]);
    }
  }
  build_lbvh((uint32_t)sc.tlas_inst.size(), uc.data(), ul.data(), uh.data(), sc.tlas);
}

// ------------------------------------------------------- DI frame state
struct Hit { // HIT_INFO_INPUT, shared.inl:171-179
  V3 world_hit, world_nrm;
  float distance;
  V3 scatter_dir;
  uint32_t inst, prim;
  uint32_t mat_idx;
};

struct Res { // RESERVOIR, shared.inl:368-377
  uint32_t Y = kInvalid, l_type = CT_LIGHT_NONE, seed = 0;
  float W_y = 0, W_sum = 0, M = 0;
  V3 F = {0, 0, 0};
};

static inline Res res_from(const ct_reservoir &r) {
  Res o;
  o.Y = r.Y; o.l_type = r.l_type; o.seed = r.seed;
  o.W_y = r.W_y; o.W_sum = r.W_sum; o.M = r.M;
  o.F = load3(r.F);
  return o;
}
static inline ct_reservoir res_to(const Res &r) {
  ct_reservoir o;
  o.Y = r.Y; o.l_type = r.l_type; o.seed = r.seed;
  o.W_y = r.W_y; o.W_sum = r.W_sum; o.M = r.M;
  store3(o.F, r.F);
  return o;
}

struct Frame {
  Scene *sc = nullptr;
  // The scene before the last update (its TLAS, BLASes and instance
  // records): traced by temporal reuse's target-at-previous query while
  // remapping is active, get_previous_tlas (restir_resampling.glsl:369 ->
  // light.glsl:102, accel_struct_mngr.hpp:219-223).  Null: the current one.
  const Scene *prev_sc = nullptr;
  uint32_t W = 0, H = 0;
  std::vector<ct_di_info> di_cur, di_prev;
  std::vector<ct_reservoir> res_int, res_prev;
  std::vector<float> velocity;   // [W*H][2]
  std::vector<float> img[2];     // [W*H][4]
  std::vector<float> taa;        // [W*H][4]
  std::vector<ct_path_reservoir> pres_o, pres_t;  // output / temporal path reservoirs
  std::vector<float> indirect;   // [W*H][3] indirect colour of the first pass
  int cur = 0;
  bool first = true;
  ct_frame_stats stats{};
};

struct Ctx { // per-frame constants read by every pixel
  const Scene *sc;
  const Scene *vis_sc = nullptr;  // scene the shadow rays traverse (null: sc)
  Frame *fr;
  const ct_frame_params *p;
  std::vector<float> pt; // per instance: prev_view_proj * geometry_T
};

static inline const ct_material &material(const Ctx &c, uint32_t idx) {
  return c.sc->mats[idx];
}

// reservoir.glsl:17-25
static inline void init_res(Res &r) { r = Res(); }
// reservoir.glsl:31-50
static bool update_reservoir(Res &r, uint32_t X, uint32_t l_type, uint32_t random_seed, V3 F,
                             float w, float c, uint32_t &seed) {
  if (finite_or_nan_bad(w)) return false;
  r.M += c;
  r.W_sum += w;
  if (rnd(seed) < (w / r.W_sum)) {
    r.Y = X;
    r.l_type = l_type;
    r.seed = random_seed;
    r.F = F;
    return true;
  }
  return false;
}
// reservoir.glsl:72-74
static inline bool res_valid(const Res &r) {
  return r.M > 0.0f && r.Y != kInvalid && r.l_type != CT_LIGHT_NONE;
}

// prng.glsl:198-215
static V3 random_cube_normal(uint32_t &seed) {
  uint32_t face = std::min(urnd_interval(seed, 0, 6), 5u);
  switch (face) {
    case 0: return v3(-1, 0, 0);
    case 1: return v3(1, 0, 0);
    case 2: return v3(0, -1, 0);
    case 3: return v3(0, 1, 0);
    case 4: return v3(0, 0, -1);
    default: return v3(0, 0, 1);
  }
}
// prng.glsl:217-224
static void orthonormal_basis(V3 n, V3 &t1, V3 &t2) {
  if (std::fabs(n.x) > std::fabs(n.y))
    t1 = normalize(v3(-n.z, 0.0f, n.x));
  else
    t1 = normalize(v3(0.0f, n.z, -n.y));
  t2 = normalize(cross(n, t1));
}
// prng.glsl:235-252
static V3 random_quad(V3 n, V3 p, float sx, float sy, uint32_t &seed) {
  float u = rnd(seed);
  float v = rnd(seed);
  float qx = u * sx, qy = v * sy;
  V3 xa, ya;
  orthonormal_basis(n, xa, ya);
  return add(add(p, muls(xa, qx)), muls(ya, qy));
}
// prng.glsl:143-161 (the rejection loop is capped at 1024 draws)
static V3 random_on_hemisphere(uint32_t &seed, V3 n) {
  V3 p = v3(0, 0, 0);
  for (int k = 0; k < 1024; ++k) {
    float x = rnd_interval(seed, -1.0f, 1.0f);
    float y = rnd_interval(seed, -1.0f, 1.0f);
    float z = rnd_interval(seed, -1.0f, 1.0f);
    p = v3(x, y, z);
    if (x * x + y * y + z * z >= 1.0f) continue;
    break;
  }
  V3 u = normalize(p);
  return dot(u, n) > 0.0f ? u : neg(u);
}

// mat.glsl:58-64 (only the z component of to_local is used)
static inline float local_z(V3 n, V3 v) { return dot(v, normalize(n)); }
// light.glsl:200-229 (Lambertian, COSINE_HEMISPHERE_SAMPLING = 1)
static float sample_material_pdf(const ct_material &, V3 n, V3 wo, V3 wi) {
  float wo_z = local_z(n, wo), wi_z = local_z(n, wi);
  if (gmin(wo_z, wi_z) < kMinCosTheta) return 0.0f;
  return kInvPi * wi_z;
}
// mat.glsl:24-27, 66-88 (Lambertian)
static V3 evaluate_material(const ct_material &m, V3 n, V3 wo, V3 wi) {
  float wo_z = local_z(n, wo), wi_z = local_z(n, wi);
  if (gmin(wo_z, wi_z) < kMinCosTheta) return v3(0, 0, 0);
  return muls(muls(load3(m.diffuse), kInvPi), wi_z);
}
// light.glsl:160-164
static float geom_fact_sa(V3 P, V3 Ps, V3 ns) {
  V3 dir = normalize(sub(Ps, P));
  V3 dd = sub(Ps, P);
  float dist2 = dot(dd, dd);
  return std::fabs(dot(ns, dir)) / dist2;
}

struct LightSample {
  bool vis;
  V3 l_pos, l_nor, Le;
  float G;
};

static const ct_light *light_by_type(const Ctx &c, uint32_t type, uint32_t idx) {
  const Scene &s = *c.sc;
  switch (type) {
    case CT_LIGHT_POINT: return idx < s.point.size() ? &s.point[idx] : nullptr;
    case CT_LIGHT_ENV_MAP: return idx < s.env.size() ? &s.env[idx] : nullptr;
    case CT_LIGHT_CUBE: return idx < s.cube.size() ? &s.cube[idx] : nullptr;
    default: return nullptr;
  }
}

// sample_lights, light.glsl:295-399.  `visibility` traces the shadow ray.
// An output the shader leaves unwritten (G on the fast-discard path) is 0.
static bool sample_lights(const Ctx &c, const Hit &hit, const ct_light &l, float &pdf,
                          LightSample &out, uint32_t &seed, bool calc_pdf, bool visibility,
                          TraceCount &cnt) {
  V3 P = hit.world_hit, n = hit.world_nrm;
  float distance = -1.0f;
  V3 l_wi = v3(0, 0, 0), l_pos = v3(0, 0, 0), l_nor = v3(0, 0, 0);
  out.G = 0.0f;
  out.l_pos = l_pos;
  out.l_nor = l_nor;
  if (l.instance_info.instance_id == hit.inst && l.instance_info.primitive_id == hit.prim) {
    out.Le = v3(0, 0, 0);
    out.vis = false;
    return false;
  }
  bool vis = true, check_instance = false;
  V3 emissive = load3(l.emissive);
  if (l.type == CT_LIGHT_CUBE) {
    float half = l.size * 0.5f;
    float qs = l.size;
    uint32_t i = 0;
    do {
      vis = true;
      l_nor = random_cube_normal(seed);
      l_pos = add(load3(l.position), muls(l_nor, half));
      l_pos = random_quad(l_nor, l_pos, qs * 0.5f, qs * 0.5f, seed);
      l_wi = normalize(sub(P, l_pos));
      vis = vis && (dot(l_wi, l_nor) > 0.0f);
      vis = vis && (dot(l_wi, n) < 0.0f);
    } while (!vis && i++ < 10);
    l_pos = compute_ray_origin(l_pos, l_nor);
    l_pos = compute_ray_origin(l_pos, l_nor);
    distance = length(sub(P, l_pos));
    if (calc_pdf) {
      float area = qs * qs * 6.0f;
      pdf /= area;
    }
    check_instance = true;
    out.G = geom_fact_sa(P, l_pos, l_nor);
  } else if (l.type == CT_LIGHT_POINT) {
    l_pos = load3(l.position);
    l_nor = normalize(sub(P, l_pos));
    distance = length(sub(P, l_pos));
    check_instance = false;
    l_wi = l_nor;
    out.G = geom_fact_sa(P, l_pos, l_nor);
  } else if (l.type == CT_LIGHT_ENV_MAP) {
    V3 l_dir = random_on_hemisphere(seed, n);
    l_pos = add(P, muls(l_dir, kMaxDistance));
    l_nor = normalize(sub(P, l_pos));
    distance = length(sub(P, l_pos));
    check_instance = false;
    l_wi = l_nor;
    emissive = mulv(emissive, sky_color(c.p->time, c.p->is_afternoon != 0, l_dir));
    if (calc_pdf) pdf /= kInv4Pi;
    out.G = 1.0f;
  }
  V3 l_v = neg(l_wi);
  if (visibility && vis) {
    uint32_t ti = check_instance ? l.instance_info.instance_id : kInvalid;
    uint32_t tp = check_instance ? l.instance_info.primitive_id : kInvalid;
    vis = vis && !trace_any(c.vis_sc ? *c.vis_sc : *c.sc, P, l_v, distance, ti, tp, cnt);
  }
  out.l_pos = l_pos;
  out.l_nor = l_nor;
  out.Le = vis ? emissive : v3(0, 0, 0);
  out.vis = vis;
  return vis;
}

// calculate_sampled_light, light.glsl:401-435
static V3 calculate_sampled_light(const Ctx &c, V3 ray_d, Hit &hit, const ct_material &mat,
                                  const ct_light &light, float pdf, float &pdf_out, float &G,
                                  uint32_t &seed, bool calc_pdf, bool use_pdf, bool use_vis,
                                  TraceCount &cnt) {
  V3 surface_normal = normalize(hit.world_nrm);
  V3 wo = neg(normalize(ray_d));
  pdf_out = pdf;
  V3 result = v3(0, 0, 0);
  G = 0.0f;
  if (light.emissive[0] <= 0.0f && light.emissive[1] <= 0.0f && light.emissive[2] <= 0.0f)
    return result;
  LightSample ls;
  bool ok = sample_lights(c, hit, light, pdf_out, ls, seed, calc_pdf, use_vis, cnt);
  G = ls.G;
  if (ok) {
    V3 wi = normalize(sub(ls.l_pos, hit.world_hit));
    V3 brdf = evaluate_material(mat, surface_normal, wo, wi);
    hit.scatter_dir = wi;
    V3 bl = mulv(brdf, ls.Le);
    result = use_pdf ? divs(muls(bl, G), pdf_out) : muls(bl, G);
  }
  return result;
}

// reservoir_get_radiance, reservoir.glsl:98-112
static V3 reservoir_get_radiance(const Ctx &c, const Res &r, V3 ray_d, Hit hit,
                                 const ct_material &mat, TraceCount &cnt) {
  if (!res_valid(r)) return v3(0, 0, 0);
  const ct_light *l = light_by_type(c, r.l_type, r.Y);
  if (!l) return v3(0, 0, 0);
  float pdf_out, G;
  uint32_t s = r.seed;
  return calculate_sampled_light(c, ray_d, hit, mat, *l, 1.0f, pdf_out, G, s, false, false,
                                 true, cnt);
}

// reservoir_visibility_pass, reservoir.glsl:76-96, 114-126
static void reservoir_visibility_pass(const Ctx &c, Res &r, const Hit &hit, TraceCount &cnt) {
  bool visibility = false;
  if (res_valid(r)) {
    const ct_light *l = light_by_type(c, r.l_type, r.Y);
    if (l) {
      float l_pdf = 1.0f;
      uint32_t s = r.seed;
      LightSample ls;
      visibility = sample_lights(c, hit, *l, l_pdf, ls, s, false, true, cnt);
    }
  }
  r.W_y = visibility ? (r.W_sum / luminance(r.F)) : 0.0f;
  r.F = muls(r.F, visibility ? 1.0f : 0.0f);
}

static bool ris_brdf(const Ctx &c, Res &reservoir, V3 ray_o, Hit &hit, const ct_material &mat, uint32_t nee,
                     uint32_t n_brdf, uint32_t cube_light_count, uint32_t &seed, TraceCount &bcnt);

// RIS, restir_resampling.glsl:19-255: the point, environment and cube
// strategies, the BRDF strategy with CT_RIS_BRDF_BIT (its closest-hit rays
// count into bcnt.bounce), then the visibility pass unless the BRDF
// strategy's candidate was the last one selected (:240-252).  ray_o, ray_d:
// the camera ray.
static Res ris(const Ctx &c, float confidence, V3 ray_o, V3 ray_d, Hit &hit, const ct_material &mat,
               uint32_t &seed, TraceCount &cnt, TraceCount &bcnt) {
  const ct_restir_params &rp = c.p->restir;
  const ct_light_config &lc = c.sc->cfg;
  uint32_t active = c.p->is_active;
  confidence = clampf(confidence, 0.0f, 1.0f);
  bool point_active = (active & CT_RIS_POINT_LIGHT_BIT) != 0;
  bool env_active = (active & CT_RIS_ENV_LIGHT_BIT) != 0;
  bool cube_active = (active & CT_RIS_CUBE_LIGHT_BIT) != 0;
  uint32_t point_light_count = lc.point_light_count;
  uint32_t n_point = std::max(
      f2u_sat(gmin((float)rp.max_ris_point_samples * (1.0f - confidence), (float)point_light_count)),
      rp.min_ris_point_samples);
  uint32_t env_light_count = lc.env_map_count;
  uint32_t n_env = std::max(
      f2u_sat(gmin((float)rp.max_ris_env_samples * (1.0f - confidence), (float)env_light_count)),
      rp.min_ris_env_samples);
  uint32_t cube_light_count = lc.cube_light_count < 1 ? 1 : lc.cube_light_count;
  uint32_t n_cube = std::max(
      f2u_sat(gmin((float)rp.max_ris_cube_samples * (1.0f - confidence), (float)cube_light_count)),
      rp.min_ris_cube_samples);
  uint32_t brdf_count = lc.brdf_count;
  uint32_t n_brdf = brdf_count + n_env;
  uint32_t nee = n_point + n_cube;
  Res reservoir;
  init_res(reservoir);
  auto strategy = [&](Res &target, uint32_t count, uint32_t nsamp, uint32_t type,
                      const std::vector<ct_light> &lights) {
    float pdf = 1.0f / (float)count;
    for (uint32_t l = 0; l < nsamp; ++l) {
      uint32_t idx = std::min(urnd_interval(seed, 0, count), count - 1);
      const ct_light &light = lights[idx];
      float light_pdf = 1.0f, G = 0.0f;
      uint32_t current_seed = seed;
      V3 F = calculate_sampled_light(c, ray_d, hit, mat, light, pdf, light_pdf, G, seed, true,
                                     false, false, cnt);
      float brdf_pdf = sample_material_pdf(mat, hit.world_nrm, neg(ray_d), hit.scatter_dir) * G;
      float m_i = 1.0f / gmax((float)nee * light_pdf + (float)n_brdf * brdf_pdf, 1e-6f);
      float w_i = luminance(F) * m_i;
      update_reservoir(target, idx, type, current_seed, F, w_i, 1.0f, seed);
    }
  };
  if (point_active && !c.sc->point.empty()) strategy(reservoir, point_light_count, n_point, CT_LIGHT_POINT, c.sc->point);
  if (env_active && !c.sc->env.empty()) {
    Res er;
    init_res(er);
    strategy(er, env_light_count, n_env, CT_LIGHT_ENV_MAP, c.sc->env);
    update_reservoir(reservoir, er.Y, er.l_type, er.seed, er.F, er.W_sum, er.M, seed);
  }
  if (cube_active && !c.sc->cube.empty()) {
    Res cr;
    init_res(cr);
    strategy(cr, cube_light_count, n_cube, CT_LIGHT_CUBE, c.sc->cube);
    update_reservoir(reservoir, cr.Y, cr.l_type, cr.seed, cr.F, cr.W_sum, cr.M, seed);
  }
  bool requires_visibility = true;
  if ((active & CT_RIS_BRDF_BIT) &&
      ris_brdf(c, reservoir, ray_o, hit, mat, nee, n_brdf, cube_light_count, seed, bcnt))
    requires_visibility = false;
  if (requires_visibility) reservoir_visibility_pass(c, reservoir, hit, cnt);
  return reservoir;
}

static inline Hit hit_from_di(const ct_di_info &d) {
  Hit h;
  V3 n = load3(d.normal);
  V3 p = compute_ray_origin(load3(d.position), n);
  p = compute_ray_origin(p, n);
  h.world_hit = p;
  h.world_nrm = n;
  h.distance = d.distance;
  h.scatter_dir = v3(0, 0, 0);
  h.inst = d.instance_hit.instance_id;
  h.prim = d.instance_hit.primitive_id;
  h.mat_idx = d.mat_index;
  return h;
}

// The remapping tables' entry for an index, 0 (unchanged) outside them.
static uint32_t remap_entry(const std::vector<uint32_t> &t, uint64_t i) { return i < t.size() ? t[i] : 0u; }

// TEMPORAL_REUSE, restir_resampling.glsl:311-416.  With REMAP_BIT the
// previous hit's primitive (:338-345, looked up by the previous instance
// table's first index, primitives.glsl:133-137) and the previous cube
// light (:380-387) go through the remapping tables; the light is evaluated
// at its current index.
static void temporal_reuse(const Ctx &c, Res &r, Res prev, uint32_t pidx, V3 ray_d, Hit &hit,
                           const ct_material &mat, uint32_t &seed, TraceCount &cnt) {
  const float mcap = c.p->restir.max_influence_from_the_past;
  const bool remap = (c.p->is_active & CT_REMAP_BIT) != 0;
  const Scene &sc = *c.sc;
  prev.M = gmin(prev.M, mcap * r.M);
  const float prev_M = prev.M;
  const float new_M = r.M + prev_M;
  if (r.W_sum > 0.0f) {
    float target_lum_at_prev = 0.0f;
    if (luminance(r.F) > 1e-6f) {
      ct_di_info dp = c.fr->di_prev[pidx];
      if (remap) {
        const uint32_t pi = dp.instance_hit.instance_id;
        uint64_t first = 0;
        if (pi < sc.inst.size()) first = sc.prev_first.empty() ? sc.inst[pi].first_primitive_index : sc.prev_first[pi];
        const uint32_t m = remap_entry(sc.remap_prim, first + dp.instance_hit.primitive_id);
        if (m == kInvalid) return;
        if (m != 0u) dp.instance_hit.primitive_id = m;
      }
      Hit ph = hit_from_di(dp);
      const ct_material &pm = material(c, dp.mat_index);
      V3 pd = normalize(sub(ph.world_hit, load3(dp.ray_origin)));
      Ctx cp = c;  // the previous frame's scene while remapping (is_remapping_active)
      if (remap && c.fr->prev_sc) cp.vis_sc = c.fr->prev_sc;
      target_lum_at_prev = luminance(reservoir_get_radiance(cp, r, pd, ph, pm, cnt));
    }
    const float p_curr = r.M * luminance(r.F);
    const float m_curr = p_curr / gmax(p_curr + prev_M * target_lum_at_prev, 1e-6f);
    r.W_sum *= m_curr;
  }
  if (remap && prev.l_type == CT_LIGHT_CUBE) {
    const uint32_t m = remap_entry(sc.remap_light, prev.Y);
    if (m != 0u) prev.Y = m;
  }
  uint32_t prev_light_index = prev.Y;
  if (prev_light_index != kInvalid) {
    V3 current_target = reservoir_get_radiance(c, prev, ray_d, hit, mat, cnt);
    const float lc = luminance(current_target);
    if (lc > 0.0f) {
      const float lp = luminance(prev.F);
      const float p_prev = prev.M * lp;
      const float m_prev = p_prev / gmax(p_prev + r.M * lc, 1e-6f);
      const float w_prev = m_prev * lc * prev.W_y;
      update_reservoir(r, prev_light_index, prev.l_type, prev.seed, current_target, w_prev, 1.0f,
                       seed);
    }
  }
  float tl = luminance(r.F);
  r.W_y = tl > 0.0f ? r.W_sum / tl : 0.0f;
  r.M = new_M;
}

static void pairwise_combine(const Ctx &c, Res &r, uint32_t num, const Res *nres, const ct_di_info *const *ndi,
                             V3 ray_d, Hit &hit, const ct_material &mat, uint32_t &seed, TraceCount &cnt);

// pairwise_mis.glsl:8-114 and SPATIAL_REUSE, restir_resampling.glsl:418-540
static void spatial_reuse(const Ctx &c, Res &r, float confidence, uint32_t x, uint32_t y,
                          V3 ray_d, Hit &hit, const ct_material &mat, uint32_t &seed,
                          TraceCount &cnt) {
  const ct_restir_params &rp = c.p->restir;
  const uint32_t W = c.p->width, H = c.p->height;
  if (r.W_y == 0.0f) confidence = 0.0f;
  float radius = mixf(rp.max_neighbor_radius, rp.min_neighbor_radius,
                      clampf(hit.distance / rp.max_distance_to_hit, 0.0f, 1.0f));
  uint32_t k = f2u_sat(mixf((float)rp.max_neighbors, (float)rp.min_neighbors,
                            clampf(confidence, 0.0f, 1.0f)));
  uint32_t idx[CT_MAX_NEIGHBORS];
  uint32_t num = 0;
  for (uint32_t i = 0; i < k; ++i) {
    float ox = 2.0f * rnd(seed) - 1.0f;
    float oy = 2.0f * rnd(seed) - 1.0f;
    int64_t nx = (int64_t)x + f2i_sat(ox * radius);
    int64_t ny = (int64_t)y + f2i_sat(oy * radius);
    if (nx < 0 || nx >= (int64_t)W || ny < 0 || ny >= (int64_t)H) continue;
    if (nx == (int64_t)x && ny == (int64_t)y) continue;
    uint32_t l = (uint32_t)ny * W + (uint32_t)nx;
    if (c.fr->di_cur[l].distance <= 0.0f) continue;
    if (num < CT_MAX_NEIGHBORS) idx[num++] = l;
  }
  if (num == 0) return;
  Res nres[CT_MAX_NEIGHBORS];
  const ct_di_info *ndi[CT_MAX_NEIGHBORS];
  for (uint32_t i = 0; i < num; ++i) {
    nres[i] = res_from(c.fr->res_int[idx[i]]);
    ndi[i] = &c.fr->di_cur[idx[i]];
  }
  pairwise_combine(c, r, num, nres, ndi, ray_d, hit, mat, seed, cnt);
}

// The pairwise MIS combination of spatial reuse over `num` chosen
// neighbours (pairwise_mis.glsl:8-114): pairwise_init, one pairwise_stream
// per neighbour, pairwise_end.
static void pairwise_combine(const Ctx &c, Res &r, uint32_t num, const Res *nres, const ct_di_info *const *ndi,
                             V3 ray_d, Hit &hit, const ct_material &mat, uint32_t &seed, TraceCount &cnt) {
  // pairwise_init
  Res mres;
  init_res(mres);
  float m_c = 1.0f;
  float M_s = r.M;
  const float fk = (float)num;
  for (uint32_t i = 0; i < num; ++i) {
    Res in = nres[i];
    const ct_di_info &nd = *ndi[i];
    Hit nh = hit_from_di(nd);
    V3 nray_d = normalize(sub(nh.world_hit, load3(nd.ray_origin)));
    const ct_material &nm = material(c, nd.mat_index);
    // pairwise_stream
    V3 curr_target = v3(0, 0, 0);
    float m_i = 0.0f;
    if (res_valid(in)) {
      curr_target = reservoir_get_radiance(c, in, ray_d, hit, mat, cnt);
      const float tl = luminance(curr_target);
      const float p_i_y_i = luminance(in.F);
      float mm = in.M * p_i_y_i;
      float den = mm + (r.M / fk) * tl;
      m_i = den > 0.0f ? mm / den : 0.0f;
    }
    V3 input_target = v3(0, 0, 0);
    if (res_valid(r)) input_target = reservoir_get_radiance(c, r, nray_d, nh, nm, cnt);
    {
      const float p_i_y_c = luminance(input_target);
      const float p_c_y_c = luminance(r.F);
      const float numer = in.M * p_i_y_c;
      const bool den_ok = (p_c_y_c + numer) > 0.0f;
      m_c += 1.0f;
      m_c -= den_ok ? numer / (numer + (r.M / fk) * p_c_y_c) : 0.0f;
    }
    if (res_valid(in)) {
      const float w_i = luminance(curr_target) * in.W_y * m_i;
      update_reservoir(mres, in.Y, in.l_type, in.seed, curr_target, w_i, 1.0f, seed);
    }
    M_s += in.M;
  }
  // pairwise_end (its seed is a copy)
  uint32_t s2 = seed;
  const float w_c = luminance(r.F) * r.W_y * m_c;
  update_reservoir(mres, r.Y, r.l_type, r.seed, r.F, w_c, 1.0f, s2);
  mres.M = M_s;
  const float tl = luminance(mres.F);
  mres.W_y = tl > 0.0f ? mres.W_sum / (tl * (float)(1u + num)) : 0.0f;
  r = mres;
}

// ================================================================ ReSTIR PT
// Indirect illumination (INDIRECT_ILLUMINATION_ON with RESTIR_PT_ON):
// trace_restir_path_tracing (path_tracer.glsl:854-887), intersect()
// (bounce.glsl:9-130), the reconnection shift (shift.glsl:25-350) and the
// temporal / spatial path reuse (indirect_illumination.glsl:147-452), for
// Lambertian materials.

constexpr float kHlfMax = 6.5504e+4f;        // shared.inl:56
constexpr float kHlfMin = 1.175494351e-38f;  // shared.inl:57
constexpr float kPiF = 3.1415926535897932384626433832795f;  // DAXA_PI, shared.inl:49

// sin/cos of x in [0, 2*pi).  GLSL leaves their accuracy to the
// implementation; this build fixes one: Cody-Waite reduction by pi/2 and
// the Cephes single-precision minimax polynomials (S. L. Moshier,
// sinf.c / cosf.c) on [-pi/4, pi/4].
static void sincos_ct(float x, float &s, float &c) {
  const float j = std::floor(x * 0.636619772367581343f + 0.5f);
  const int q = (int)j;
  const float r = ((x - j * 1.5703125f) - j * 4.837512969970703125e-4f) - j * 7.54978995489188216e-8f;
  const float z = r * r;
  const float sp = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
  const float cp = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
                   0.5f * z + 1.0f;
  switch (q & 3) {
    case 0: s = sp; c = cp; break;
    case 1: s = cp; c = -sp; break;
    case 2: s = -sp; c = -cp; break;
    default: s = -cp; c = sp; break;
  }
}

struct Isect {  // INTERSECT, shared.inl:323-334
  bool is_hit = false;
  float distance = 0.0f;
  V3 P{0, 0, 0}, n{0, 0, 0}, wo{0, 0, 0}, wi{0, 0, 0};
  uint32_t inst = CT_MAX_INSTANCES, prim = CT_MAX_PRIMITIVES, mat_idx = 10000u;  // MAX_MATERIALS
  ct_material mat{};
};

static inline uint32_t pf_path_length(uint32_t f) { return f & 0xFu; }
static inline uint32_t pf_rc_length(uint32_t f) { return (f >> 4) & 0xFu; }
static inline void pf_set_path_length(uint32_t &f, uint32_t l) { f &= ~0xFu; f |= (l & 0xFu); }
static inline void pf_set_rc_length(uint32_t &f, uint32_t l) { f &= ~0xF0u; f |= (l & 0xFu) << 4; }
static inline bool pf_delta(uint32_t f, bool before) { return (f >> (before ? 20 : 21)) & 1u; }
static inline bool pf_transmission(uint32_t f, bool before) { return (f >> (before ? 22 : 23)) & 1u; }
static inline void pf_set_last_nee(uint32_t &f, bool b) { f &= ~0x10000u; f |= (uint32_t)(b ? 1 : 0) << 16; }
static inline bool pf_last_nee(uint32_t f) { return (f >> 16) & 1u; }
static inline void pf_set_light_type(uint32_t &f, uint32_t t) { f &= ~0xc0000u; f |= (t & 3u) << 18; }
static inline uint32_t pf_light_type(uint32_t f) { return (f >> 18) & 3u; }

// path_F_to_scalar (Rec. 601 weights), path_reservoir.glsl:160-162
static inline float lum601(V3 c) { return dot(c, v3(0.299f, 0.587f, 0.114f)); }
// eval_mis / power_heuristic, light.glsl:171-191
static inline float eval_mis(float nf, float f_pdf, float ng, float g_pdf) {
  float a = nf * f_pdf, b = ng * g_pdf;
  float f = a * a, g = b * b;
  return f / (f + g);
}
static inline bool bad3(V3 v) {
  return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z) || std::isinf(v.x) || std::isinf(v.y) ||
         std::isinf(v.z);
}
static inline bool jac_invalid(float j) { return j <= 0.0f || std::isnan(j) || std::isinf(j); }

static void pres_init(ct_path_reservoir &r, uint32_t max_depth) {  // path_reservoir.glsl:151-165
  r = ct_path_reservoir{};
  r.M = 0.0f;
  r.weight = 0.0f;
  r.path_flags = 0;
  pf_set_rc_length(r.path_flags, max_depth);
  r.rc_vertex_hit.object.instance_id = CT_MAX_INSTANCES;
  r.rc_vertex_hit.object.primitive_id = CT_MAX_PRIMITIVES;
}
static bool pres_update(ct_path_reservoir &r, V3 in_F, float p, uint32_t &seed) {  // :183-202
  r.M += 1.0f;
  float w = lum601(in_F) / p;
  if (std::isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (rnd(seed) * r.weight <= w) {
    store3(r.F, in_F);
    return true;
  }
  return false;
}
static void pres_copy_from(ct_path_reservoir &r, const ct_path_reservoir &in, V3 in_F) {
  r.path_flags = in.path_flags;
  r.rc_random_seed = in.rc_random_seed;
  r.init_random_seed = in.init_random_seed;
  std::memcpy(r.cached_jacobian, in.cached_jacobian, sizeof(r.cached_jacobian));
  r.light_pdf = in.light_pdf;
  std::memcpy(r.rc_vertex_wi, in.rc_vertex_wi, sizeof(r.rc_vertex_wi));
  r.rc_vertex_hit = in.rc_vertex_hit;
  std::memcpy(r.rc_vertex_irradiance, in.rc_vertex_irradiance, sizeof(r.rc_vertex_irradiance));
  store3(r.F, in_F);
}
// path_reservoir_merge, path_reservoir.glsl:204-232
static bool pres_merge(ct_path_reservoir &r, V3 in_F, float jac, const ct_path_reservoir &in, uint32_t &seed,
                       float mis, bool force) {
  float w = lum601(in_F) * jac * in.M * in.weight * mis;
  r.M += in.M;
  if (std::isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (force || rnd(seed) * r.weight <= w) {
    pres_copy_from(r, in, in_F);
    return true;
  }
  return false;
}
// path_reservoir_merge_with_resampling_MIS, path_reservoir.glsl:234-263
static bool pres_merge_mis(ct_path_reservoir &r, V3 in_F, float jac, const ct_path_reservoir &in, uint32_t &seed,
                           float mis) {
  float w = lum601(in_F) * jac * in.weight * mis;
  r.M += in.M;
  if (std::isnan(w) || w == 0.0f) return false;
  r.weight += w;
  if (rnd(seed) * r.weight <= w) {
    pres_copy_from(r, in, in_F);
    return true;
  }
  return false;
}

// transpose(world2obj) * vec4(n, 0)
static V3 normal_to_world(const float *m, V3 n) {
  V3 o;
  float *po = &o.x;
  for (int r = 0; r < 3; ++r)
    po[r] = m[4 * r + 0] * n.x + m[4 * r + 1] * n.y + m[4 * r + 2] * n.z + m[4 * r + 3] * 0.0f;
  return o;
}
static V3 voxel_center(const Scene &sc, uint32_t gp) {
  const ct_aabb &bb = sc.aabbs[gp];
  return muls(add(load3(bb.minimum), load3(bb.maximum)), 0.5f);
}
// is_hit_from_ray, primitives.glsl:412-453, rayCanStartInBox = true
// (its `oriented` argument is not forwarded: intersect_box runs
// non-oriented with safeInverse of the object-space direction).
static bool hit_from_ray(const Scene &sc, uint32_t ii, uint32_t pid, V3 ro, V3 rd, float &t, V3 &pos, V3 &nrm) {
  if (ii >= sc.inst.size() || pid >= sc.inst[ii].primitive_count) return false;
  const float *m = &sc.w2o[16 * (size_t)ii];
  V3 oo = xform_point(m, ro), od = xform_dir(m, rd);
  V3 rad = v3(kHalfVoxel, kHalfVoxel, kHalfVoxel);
  V3 irad = v3(safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel), safe_inverse(kHalfVoxel));
  V3 inv = v3(safe_inverse(od.x), safe_inverse(od.y), safe_inverse(od.z));
  V3 on;
  bool h = intersect_box(voxel_center(sc, sc.inst[ii].first_primitive_index + pid), rad, irad, oo, od, t, on, true,
                         false, inv);
  pos = add(ro, muls(rd, t));
  nrm = normal_to_world(m, on);
  return h;
}
static uint32_t material_index_of(const Scene &sc, uint32_t ii, uint32_t pid) {
  return sc.prims[sc.inst[ii].first_primitive_index + pid].material_index;
}
static const ct_material &material_or_zero(const Scene &sc, uint32_t idx) {
  static const ct_material zero{};
  return idx < sc.mats.size() ? sc.mats[idx] : zero;
}

// intersect(), bounce.glsl:9-130
static Isect intersect_bounce(const Ctx &c, V3 origin, V3 dir, TraceCount &cnt) {
  const Scene &sc = *c.sc;
  V3 ro = compute_ray_origin(origin, dir);
  ClosestState st = trace_bounce(sc, ro, dir, cnt);
  Isect i;
  i.mat_idx = 0;
  V3 int_hit = v3(0, 0, 0);
  if (st.inst != kInvalid) {
    float t;
    V3 pos, nrm;
    hit_from_ray(sc, st.inst, st.prim, ro, dir, t, pos, nrm);
    int_hit = pos;
    i.is_hit = true;
    i.P = pos;
    i.n = nrm;
    i.inst = st.inst;
    i.prim = st.prim;
    i.mat_idx = material_index_of(sc, st.inst, st.prim);
    i.mat = material_or_zero(sc, i.mat_idx);
    i.distance = length(sub(origin, int_hit));
  }
  i.wo = normalize(sub(origin, int_hit));
  return i;
}

// is_vertex_visible, light.glsl:82-142
static bool vertex_visible(const Ctx &c, V3 o, V3 d, float dist, uint32_t ti, uint32_t tp, TraceCount &cnt) {
  return !trace_any(*c.sc, o, d, dist, ti, tp, cnt);
}

// load_intersection_data_vertex_position, primitives.glsl:567-611
static Isect load_vertex(const Ctx &c, const ct_object_hit &h, V3 world_pos) {
  V3 rd = sub(load3(h.hit), world_pos);
  float t;
  V3 pos, nor;
  Isect i;
  if (!hit_from_ray(*c.sc, h.object.instance_id, h.object.primitive_id, world_pos, rd, t, pos, nor)) return i;
  pos = compute_ray_origin(pos, nor);
  pos = compute_ray_origin(pos, nor);
  i.is_hit = true;
  i.distance = length(sub(world_pos, pos));
  i.P = pos;
  i.n = nor;
  i.wo = normalize(sub(world_pos, pos));
  i.inst = h.object.instance_id;
  i.prim = h.object.primitive_id;
  i.mat_idx = material_index_of(*c.sc, i.inst, i.prim);
  i.mat = material_or_zero(*c.sc, i.mat_idx);
  return i;
}

// random_cosine_direction, prng.glsl:84-108
static V3 random_cosine_direction(uint32_t &seed, V3 normal) {
  float r1 = rnd(seed);
  float r2 = rnd(seed);
  float z = std::sqrt(1.0f - r2);
  float phi = 2.0f * kPiF * r1;
  float sp, cp;
  sincos_ct(phi, sp, cp);
  float x = cp * std::sqrt(r2);
  float y = sp * std::sqrt(r2);
  V3 u = std::fabs(normal.x) > 0.1f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
  u = normalize(cross(normal, u));
  V3 v = cross(normal, u);
  return add(add(muls(u, x), muls(v, y)), muls(normal, z));
}

// The BRDF strategy of RIS, restir_resampling.glsl:190-244: brdf_count
// directions from sample_material (light.glsl:232-270 with the Lambertian
// callable, rcall_scatter.glsl:73-84, which also leaves the direction in
// hit.scatter_dir), each traced by intersect() (bounce.glsl:9-130).  A
// candidate whose closest voxel emits has F = brdf * Le * G with G =
// geom_fact_sa at the hit (NaN, inf and negative G taken as 0) and
// m_i = 1 / max(n_brdf * pdf * G + nee * p_light * G, 1e-6), p_light =
// (1 / cube lights) / (6 VOXEL_EXTENT^2) (sample_lights_pdf,
// light.glsl:272-293), and names the voxel's cube light
// (PRIMITIVE.light_index); any other candidate has weight 0 and light -1.
// All are CUBE candidates.  Returns whether the merge into `reservoir`
// selected the strategy's reservoir.
static bool ris_brdf(const Ctx &c, Res &reservoir, V3 ray_o, Hit &hit, const ct_material &mat, uint32_t nee,
                     uint32_t n_brdf, uint32_t cube_light_count, uint32_t &seed, TraceCount &bcnt) {
  const Scene &sc = *c.sc;
  Res br;
  init_res(br);
  for (uint32_t l = 0; l < sc.cfg.brdf_count; ++l) {
    float w_i = 0.0f;
    V3 F = v3(0, 0, 0);
    const V3 P = hit.world_hit, n = hit.world_nrm;
    const V3 wo = normalize(sub(ray_o, P));
    uint32_t light_index = kInvalid;
    const uint32_t current_seed = seed;
    V3 m_wi = random_cosine_direction(seed, n);
    if (std::fabs(m_wi.x) < 1e-8f && std::fabs(m_wi.y) < 1e-8f && std::fabs(m_wi.z) < 1e-8f) m_wi = n;
    const float mat_pdf = sample_material_pdf(mat, n, wo, m_wi);
    hit.scatter_dir = m_wi;
    TraceCount tc;
    const Isect i = intersect_bounce(c, P, m_wi, tc);
    bcnt.bounce += tc.bounce;
    if (i.is_hit && (i.mat.emission[0] > 0.0f || i.mat.emission[1] > 0.0f || i.mat.emission[2] > 0.0f)) {
      float G = geom_fact_sa(P, i.P, i.n);
      if (std::isnan(G) || std::isinf(G) || G < 0.0f) G = 0.0f;
      const float light_pdf = (1.0f / (float)cube_light_count) / (kVoxelExtent * kVoxelExtent * 6.0f);
      const float m_i = 1.0f / gmax((float)n_brdf * mat_pdf * G + (float)nee * light_pdf * G, 1e-6f);
      const V3 brdf = evaluate_material(mat, n, wo, m_wi);
      F = muls(mulv(brdf, load3(i.mat.emission)), G);
      w_i = luminance(F) * m_i;
      light_index = sc.prims[sc.inst[i.inst].first_primitive_index + i.prim].light_index;
    }
    update_reservoir(br, light_index, CT_LIGHT_CUBE, current_seed, F, w_i, 1.0f, seed);
  }
  return update_reservoir(reservoir, br.Y, br.l_type, br.seed, br.F, br.W_sum, br.M, seed);
}

struct PathBuilder {  // path_builder.glsl:8-17
  ct_object_hit rc_vertex_hit{};
  V3 rc_vertex_wi{0, 0, 0};
  uint32_t cached_random_seed = 0, seed = 0, rc_vertex_length = 0, path_flags = 0;
  V3 cached_jacobian{0, 0, 0};
};
struct PathState {  // path_state.glsl:23-78
  uint32_t max_depth = 0, flags = 0, path_length = 0, bounce_counters = 0;
  float scene_length = 0.0f;
  V3 origin{0, 0, 0}, dir{0, 0, 0};
  float pdf = 0.0f;
  V3 normal{0, 0, 0};
  ct_object_hit hit{};
  V3 thp{1, 1, 1}, prefix_thp{1, 1, 1}, L{0, 0, 0};
  float rr_pdf = 1.0f, prev_scatter_pdf = 1.0f;
  uint32_t seed = 0;
  PathBuilder pb;
  ct_path_reservoir res{};
  bool last_rough = true;
  // Hybrid shift: the path builder picks its reconnection vertex by the
  // hybrid rule (use_hybrid_shift).  A replay (is_replay_for_hybrid_shift)
  // retraces a source path's random numbers up to the vertex before the
  // source's reconnection vertex (replay_rc = its index) and stops there.
  bool hybrid = false, replay = false;
  uint32_t replay_rc = 0;
  bool replay_done = false;  // stopped at vertex replay_rc - 1, invertibly
  Isect replay_prev;         // that vertex
  V3 replay_tp{0, 0, 0};     // path_get_current_thp on arriving there
};
constexpr uint32_t kPfActive = 0x1, kPfHit = 0x2, kPfTransmission = 0x4, kPfDelta = 0x10, kPfUpper = 0x80,
                   kPfLower = 0x100, kPfDiffusePrimary = 0x200;  // defines.glsl:64-76

static inline void path_set_flag(PathState &p, uint32_t f, bool v) {  // path_state.glsl:81-91
  if (v) p.flags |= f;
  else p.flags &= ~f;
}
static inline void set_objhit(ct_object_hit &h, uint32_t i, uint32_t pr, V3 pos) {
  h.object.instance_id = i;
  h.object.primitive_id = pr;
  store3(h.hit, pos);
}

// generate_scatter_ray (Lambertian callable, rcall_scatter.glsl:73-84) and
// path_generate_scatter_ray, path_tracer.glsl:59-112
static bool path_generate_scatter_ray(PathState &p, Isect &i) {
  V3 sd = random_cosine_direction(p.seed, i.n);
  if (std::fabs(sd.x) < 1e-8f && std::fabs(sd.y) < 1e-8f && std::fabs(sd.z) < 1e-8f) sd = i.n;
  i.wi = sd;
  V3 weight = evaluate_material(i.mat, i.n, i.wo, i.wi);
  float pdf = sample_material_pdf(i.mat, i.n, i.wo, i.wi);
  weight = divs(weight, pdf);
  p.dir = i.wi;
  if (p.path_length == 0) path_set_flag(p, kPfDiffusePrimary, true);
  p.thp = mulv(p.thp, weight);
  p.pdf = pdf;
  if (p.path_length == p.pb.rc_vertex_length) p.pb.cached_jacobian.y = pdf;
  p.prev_scatter_pdf = pdf;
  if (pdf <= 0.0f) return false;
  p.bounce_counters += 1u << 0;  // BOUNCE_TYPE_DIFFUSE
  p.normal = i.n;
  bool valid = p.thp.x > 0.0f || p.thp.y > 0.0f || p.thp.z > 0.0f;
  if (valid && p.path_length <= p.pb.rc_vertex_length) {
    p.prefix_thp = p.thp;
    p.thp = v3(1.0f, 1.0f, 1.0f);
  }
  return valid;
}

// path_builder_add_escape_vertex, path_builder.glsl:36-75
static bool add_escape_vertex(PathBuilder &pb, uint32_t path_length, V3 wi, V3 path_weight, V3 postfix,
                              float rr_pdf, float mis_weight, float light_pdf, uint32_t light_type,
                              ct_path_reservoir &res) {
  if (path_length < 1) return false;
  bool is_rc = path_length == pb.rc_vertex_length;
  if (!pres_update(res, path_weight, rr_pdf, pb.seed)) return false;
  res.path_flags = pb.path_flags;
  pf_set_path_length(res.path_flags, path_length);
  pf_set_rc_length(res.path_flags, pb.rc_vertex_length);
  pf_set_last_nee(res.path_flags, false);
  store3(res.rc_vertex_irradiance, postfix);
  store3(res.F, path_weight);
  store3(res.rc_vertex_wi, pb.rc_vertex_wi);
  res.rc_vertex_hit = pb.rc_vertex_hit;
  store3(res.cached_jacobian, pb.cached_jacobian);
  res.rc_random_seed = pb.cached_random_seed;
  if (is_rc) {
    store3(res.rc_vertex_wi, wi);
    store3(res.rc_vertex_irradiance, divs(load3(res.rc_vertex_irradiance), mis_weight));
  }
  res.light_pdf = light_pdf;
  pf_set_light_type(res.path_flags, light_type);
  return true;
}

// path_builder_mark_escape_vertex_as_rc_vertex, path_builder.glsl:114-139:
// the selected escape vertex becomes the reconnection vertex of the
// reservoir's sample (Lambertian: no delta, transmission or specular bits).
static void mark_escape_vertex_as_rc(ct_path_reservoir &res, uint32_t length, const ct_object_hit &hit,
                                     float light_pdf, uint32_t light_type, V3 irradiance, V3 wi,
                                     float prev_scatter_pdf, float geometry_factor) {
  pf_set_rc_length(res.path_flags, length);
  res.rc_vertex_hit = hit;
  res.path_flags &= ~((1u << 20) | (1u << 22));
  res.light_pdf = light_pdf;
  res.cached_jacobian[0] = prev_scatter_pdf;
  pf_set_light_type(res.path_flags, light_type);
  store3(res.rc_vertex_irradiance, irradiance);
  store3(res.rc_vertex_wi, wi);
  res.cached_jacobian[2] = geometry_factor;
}

// eval_voxel_pdf, path_tracer.glsl:185-211
static float eval_voxel_pdf(V3 pos_W, const Isect &i) {
  V3 to_light = sub(i.P, pos_W);
  const float dist_sqr = dot(to_light, to_light);
  if (dist_sqr <= kHlfMin) return 0.0f;
  V3 L = divs(to_light, std::sqrt(dist_sqr));
  float cos_theta = dot(i.n, neg(L));
  if (cos_theta <= 0.0f) return 0.0f;
  const float side = kVoxelExtent;
  const float area = side * side;
  float denom = gmax(kHlfMin, cos_theta * area);
  return dist_sqr / denom;
}

// path_handle_emissive_hit, path_tracer.glsl:258-337
static void path_handle_emissive_hit(const Ctx &c, PathState &p, const Isect &i) {
  float light_pdf = 0.0f, mis = 1.0f;
  V3 Lr = v3(0, 0, 0);
  V3 em = load3(i.mat.emission);
  if (p.path_length != 0) {
    float light_count = (float)c.sc->cube.size();
    light_pdf = (1.0f / light_count) * eval_voxel_pdf(p.origin, i);
    mis = eval_mis(1, p.pdf, 1, light_pdf);
    Lr = muls(mulv(mulv(p.thp, p.prefix_thp), em), mis);
  }
  V3 att = mulv(mulv(p.thp, p.prefix_thp), em);
  if (p.replay) return;  // a replay only carries throughput to its stop vertex
  if (p.path_length > 1) p.L = add(p.L, Lr);
  V3 postfix = muls(mulv(p.thp, att), mis);
  uint32_t plen = (uint32_t)std::max(0, (int)p.path_length - 1);
  const bool selected = add_escape_vertex(p.pb, plen, p.dir, Lr, postfix, p.rr_pdf, mis, light_pdf, CT_LIGHT_CUBE, p.res);
  // an escaped vertex at length >= 2 can be the reconnection vertex
  // (path_tracer.glsl:303-335)
  if (selected && p.hybrid && p.path_length >= 2 && p.path_length < p.pb.rc_vertex_length) {
    const bool far = length(sub(i.P, p.origin)) >= c.p->restir.near_field_distance;
    if (far && p.last_rough) {
      V3 disp = sub(p.origin, i.P);
      float g = std::fabs(dot(i.n, i.wo)) / dot(disp, disp);
      if (std::isnan(g)) g = 0.0f;
      mark_escape_vertex_as_rc(p.res, p.path_length, p.hit, light_pdf, CT_LIGHT_CUBE, em, v3(0, 0, 0),
                               p.prev_scatter_pdf, g);
    }
  }
}

// path_generate_light_sample, path_tracer.glsl:222-256: the light sample
// draws its random numbers; its MIS weight uses ls.pdf, which the shader
// never assigns (0), so the NEE term it feeds to the reservoir is 0 or NaN
// and is rejected without a draw (path_reservoir.glsl:189-190).  The light
// list is the cube lights (the shader's `light_count` is undeclared in the
// prepass, rgen.glsl:281).
static void path_light_sample(const Ctx &c, const Isect &i, uint32_t &seed) {
  const uint32_t n = (uint32_t)c.sc->cube.size();
  uint32_t li = std::min(urnd_interval(seed, 0, n), n - 1u);
  if (n == 0) return;
  Hit h;
  h.world_hit = i.P;
  h.world_nrm = i.n;
  h.distance = i.distance;
  h.scatter_dir = neg(i.wo);
  h.inst = i.inst;
  h.prim = i.prim;
  h.mat_idx = i.mat_idx;
  float pdf = 1.0f / (float)n;
  LightSample ls;
  TraceCount none;
  sample_lights(c, h, c.sc->cube[li], pdf, ls, seed, true, false, none);
}

// path_handle_hit, path_tracer.glsl:599-727 (use_hybrid_shift = false,
// path_state.glsl:245-247)
static void path_handle_hit(const Ctx &c, PathState &p, Isect &i) {
  const ct_restir_params &rp = c.p->restir;
  bool compute_emissive = true;
  if (p.path_length == 1 && !(p.flags & kPfTransmission) && !(p.flags & kPfDelta)) compute_emissive = false;
  V3 em = load3(i.mat.emission);
  if (compute_emissive && (em.x > 0.0f || em.y > 0.0f || em.z > 0.0f)) path_handle_emissive_hit(c, p, i);
  if ((p.bounce_counters & 0xffu) > p.max_depth) {
    path_set_flag(p, kPfActive, false);
    return;
  }
  V3 prev_origin = p.origin;
  p.origin = compute_ray_origin(i.P, i.n);
  p.origin = compute_ray_origin(p.origin, i.n);
  bool cur_rough = i.mat.roughness > rp.roughness_threshold;
  // path_check_rc_vertex, path_tracer.glsl:474-597
  const bool far = length(sub(p.origin, prev_origin)) >= rp.near_field_distance;
  const bool connectable = far && cur_rough && p.last_rough;
  if (!p.replay) {
    const bool can_connect = p.path_length >= 1 && p.path_length < p.pb.rc_vertex_length && connectable;
    if ((!p.hybrid && p.path_length == 1) || (p.hybrid && can_connect)) {
      p.pb.rc_vertex_hit = p.hit;
      p.pb.cached_jacobian.x = p.prev_scatter_pdf;
      V3 disp = sub(prev_origin, i.P);
      p.pb.cached_jacobian.z = std::fabs(dot(i.n, i.wo)) / dot(disp, disp);
      if (std::isnan(p.pb.cached_jacobian.z)) p.pb.cached_jacobian.z = 0.0f;
      p.pb.rc_vertex_length = p.path_length;
    }
  } else {
    // The source path's reconnection vertex is the first connectable one,
    // so each replayed vertex before it must not be connectable; the one
    // right before it must be rough to connect from (invertibility, GRIS).
    bool invertible = !(p.path_length >= 1 && p.path_length + 1 <= p.replay_rc && connectable);
    if (invertible && p.path_length + 1 == p.replay_rc) {
      if (!cur_rough) {
        invertible = false;
      } else {
        p.replay_done = true;
        p.replay_prev = i;
        p.replay_tp = mulv(p.thp, p.prefix_thp);
      }
    }
    if (!invertible || p.replay_done) {
      path_set_flag(p, kPfActive, false);
      return;
    }
  }
  path_set_flag(p, kPfUpper, false);
  path_set_flag(p, kPfLower, false);
  // path_handle_sample_light, path_tracer.glsl:339-438
  if (p.path_length == p.pb.rc_vertex_length) p.pb.cached_random_seed = p.seed;
  path_light_sample(c, i, p.seed);
  path_set_flag(p, kPfUpper, true);
  path_set_flag(p, kPfLower, false);
  if (rp.pt_russian_roulette && p.path_length >= 1 && p.replay) {
    (void)rnd(p.seed);  // a replay draws the roulette's number and goes on (path_tracer.glsl:696-699)
  } else if (rp.pt_russian_roulette && p.path_length >= 1) {
    // terminate_path_by_russian_roulette, path_tracer.glsl:165-177
    float rr = luminance(mulv(p.thp, p.prefix_thp));
    float prob = gmax(0.0f, 1.0f - rr);
    if (rnd(p.seed) < prob) {
      path_set_flag(p, kPfActive, false);
      return;
    }
    p.rr_pdf *= 1.0f - prob;
  }
  bool valid = path_generate_scatter_ray(p, i);
  if (p.path_length == p.pb.rc_vertex_length) p.pb.cached_random_seed = p.seed;
  p.pb.rc_vertex_wi = p.dir;
  p.last_rough = cur_rough;
  if (!(p.flags & kPfActive)) return;
  if ((p.bounce_counters & 0xffu) > p.max_depth) valid = false;
  if (!valid) path_set_flag(p, kPfActive, false);
}

// path_handle_miss, path_tracer.glsl:729-815
static void path_handle_miss(const Ctx &c, PathState &p) {
  if (p.replay) {  // a replay that escapes before its stop vertex is not invertible
    path_set_flag(p, kPfActive, false);
    return;
  }
  float light_pdf = kInv4Pi;  // env_map_sampler_eval_pdf, light.glsl:71
  float mis = eval_mis(1, p.pdf, 1, light_pdf);
  V3 Le = sky_color(c.p->time, c.p->is_afternoon != 0, p.dir);
  V3 Lr = muls(mulv(mulv(p.thp, p.prefix_thp), Le), mis);
  if (p.path_length > 0) p.L = add(p.L, Lr);
  const bool selected = add_escape_vertex(p.pb, p.path_length, p.dir, Lr, muls(mulv(p.thp, Le), mis), p.rr_pdf,
                                          mis, light_pdf, CT_LIGHT_ENV_MAP, p.res);
  // escaped to the sky: a reconnection vertex at infinity (path_tracer.glsl:805-833)
  if (selected && p.hybrid && p.path_length >= 1 && p.path_length + 1 < p.pb.rc_vertex_length && p.last_rough) {
    ct_object_hit none{};
    none.object.instance_id = CT_MAX_INSTANCES;
    none.object.primitive_id = CT_MAX_PRIMITIVES;
    mark_escape_vertex_as_rc(p.res, p.path_length + 1, none, light_pdf, CT_LIGHT_ENV_MAP, Le, p.dir,
                             p.prev_scatter_pdf, 1.0f);
  }
  path_set_flag(p, kPfActive, false);
}

// trace_restir_path_tracing, path_tracer.glsl:854-887
static ct_path_reservoir trace_restir_path(const Ctx &c, Isect i, uint32_t &seed, V3 &color, TraceCount &bcnt) {
  const uint32_t D = c.p->restir.max_depth;
  PathState p;
  // generate_path, path_state.glsl:208-249
  p.max_depth = D;
  p.flags = 0;
  p.path_length = 0;
  path_set_flag(p, kPfActive, true);
  p.bounce_counters = 0;
  set_objhit(p.hit, i.inst, i.prim, i.P);
  path_set_flag(p, kPfHit, true);
  p.seed = seed;
  p.pb.seed = seed;
  p.pb.rc_vertex_hit.object.instance_id = CT_MAX_INSTANCES;
  p.pb.rc_vertex_hit.object.primitive_id = CT_MAX_PRIMITIVES;
  p.pb.rc_vertex_length = D;
  p.hybrid = c.p->restir.pt_shift_mapping == CT_SHIFT_HYBRID;
  pres_init(p.res, D);
  // path_handle_primary_hit, path_tracer.glsl:440-469
  p.origin = compute_ray_origin(i.P, i.n);
  p.origin = compute_ray_origin(p.origin, i.n);
  path_set_flag(p, kPfUpper, false);
  path_set_flag(p, kPfLower, false);
  p.res.init_random_seed = p.seed;
  bool valid = path_generate_scatter_ray(p, i);
  if (p.path_length == p.pb.rc_vertex_length) p.pb.cached_random_seed = p.seed;
  p.pb.rc_vertex_wi = p.dir;
  if (valid) {
    do {
      i = intersect_bounce(c, p.origin, p.dir, bcnt);  // path_next_vertex, :136-149
      if (i.is_hit) {
        set_objhit(p.hit, i.inst, i.prim, i.P);
        path_set_flag(p, kPfHit, true);
        p.path_length++;
        p.scene_length += i.distance;
        path_handle_hit(c, p, i);
      } else {
        path_set_flag(p, kPfHit, false);
        p.scene_length = kHlfMax;
        path_handle_miss(c, p);
      }
    } while (p.flags & kPfActive);
  }
  p.res.M = 1.0f;  // path_builder_finilize
  // path_output, path_tracer.glsl:822-835 (path_reservoir_finalize_RIS)
  float p_hat = lum601(load3(p.res.F));
  if (p_hat == 0.0f || p.res.M == 0.0f)
    p.res.weight = 0.0f;
  else
    p.res.weight = p.res.weight / (p_hat * p.res.M);
  color = muls(load3(p.res.F), p.res.weight);
  if (bad3(color)) color = v3(0, 0, 0);
  seed = p.seed;
  return p.res;
}

// compute_new_ray_origin, primitives.glsl:259-264
static V3 compute_new_ray_origin(V3 pos, V3 n, bool view_side) {
  pos = compute_ray_origin(pos, view_side ? n : neg(n));
  return compute_ray_origin(pos, view_side ? n : neg(n));
}

// compute_shifted_integrand_reconnection, shift.glsl:25-350.  With
// use_hybrid the reconnection vertex index comes from the path flags and a
// near-field connection is rejected; with use_cached the source side's
// geometry term and pdfs come from the reservoir's cached Jacobian, and a
// successful shift leaves the destination's there.
static V3 shift_reconnection(const Ctx &c, float &dst_jacobian, Isect dst, Isect src, ct_path_reservoir &s,
                             bool eval_visibility, TraceCount &acnt, bool use_hybrid = false,
                             bool use_cached = false) {
  dst_jacobian = 0.0f;
  float dst_cached[3] = {0.0f, 0.0f, 0.0f};
  const uint32_t rc_vertex_length = use_hybrid ? pf_rc_length(s.path_flags) : 1u;
  const ct_object_hit &rch = s.rc_vertex_hit;
  V3 irr = load3(s.rc_vertex_irradiance), wi = load3(s.rc_vertex_wi);
  bool exists = rch.object.instance_id < CT_MAX_INSTANCES && rch.object.primitive_id < CT_MAX_PRIMITIVES;
  uint32_t path_length = pf_path_length(s.path_flags);
  bool last_nee = pf_last_nee(s.path_flags);
  bool trans = pf_transmission(s.path_flags, true);
  src.P = compute_new_ray_origin(src.P, src.n, !trans);
  dst.P = compute_new_ray_origin(dst.P, dst.n, !trans);
  if (!exists) {
    V3 integrand = v3(0, 0, 0);
    if (pf_light_type(s.path_flags) == CT_LIGHT_ENV_MAP && path_length + 1 == rc_vertex_length && !last_nee) {
      if (vertex_visible(c, dst.P, wi, kMaxDistance, kInvalid, kInvalid, acnt)) {  // is_segment_visible
        float src_pdf1 = use_cached ? s.cached_jacobian[0] : sample_material_pdf(src.mat, src.n, src.wo, wi);
        float dst_pdf1 = sample_material_pdf(dst.mat, dst.n, dst.wo, wi);
        dst_cached[0] = dst_pdf1;
        V3 dst_f1 = evaluate_material(dst.mat, dst.n, dst.wo, wi);
        float mis = eval_mis(1, dst_pdf1, 1, s.light_pdf);
        integrand = mulv(muls(divs(dst_f1, dst_pdf1), mis), irr);
        dst_jacobian = dst_pdf1 / src_pdf1;
      }
    }
    if (use_cached) store3(s.cached_jacobian, v3(dst_cached[0], dst_cached[1], dst_cached[2]));
    if (jac_invalid(dst_jacobian)) dst_jacobian = 0.0f;
    if (bad3(integrand)) return v3(0, 0, 0);
    return integrand;
  }
  bool is_final = path_length == rc_vertex_length;
  bool escaped = path_length + 1 == rc_vertex_length && !last_nee;
  bool rc_nee = is_final && last_nee;
  if (pf_delta(s.path_flags, true) || pf_delta(s.path_flags, false)) return v3(0, 0, 0);
  Isect rc = load_vertex(c, rch, dst.P);
  V3 dst_cv = neg(rc.wo);
  V3 src_cv = normalize(sub(rc.P, src.P));
  V3 shifted_disp = sub(rc.P, dst.P);
  float shifted_dist2 = dot(shifted_disp, shifted_disp);
  float shifted_cosine = std::fabs(dot(rc.n, neg(dst_cv)));
  if (use_hybrid && !(std::sqrt(shifted_dist2) >= c.p->restir.near_field_distance)) return v3(0, 0, 0);
  float cj_z = shifted_cosine / shifted_dist2;
  dst_cached[2] = cj_z;
  float jacobian;
  if (use_cached) {
    jacobian = cj_z / s.cached_jacobian[2];
  } else {
    V3 original_disp = sub(rc.P, src.P);
    float original_dist2 = dot(original_disp, original_disp);
    float original_cosine = std::fabs(dot(rc.n, neg(src_cv)));
    jacobian = cj_z * original_dist2 / original_cosine;
  }
  if (jac_invalid(jacobian)) return v3(0, 0, 0);
  float dst_pdf1 = sample_material_pdf(dst.mat, dst.n, dst.wo, dst_cv);
  float dst_pdf1_all = dst_pdf1;
  dst_cached[0] = dst_pdf1;
  float src_pdf1 = use_cached ? s.cached_jacobian[0] : sample_material_pdf(src.mat, src.n, src.wo, src_cv);
  jacobian *= dst_pdf1 / src_pdf1;
  if (jac_invalid(jacobian)) return v3(0, 0, 0);
  V3 dst_f1 = evaluate_material(dst.mat, dst.n, dst.wo, dst_cv);
  float dst_rc_all = 0.0f, dst_pdf2 = 1.0f, dst_rc = 1.0f, src_rc = 1.0f;
  if (!escaped) {
    dst_rc = sample_material_pdf(rc.mat, rc.n, rc.wo, wi);
    dst_rc_all = dst_rc;
    dst_cached[1] = dst_rc;
    src_rc = use_cached ? s.cached_jacobian[1] : sample_material_pdf(rc.mat, rc.n, neg(src_cv), wi);
    dst_pdf2 = !rc_nee ? dst_rc : s.light_pdf;
  }
  V3 dst_f2 = v3(1.0f, 1.0f, 1.0f);
  if (!escaped) dst_f2 = evaluate_material(rc.mat, rc.n, rc.wo, wi);
  if ((dst_f1.x == 0.0f && dst_f1.y == 0.0f && dst_f1.z == 0.0f) ||
      (dst_f2.x == 0.0f && dst_f2.y == 0.0f && dst_f2.z == 0.0f))
    return v3(0, 0, 0);
  V3 no_f1 = mulv(divs(dst_f2, dst_pdf2), irr);
  V3 integrand = mulv(divs(dst_f1, dst_pdf1), no_f1);
  if (escaped) integrand = muls(integrand, eval_mis(1, dst_pdf1_all, 1, s.light_pdf));
  if (is_final) {
    if (pf_light_type(s.path_flags) != 4u) {  // GEOMETRY_LIGHT_ANALITIC
      float lp = s.light_pdf;
      float mis = eval_mis(1, rc_nee ? lp : dst_rc_all, 1, rc_nee ? dst_rc_all : lp);
      integrand = muls(integrand, mis);
      no_f1 = muls(no_f1, mis);
      if (!rc_nee) jacobian *= dst_rc / src_rc;
    }
  }
  if (!is_final && !escaped) jacobian *= dst_rc / src_rc;
  if (jac_invalid(jacobian)) return v3(0, 0, 0);
  if (eval_visibility) {
    V3 dir = normalize(sub(rc.P, dst.P));
    float distance = length(sub(dst.P, rc.P));
    if (!vertex_visible(c, dst.P, dir, distance, rc.inst, rc.prim, acnt)) return v3(0, 0, 0);
  }
  if (bad3(integrand)) return v3(0, 0, 0);
  dst_jacobian = jacobian;
  if (use_cached) store3(s.cached_jacobian, v3(dst_cached[0], dst_cached[1], dst_cached[2]));
  return integrand;
}

// RECONNECTION_DATA (shared.inl:380-394) as the retrace produces it: the
// destination path's vertex before the reconnection vertex and the
// throughput that reaches it.
struct RcData {
  bool ok = false;
  Isect prev;
  V3 tp{0, 0, 0};
};

// The retrace of trace_hybrid_shift_rays / trace_random_replay_path_hybrid_simple
// (indirect_illumination.glsl:12-28, path_tracer.glsl:889-966): replay the
// source reservoir's random numbers from the destination's primary
// intersection through the frame's own sampling steps until the vertex
// before the source's reconnection vertex.  The replay is not invertible
// (ok = false) if it misses, stops early, reaches a connectable vertex
// before that one or that vertex is not rough.
static RcData hybrid_retrace(const Ctx &c, Isect i, const ct_path_reservoir &src, TraceCount &bcnt) {
  RcData out;
  const uint32_t D = c.p->restir.max_depth;
  PathState p;
  p.max_depth = D;
  p.flags = 0;
  p.path_length = 0;
  path_set_flag(p, kPfActive, true);
  p.bounce_counters = 0;
  set_objhit(p.hit, i.inst, i.prim, i.P);
  path_set_flag(p, kPfHit, true);
  p.seed = src.init_random_seed;
  p.pb.seed = p.seed;
  p.pb.rc_vertex_hit.object.instance_id = CT_MAX_INSTANCES;
  p.pb.rc_vertex_hit.object.primitive_id = CT_MAX_PRIMITIVES;
  p.pb.rc_vertex_length = pf_rc_length(src.path_flags);
  p.hybrid = true;
  p.replay = true;
  p.replay_rc = pf_rc_length(src.path_flags);
  pres_init(p.res, D);
  p.origin = compute_ray_origin(i.P, i.n);
  p.origin = compute_ray_origin(p.origin, i.n);
  bool valid = path_generate_scatter_ray(p, i);
  p.last_rough = i.mat.roughness > c.p->restir.roughness_threshold;  // generate_replay_primary_scatter_ray
  if (valid) {
    do {
      Isect h = intersect_bounce(c, p.origin, p.dir, bcnt);
      if (h.is_hit) {
        set_objhit(p.hit, h.inst, h.prim, h.P);
        path_set_flag(p, kPfHit, true);
        p.path_length++;
        p.scene_length += h.distance;
        path_handle_hit(c, p, h);
      } else {
        path_set_flag(p, kPfHit, false);
        path_handle_miss(c, p);
      }
    } while (p.flags & kPfActive);
  }
  if (p.replay_done) {
    out.ok = true;
    out.prev = p.replay_prev;
    out.tp = p.replay_tp;
  }
  return out;
}

// compute_shifted_integrand_hybrid, shift.glsl:352-444: random replay up to
// the vertex before the reconnection vertex, then the reconnection from
// there with the cached Jacobian (or from the destination's primary vertex
// when the reconnection vertex is the first bounce).  Two points differ
// from the text, which as written rejects every shift: the Jacobian starts
// at 1 (the replayed prefix has unit Jacobian in primary-sample space) where
// the text sets 0 and multiplies into it (:363, :437), and reconnection
// length 1 reconnects from the destination's primary vertex where the text
// passes an invalid hit (:376-378).  A source whose reconnection vertex
// does not exist (the path ended first) is not shifted.
static V3 hybrid_integrand(const Ctx &c, float &dst_jacobian, const Isect &dst, const Isect &src,
                           ct_path_reservoir &s, bool eval_visibility, TraceCount &acnt) {
  dst_jacobian = 0.0f;
  if (s.weight == 0.0f) return v3(0, 0, 0);
  const uint32_t k = pf_rc_length(s.path_flags), plen = pf_path_length(s.path_flags);
  const bool escaped = plen + 1 == k;
  if (!(k <= plen || escaped) || k == 0) return v3(0, 0, 0);
  V3 tp = v3(1.0f, 1.0f, 1.0f);
  Isect prev = dst;
  if (k > 1) {
    RcData rc = hybrid_retrace(c, dst, s, acnt);
    if (!rc.ok) return v3(0, 0, 0);
    tp = rc.tp;
    prev = rc.prev;
  }
  if (!(tp.x > 0.0f || tp.y > 0.0f || tp.z > 0.0f)) return v3(0, 0, 0);
  float rj = 0.0f;
  const V3 rc_tp = shift_reconnection(c, rj, prev, src, s, eval_visibility, acnt, true, k > 1);
  dst_jacobian = 1.0f * rj;
  return mulv(tp, rc_tp);
}

// compute_shifted_integrand_, shift.glsl:446-483.  `s` is the caller's
// copy of the source reservoir: the hybrid shift leaves the destination's
// cached Jacobian in it.
static V3 shifted_integrand(const Ctx &c, float &dst_jacobian, const Isect &dst, const Isect &src,
                            ct_path_reservoir &s, bool eval_visibility, TraceCount &acnt) {
  dst_jacobian = 0.0f;
  if (s.weight == 0.0f) return v3(0, 0, 0);
  const uint32_t shift = c.p->restir.pt_shift_mapping;
  if (shift == CT_SHIFT_RECONNECTION) return shift_reconnection(c, dst_jacobian, dst, src, s, eval_visibility, acnt);
  if (shift == CT_SHIFT_HYBRID) return hybrid_integrand(c, dst_jacobian, dst, src, s, eval_visibility, acnt);
  return v3(1, 1, 1);
}
// compute_shifted_integrand, shift.glsl:485-497: on a copy of the reservoir.
static V3 shifted_integrand_copy(const Ctx &c, float &dst_jacobian, const Isect &dst, const Isect &src,
                                 const ct_path_reservoir &s, bool eval_visibility, TraceCount &acnt) {
  ct_path_reservoir t = s;
  return shifted_integrand(c, dst_jacobian, dst, src, t, eval_visibility, acnt);
}

// shift_and_merge_reservoir, shift.glsl:485-510 (force_merge)
static bool shift_and_merge(const Ctx &c, float &dst_jacobian, const Isect &dst, ct_path_reservoir &dst_res,
                            const Isect &src, const ct_path_reservoir &src_res, uint32_t &seed, TraceCount &acnt) {
  ct_path_reservoir tmp = src_res;
  V3 integrand = shifted_integrand(c, dst_jacobian, dst, src, tmp, true, acnt);
  bool selected = pres_merge(dst_res, integrand, dst_jacobian, tmp, seed, 1.0f, true);
  if (!selected) store3(dst_res.F, v3(0, 0, 0));
  dst_res.M = src_res.M;
  dst_res.weight = src_res.weight;
  return selected;
}

static void pres_finalize_gris(ct_path_reservoir &r) {  // path_reservoir.glsl:273-279
  float p_hat = lum601(load3(r.F));
  if (p_hat == 0.0f) r.weight = 0.0f;
  else r.weight = r.weight / p_hat;
}

// temporal_path_reuse, indirect_illumination.glsl:147-279
static ct_path_reservoir temporal_path_reuse(const Ctx &c, const ct_path_reservoir &central,
                                             ct_path_reservoir temporal, const Isect &cur_i, const Isect &prev_i,
                                             uint32_t &seed, TraceCount &acnt) {
  ct_path_reservoir dest;
  pres_init(dest, c.p->restir.max_depth);
  const float current_M = central.M;
  temporal.M = gmin(c.p->restir.pt_m_cap * current_M, temporal.M);
  float dst_jacobian = 0.0f;
  for (int i = -1; i <= 0; ++i) {
    float p_sum = 0.0f, p_self = 0.0f;
    ct_path_reservoir temp_dst = dest;
    bool possible = false;
    if (i == -1) {
      temp_dst = central;
      dst_jacobian = 1.0f;
      possible = temp_dst.weight > 0.0f;
    } else {
      possible = shift_and_merge(c, dst_jacobian, cur_i, temp_dst, prev_i, temporal, seed, acnt);
    }
    if (possible) {
      for (int j = -1; j <= 0; ++j) {
        if (j == -1) {
          float cur_p = lum601(load3(temp_dst.F)) * current_M;
          p_sum += cur_p;
          if (i == -1) p_self = cur_p;
        } else {
          if (i == j) {
            p_self = lum601(load3(temporal.F)) / dst_jacobian * temporal.M;
            p_sum += p_self;
            continue;
          }
          float tnj = 0.0f;
          V3 tni = shifted_integrand_copy(c, tnj, prev_i, cur_i, temp_dst, true, acnt);
          float pp = lum601(tni) * tnj;
          p_sum += pp * temporal.M;
        }
      }
    }
    float mis = p_sum == 0.0f ? 0.0f : p_self / p_sum;
    pres_merge_mis(dest, load3(temp_dst.F), dst_jacobian, temp_dst, seed, mis);
  }
  if (dest.weight > 0.0f) pres_finalize_gris(dest);
  if (dest.weight < 0.0f || std::isinf(dest.weight) || std::isnan(dest.weight)) dest.weight = 0.0f;
  return dest;
}

static Isect isect_from_di(const Ctx &c, const ct_di_info &d, bool normalize_wo) {
  Isect i;
  i.is_hit = d.distance > 0.0f;
  i.distance = d.distance;
  i.P = load3(d.position);
  i.n = load3(d.normal);
  i.wo = sub(load3(d.ray_origin), i.P);
  if (normalize_wo) i.wo = normalize(i.wo);
  i.wi = v3(0, 0, 0);
  i.inst = d.instance_hit.instance_id;
  i.prim = d.instance_hit.primitive_id;
  i.mat_idx = d.mat_index;
  i.mat = material_or_zero(*c.sc, d.mat_index);
  return i;
}

// indirect_illumination_spatial_reuse, indirect_illumination.glsl:312-452
static ct_path_reservoir spatial_path_reuse(const Ctx &c, uint32_t x, uint32_t y, const Isect &ci, V3 camera_pos,
                                            uint32_t &seed, TraceCount &acnt) {
  const ct_frame_params &p = *c.p;
  const uint32_t W = p.width, H = p.height;
  const ct_path_reservoir &central = c.fr->pres_o[y * W + x];
  ct_path_reservoir dest;
  pres_init(dest, p.restir.max_depth);
  const int ncount = (int)p.restir.pt_neighbor_count;
  int valid = 0;
  float canonical = 1.0f;
  for (int i = 0; i < ncount; ++i) {
    // get_next_neighbor_pixel, prng.glsl:272-296
    float ox = 2.0f * rnd(seed) - 1.0f;
    float oy = 2.0f * rnd(seed) - 1.0f;
    int nx = (int)x + f2i_sat(ox * (float)p.restir.pt_neighbor_radius);
    int ny = (int)y + f2i_sat(oy * (float)p.restir.pt_neighbor_radius);
    if (nx == (int)x && ny == (int)y) nx = ny = -1;
    if (nx < 0 || ny < 0 || nx >= (int)W || ny >= (int)H) continue;
    const uint32_t ni = (uint32_t)ny * W + (uint32_t)nx;
    Isect nis = isect_from_di(c, c.fr->di_cur[ni], false);
    if (!(nis.inst < CT_MAX_INSTANCES && nis.prim < CT_MAX_PRIMITIVES)) continue;
    // is_valid_geometry, primitives.glsl:547-556
    float cd = length(sub(camera_pos, ci.P)), nd = length(sub(camera_pos, nis.P));
    if (!(dot(ci.n, nis.n) >= 0.5f && std::fabs(cd - nd) < 0.1f * cd)) continue;
    const ct_path_reservoir &nres = c.fr->pres_o[ni];
    valid++;
    float prefix_jac = 0.0f;
    V3 prefix = shifted_integrand_copy(c, prefix_jac, nis, ci, central, true, acnt);
    float prefix_pdf = lum601(prefix) * prefix_jac;
    canonical += 1.0f;
    if (prefix_pdf > 0.0f)
      canonical -= prefix_pdf * nres.M /
                   (prefix_pdf * nres.M + central.M * lum601(load3(central.F)) / (float)ncount);
    ct_path_reservoir tdst = dest;
    float dst_jacobian = 0.0f;
    bool possible = shift_and_merge(c, dst_jacobian, ci, tdst, nis, nres, seed, acnt);
    float nw = 0.0f;
    if (possible) {
      float a = lum601(load3(nres.F)), b = lum601(load3(tdst.F));
      nw = a / dst_jacobian * nres.M / ((a / dst_jacobian) * nres.M + b * central.M / (float)ncount);
      if (std::isnan(nw) || std::isinf(nw)) nw = 0.0f;
    }
    pres_merge_mis(dest, load3(tdst.F), dst_jacobian, tdst, seed, nw);
  }
  pres_merge_mis(dest, load3(central.F), 1.0f, central, seed, canonical);
  if (dest.weight > 0.0f) {
    pres_finalize_gris(dest);
    dest.weight /= (float)(valid + 1);
  }
  if (dest.weight < 0.0f) dest.weight = 0.0f;
  if (std::isnan(dest.weight) || std::isinf(dest.weight)) dest.weight = 0.0f;
  return dest;
}

// get_ray_from_current_pixel, rgen.glsl:18-42 (jitter = 0)
static void camera_ray(const ct_frame_params &p, uint32_t x, uint32_t y, V3 &o, V3 &d) {
  float pcx = ((float)x + 0.0f) + 0.5f, pcy = ((float)y + 0.0f) + 0.5f;
  float ux = pcx / (float)p.width, uy = pcy / (float)p.height;
  float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
  float org[4] = {0, 0, 0, 1}, ov[4];
  mat4_mul_vec(p.inv_view, org, ov);
  o = v3(ov[0], ov[1], ov[2]);
  float tv[4] = {dx, dy, 1.0f, 1.0f}, t4[4];
  mat4_mul_vec(p.inv_proj, tv, t4);
  V3 nt = normalize(v3(t4[0], t4[1], t4[2]));
  float dv[4] = {nt.x, nt.y, nt.z, 0.0f}, d4[4];
  mat4_mul_vec(p.inv_view, dv, d4);
  d = v3(d4[0], d4[1], d4[2]);
}

// The model matrix rchit.glsl:29-33 builds from gl_ObjectToWorldEXT
// (element [0][1] where [1][0] belongs).
static void rchit_model(const float *t, float *m) {
  const float q[16] = {t[0], t[1], t[2], 0.0f,  t[1],  t[5],  t[6],  0.0f,
                       t[8], t[9], t[10], 0.0f, t[12], t[13], t[14], 1.0f};
  std::memcpy(m, q, sizeof(q));
}

// RESTIR_PREPASS_AND_FIRST_VISIBILITY_TEST, rgen.glsl:46-320
static void prepass_pixel(const Ctx &c, uint32_t x, uint32_t y, TraceCount &cnt, TraceCount &pc,
                          TraceCount &ptc, TraceCount &pac) {
  const ct_frame_params &p = *c.p;
  Frame &fr = *c.fr;
  const uint32_t W = p.width, H = p.height;
  const uint32_t idx = y * W + x;
  uint32_t seed = tea(y * W + x, p.frame_number);
  V3 ro, rd;
  camera_ray(p, x, y, ro, rd);
  ClosestState hs = trace_closest(*c.sc, ro, rd, kMaxDistance, pc);
  ct_di_info di{};
  V3 hv = v3(1.0f, 1.0f, 1.0f);  // prd.hit_value
  float distance = -1.0f;
  V3 world_pos = v3(0, 0, 0), world_nrm = v3(0, 0, 0);
  uint32_t hi = CT_MAX_INSTANCES, hp = CT_MAX_PRIMITIVES, mat_index = kInvalid;
  if (hs.inst != kInvalid) {
    // rchit.glsl:18-54 + packed_intersection_info, primitives.glsl:509-534
    const ct_instance &in = c.sc->inst[hs.inst];
    float model[16];
    rchit_model(in.transform, model);
    world_pos = add(ro, muls(rd, hs.t));
    uint32_t gp = in.first_primitive_index + hs.prim;
    const ct_aabb &bb = c.sc->aabbs[gp];
    V3 center = muls(add(load3(bb.minimum), load3(bb.maximum)), 0.5f);
    center = xform_point(model, center);
    world_nrm = cube_like_normal(normalize(sub(world_pos, center)));
    distance = length(sub(world_pos, ro));
    mat_index = c.sc->prims[gp].material_index;
    hi = hs.inst;
    hp = hs.prim;
  } else {
    hv = mulv(hv, sky_color(p.time, p.is_afternoon != 0, rd));  // rmiss.glsl:19-29
  }
  store3(di.position, world_pos);
  di.distance = distance;
  store3(di.normal, world_nrm);
  store3(di.ray_origin, ro);
  di.seed = seed;
  di.instance_hit.instance_id = hi;
  di.instance_hit.primitive_id = hp;
  di.mat_index = mat_index;
  di.confidence = 1.0f;
  const bool is_hit = di.distance > 0.0f;
  float *img = &fr.img[fr.cur][4 * (size_t)idx];
  if (!is_hit) {
    store3(img, hv);
    img[3] = 1.0f;
    if (p.is_active & CT_TAA_BIT) {
      float *t = &fr.taa[4 * (size_t)idx];
      store3(t, hv);
      t[3] = 1.0f;
    }
  } else {
    // motion_vectors.glsl:70-89
    const float *pt = &c.pt[16 * (size_t)di.instance_hit.instance_id];
    float xi_x = (float)x + 0.5f, xi_y = (float)y + 0.5f;
    float S[4] = {di.position[0], di.position[1], di.position[2], 1.0f}, ps[4];
    mat4_mul_vec(pt, S, ps);
    float x1 = ps[0] / ps[3], y1 = ps[1] / ps[3];
    x1 = 0.5f * x1 + 0.5f;
    y1 = 0.5f * y1 + 0.5f;
    float mvx = x1 * (float)W - xi_x, mvy = y1 * (float)H - xi_y;
    fr.velocity[2 * (size_t)idx] = mvx;
    fr.velocity[2 * (size_t)idx + 1] = mvy;
    const ct_material &mat = material(c, di.mat_index);
    uint32_t px = f2u_sat(xi_x + mvx), py = f2u_sat(xi_y + mvy);
    Hit hit = hit_from_di(di);
    // GATHER_TEMPORAL_RESERVOIR, restir_resampling.glsl:269-309
    Res prev;
    init_res(prev);
    uint32_t pidx = px + py * W;
    if (pidx < W * H) prev = res_from(fr.res_prev[pidx]);
    float predicted = clampf(prev.M / p.restir.max_influence_from_the_past, 0.0f, 1.0f);
    float confidence = (prev.W_y > 0.0f) ? predicted : 0.0f;
    di.confidence = confidence;
    Res r = ris(c, confidence, ro, rd, hit, mat, di.seed, cnt, ptc);
    if (prev.W_y > 0.0f) temporal_reuse(c, r, prev, pidx, rd, hit, mat, di.seed, cnt);
    fr.res_int[idx] = res_to(r);
    if (p.restir.indirect) {
      // rgen.glsl:271-315: indirect_illumination_restir_path_tracing
      // (indirect_illumination.glsl:454-475)
      const Isect ci = isect_from_di(c, di, true);
      V3 ind;
      ct_path_reservoir out = trace_restir_path(c, ci, di.seed, ind, ptc);
      if (p.restir.pt_temporal) {
        // temporal_path_get_reprojected, indirect_illumination.glsl:29-57
        const float vx = fr.velocity[2 * (size_t)idx], vy = fr.velocity[2 * (size_t)idx + 1];
        const float r1 = rnd(di.seed);
        const float r2 = rnd(di.seed);
        uint32_t pcx = f2u_sat(((float)x + vx) + r1), pcy = f2u_sat(((float)y + vy) + r2);
        uint32_t prev_idx = std::min(W * H - 1u, pcx + pcy * W);
        const Isect pi = isect_from_di(c, fr.di_prev[prev_idx], false);
        out = temporal_path_reuse(c, out, fr.pres_t[prev_idx], ci, pi, di.seed, pac);
        ind = muls(load3(out.F), out.weight);
      }
      if (bad3(ind)) ind = v3(0, 0, 0);
      fr.pres_o[idx] = out;
      store3(&fr.indirect[3 * (size_t)idx], ind);
    }
  }
  fr.di_cur[idx] = di;
}

// THIRD_VISIBILITY_TEST_AND_SHADING_PASS, rgen.glsl:329-510
static void shade_pixel(const Ctx &c, uint32_t x, uint32_t y, TraceCount &cnt, TraceCount &ptc) {
  const ct_frame_params &p = *c.p;
  Frame &fr = *c.fr;
  const uint32_t idx = y * p.width + x;
  ct_di_info di = fr.di_cur[idx];
  if (di.distance > 0.0f) {
    V3 ray_d = normalize(sub(load3(di.position), load3(di.ray_origin)));
    const ct_material &mat = material(c, di.mat_index);
    Hit hit = hit_from_di(di);
    Res r = res_from(fr.res_int[idx]);
    float confidence = di.confidence;
    Res sr = r;
    spatial_reuse(c, sr, confidence, x, y, ray_d, hit, mat, di.seed, cnt);
    V3 hv = v3(0, 0, 0);
    hv = add(hv, muls(sr.F, sr.W_y));
    if (p.restir.indirect) {
      // rgen.glsl:427-452
      V3 ind;
      if (p.restir.pt_spatial) {
        const Isect ci = isect_from_di(c, di, true);
        const V3 cam = v3(p.inv_view[12], p.inv_view[13], p.inv_view[14]);
        ct_path_reservoir fin = spatial_path_reuse(c, x, y, ci, cam, di.seed, ptc);
        fr.pres_t[idx] = fin;
        ind = muls(load3(fin.F), fin.weight);
      } else {
        if (p.restir.pt_temporal) fr.pres_t[idx] = fr.pres_o[idx];
        ind = load3(&fr.indirect[3 * (size_t)idx]);
      }
      hv = add(hv, ind);
    }
    if (p.restir.direct_emittance) hv = add(hv, load3(mat.emission));
    if (std::isinf(hv.x) || std::isnan(hv.x)) hv.x = 0.0f;
    if (std::isinf(hv.y) || std::isnan(hv.y)) hv.y = 0.0f;
    if (std::isinf(hv.z) || std::isnan(hv.z)) hv.z = 0.0f;
    float *img = &fr.img[fr.cur][4 * (size_t)idx];
    store3(img, hv);
    img[3] = 1.0f;
    if (p.is_active & CT_TAA_BIT) {
      float *t = &fr.taa[4 * (size_t)idx];
      store3(t, hv);
      t[3] = 1.0f;
    }
    fr.res_prev[idx] = res_to(sr);
    fr.di_prev[idx] = di;
  } else {
    fr.di_prev[idx].distance = 0.0f;
  }
}

// taa.glsl:16-108
static void taa_pixel(const Ctx &c, uint32_t x, uint32_t y) {
  const ct_frame_params &p = *c.p;
  Frame &fr = *c.fr;
  const uint32_t W = p.width, H = p.height;
  const uint32_t idx = x + y * W;
  const ct_di_info &di = fr.di_cur[idx];
  if (!(di.distance > 0.0f)) return;
  float b[4] = {0, 0, 0, 0};
  uint32_t count = 0;
  V3 dn = load3(di.normal);
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j) {
      int64_t nx = (int64_t)x + i, ny = (int64_t)y + j;
      if (nx < 0 || nx >= (int64_t)W || ny < 0 || ny >= (int64_t)H) continue;
      uint32_t ni = (uint32_t)nx + (uint32_t)ny * W;
      const ct_di_info &nd = fr.di_cur[ni];
      if (di.distance <= 0.0f || di.mat_index != nd.mat_index ||
          (dot(dn, load3(nd.normal)) < 0.906f) ||
          (std::fabs(di.distance - nd.distance) > 0.1f * di.distance))
        continue;
      ++count;
      const float *t = &fr.taa[4 * (size_t)ni];
      for (int k = 0; k < 4; ++k) b[k] = b[k] + t[k];
    }
  for (int k = 0; k < 4; ++k) b[k] = b[k] / (float)count;
  float vx = fr.velocity[2 * (size_t)idx], vy = fr.velocity[2 * (size_t)idx + 1];
  uint32_t s = di.seed;
  float jx = rnd(s);
  float jy = rnd(s);
  float x1 = ((float)x + vx) + jx, y1 = ((float)y + vy) + jy;
  int32_t ix = f2i_sat(x1), iy = f2i_sat(y1);
  if (ix >= 0 && ix < (int32_t)W && iy >= 0 && iy < (int32_t)H) {
    uint32_t pi = (uint32_t)ix + (uint32_t)iy * W;
    const float *pp = &fr.img[fr.cur ^ 1][4 * (size_t)pi];
    const ct_di_info &pd = fr.di_prev[pi];
    bool valid = dot(load3(pd.normal), dn) >= 0.99f && pd.mat_index == di.mat_index;
    if (valid) {
      const float a = p.restir.taa_alpha, oma = 1.0f - p.restir.taa_alpha;
      for (int k = 0; k < 4; ++k) b[k] = a * b[k] + oma * pp[k];
    }
  }
  float *o = &fr.img[fr.cur][4 * (size_t)idx];
  for (int k = 0; k < 4; ++k) o[k] = b[k];
}

}  // namespace

// ================================================================ adjoint
// Differentiable voxel path tracer (the build's own; ct_abi.h documents
// the model) and its adjoint, restated for the CPU.

struct AdjSample {
  uint32_t g[8];
  V3 beta[8];
  uint32_t nv = 0;
};

static V3 adj_trace(const Scene &sc, const ct_adjoint_desc &d, const ct_adjoint_view &V, uint32_t x, uint32_t y,
                    uint32_t seed, const float *albedo, const float *emission, AdjSample *path) {
  // camera ray through (x + u, y + v)
  const float jx = rnd(seed);
  const float jy = rnd(seed);
  float pcx = (float)x + jx, pcy = (float)y + jy;
  float ux = pcx / (float)V.width, uy = pcy / (float)V.height;
  float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
  float org[4] = {0, 0, 0, 1}, ov[4];
  mat4_mul_vec(V.inv_view, org, ov);
  float tv[4] = {dx, dy, 1.0f, 1.0f}, t4[4];
  mat4_mul_vec(V.inv_proj, tv, t4);
  V3 nt = normalize(v3(t4[0], t4[1], t4[2]));
  float dv[4] = {nt.x, nt.y, nt.z, 0.0f}, d4[4];
  mat4_mul_vec(V.inv_view, dv, d4);
  V3 o = v3(ov[0], ov[1], ov[2]), dir = v3(d4[0], d4[1], d4[2]);
  V3 beta = v3(1, 1, 1), L = v3(0, 0, 0);
  uint32_t nv = 0;
  for (uint32_t k = 0; k <= d.max_depth; ++k) {
    V3 ro = compute_ray_origin(o, dir);
    TraceCount cnt;
    ClosestState st = trace_bounce(sc, ro, dir, cnt);
    if (st.inst == kInvalid) break;
    float t;
    V3 pos, nrm;
    hit_from_ray(sc, st.inst, st.prim, ro, dir, t, pos, nrm);
    uint32_t g = sc.inst[st.inst].first_primitive_index + st.prim;
    if (path) {
      path->g[k] = g;
      path->beta[k] = beta;
    }
    nv = k + 1;
    L = add(L, muls(beta, emission[g]));
    if (k == d.max_depth) break;
    beta = mulv(beta, load3(&albedo[3 * (size_t)g]));
    V3 sd = random_cosine_direction(seed, nrm);
    if (std::fabs(sd.x) < 1e-8f && std::fabs(sd.y) < 1e-8f && std::fabs(sd.z) < 1e-8f) sd = nrm;
    dir = sd;
    o = compute_ray_origin(compute_ray_origin(pos, nrm), nrm);
  }
  if (path) path->nv = nv;
  return L;
}

// ======================================================= plain path tracing
// indirect_illumination_path_tracing (indirect_illumination.glsl:454-539)
// with the material callables (rcall_scatter.glsl:11-88): the adjoint's
// camera rays and segments, a MATERIAL per voxel, emission at every
// vertex, the callable's direction and attenuation.

static V3 reflection(V3 v, V3 n) { return sub(v, muls(n, 2.0f * dot(v, n))); }  // prng.glsl:209-211

static float pow5(float x) {
  const float x2 = x * x;
  return x2 * x2 * x;
}

static float reflectance(float cosine, float ref_idx) {  // prng.glsl:214-218
  float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
  r0 = r0 * r0;
  return r0 + (1.0f - r0) * pow5(1.0f - cosine);
}

static V3 refraction(V3 uv, V3 n, float eta) {  // prng.glsl:220-225
  const float cos_theta = std::fmin(dot(neg(uv), n), 1.0f);
  const V3 perp = muls(add(uv, muls(n, cos_theta)), eta);
  const float ls = perp.x * perp.x + perp.y * perp.y + perp.z * perp.z;
  const V3 par = muls(n, -std::sqrt(std::fabs(1.0f - ls)));
  return add(par, perp);
}

static V3 random_unit_vector(uint32_t &seed, V3 fallback) {  // prng.glsl:63-74, at most 64 draws
  for (int k = 0; k < 64; ++k) {
    const float x = -1.0f + rnd(seed) * 2.0f;
    const float y = -1.0f + rnd(seed) * 2.0f;
    const float z = -1.0f + rnd(seed) * 2.0f;
    const float l2 = x * x + y * y + z * z;
    if (l2 >= 1.0f) continue;
    return normalize(v3(x, y, z));
  }
  return fallback;
}

static bool near_zero(V3 v) { return std::fabs(v.x) < 1e-8f && std::fabs(v.y) < 1e-8f && std::fabs(v.z) < 1e-8f; }

static bool scatter_material(const ct_material &m, V3 dir, V3 nrm, uint32_t &seed, V3 &sd, V3 &att) {
  const uint32_t type = m.type & CT_MATERIAL_TYPE_MASK;
  if (type == CT_MATERIAL_METAL) {
    const V3 refl = reflection(dir, nrm);
    sd = add(muls(random_cosine_direction(seed, refl), std::fmin(m.roughness, 1.0f)), muls(refl, 1.0f - m.roughness));
    att = load3(m.specular);
    return dot(sd, nrm) > 0.0f;
  }
  if (type == CT_MATERIAL_DIELECTRIC) {
    V3 n = nrm;
    float eta = m.ior;
    if (dot(dir, n) > 0.0f) {
      n = neg(n);
      eta = 1.0f / eta;
    }
    const float cos_theta = std::fmin(dot(neg(dir), n), 1.0f);
    const float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
    const bool cannot_refract = eta * sin_theta > 1.0f;
    if (cannot_refract || reflectance(cos_theta, eta) > rnd(seed)) sd = reflection(dir, n);
    else sd = refraction(dir, n, eta);
    att = v3(1.0f, 1.0f, 1.0f);
    return true;
  }
  if (type == CT_MATERIAL_CONSTANT_MEDIUM) {
    sd = random_unit_vector(seed, nrm);
    if (near_zero(sd)) sd = nrm;
  } else {
    sd = random_cosine_direction(seed, nrm);
    if (near_zero(sd)) sd = nrm;
  }
  att = load3(m.diffuse);
  return true;
}

static V3 path_trace_sample(const Scene &sc, const ct_adjoint_desc &d, const ct_adjoint_view &V, uint32_t x, uint32_t y,
                            uint32_t seed, const ct_material *mats, uint32_t n_mats) {
  const float jx = rnd(seed);
  const float jy = rnd(seed);
  float pcx = (float)x + jx, pcy = (float)y + jy;
  float ux = pcx / (float)V.width, uy = pcy / (float)V.height;
  float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
  float org[4] = {0, 0, 0, 1}, ov[4];
  mat4_mul_vec(V.inv_view, org, ov);
  float tv[4] = {dx, dy, 1.0f, 1.0f}, t4[4];
  mat4_mul_vec(V.inv_proj, tv, t4);
  V3 nt = normalize(v3(t4[0], t4[1], t4[2]));
  float dv[4] = {nt.x, nt.y, nt.z, 0.0f}, d4[4];
  mat4_mul_vec(V.inv_view, dv, d4);
  V3 o = v3(ov[0], ov[1], ov[2]), dir = v3(d4[0], d4[1], d4[2]);
  V3 beta = v3(1, 1, 1), L = v3(0, 0, 0);
  for (uint32_t k = 0; k <= d.max_depth; ++k) {
    V3 ro = compute_ray_origin(o, dir);
    TraceCount cnt;
    ClosestState st = trace_bounce(sc, ro, dir, cnt);
    if (st.inst == kInvalid) break;
    float t;
    V3 pos, nrm;
    hit_from_ray(sc, st.inst, st.prim, ro, dir, t, pos, nrm);
    const uint32_t g = sc.inst[st.inst].first_primitive_index + st.prim;
    const uint32_t mi = sc.prims[g].material_index;
    if (mi >= n_mats) break;
    const ct_material &m = mats[mi];
    L = add(L, mulv(beta, load3(m.emission)));
    if (k == d.max_depth) break;
    V3 sd, att;
    if (!scatter_material(m, dir, nrm, seed, sd, att)) break;
    beta = mulv(beta, att);
    const V3 side = ((m.type & CT_MATERIAL_TYPE_MASK) == CT_MATERIAL_LAMBERTIAN || dot(sd, nrm) >= 0.0f) ? nrm : neg(nrm);
    dir = sd;
    o = compute_ray_origin(compute_ray_origin(pos, side), side);
  }
  return L;
}

// ================================================================ C API
struct oc_scene : Scene {};
struct oc_frame : Frame {};

extern "C" {

uint32_t oc_tea(uint32_t v0, uint32_t v1) { return tea(v0, v1); }
uint32_t oc_lcg(uint32_t *state) { return lcg(*state); }
float oc_rnd(uint32_t *state) { return rnd(*state); }
uint32_t oc_urnd_interval(uint32_t *state, uint32_t mn, uint32_t mx) {
  return urnd_interval(*state, mn, mx);
}
void oc_compute_ray_origin(const float pos[3], const float nrm[3], float out[3]) {
  store3(out, compute_ray_origin(load3(pos), load3(nrm)));
}
int oc_intersect_box(const float center[3], float radius, const float origin[3],
                     const float direction[3], int can_start_in_box, int oriented, float *t,
                     float normal[3]) {
  V3 rd = load3(direction);
  V3 inv = v3(safe_inverse(rd.x), safe_inverse(rd.y), safe_inverse(rd.z));
  V3 nrm;
  float tt = 0.0f;
  bool hit = intersect_box(load3(center), v3(radius, radius, radius),
                           v3(safe_inverse(radius), safe_inverse(radius), safe_inverse(radius)),
                           load3(origin), rd, tt, nrm, can_start_in_box != 0, oriented != 0, inv);
  *t = tt;
  store3(normal, nrm);
  return hit ? 1 : 0;
}
void oc_cube_like_normal(const float n[3], float out[3]) {
  store3(out, cube_like_normal(load3(n)));
}
void oc_sky_color(float time, int is_afternoon, const float dir[3], float out[3]) {
  store3(out, sky_color(time, is_afternoon != 0, load3(dir)));
}
uint32_t oc_morton3(uint32_t qx, uint32_t qy, uint32_t qz) { return morton3(qx, qy, qz); }

static void export_bvh(const Bvh &b, ct_bvh_arrays *a) {
  uint32_t n = b.n;
  if (a->keys) std::memcpy(a->keys, b.keys.data(), n * sizeof(uint32_t));
  if (a->ids) std::memcpy(a->ids, b.ids.data(), n * sizeof(uint32_t));
  if (n > 1) {
    if (a->child) std::memcpy(a->child, b.child.data(), b.child.size() * sizeof(uint32_t));
    if (a->boxes) std::memcpy(a->boxes, b.box.data(), b.box.size() * sizeof(float));
  }
  if (a->parent) std::memcpy(a->parent, b.parent.data(), b.parent.size() * sizeof(uint32_t));
  a->n = n;
  a->depth = b.depth;
  a->root_ref = b.root_ref;
}

int oc_lbvh_build(uint32_t n, const float *centers, const float *lo, const float *hi,
                  uint32_t *keys, uint32_t *ids, uint32_t *child, float *boxes,
                  uint32_t *parent) {
  Bvh b;
  if (!build_lbvh(n, centers, lo, hi, b)) return -1;
  ct_bvh_arrays a{keys, ids, child, boxes, parent, 0, 0, 0, 0};
  export_bvh(b, &a);
  return (int)b.depth;
}

oc_scene *oc_scene_create(uint32_t n_inst, const ct_instance *instances,
                          const float *world_to_object, uint32_t prim_capacity,
                          const ct_aabb *aabbs, const ct_primitive *prims, uint32_t n_mat,
                          const ct_material *materials, uint32_t n_cube,
                          const ct_light *cube_lights, uint32_t n_point,
                          const ct_light *point_lights, uint32_t n_env,
                          const ct_light *env_lights, const ct_light_config *config) {
  if (n_inst == 0 || !instances || !world_to_object || !aabbs || !prims || !config) return nullptr;
  for (uint32_t i = 0; i < n_mat; ++i)
    if ((materials[i].type & CT_MATERIAL_TYPE_MASK) != CT_MATERIAL_LAMBERTIAN) return nullptr;
  oc_scene *s = new oc_scene();
  s->inst.assign(instances, instances + n_inst);
  s->w2o.assign(world_to_object, world_to_object + 16 * (size_t)n_inst);
  s->prim_capacity = prim_capacity;
  s->aabbs.assign(aabbs, aabbs + prim_capacity);
  s->prims.assign(prims, prims + prim_capacity);
  if (n_mat) s->mats.assign(materials, materials + n_mat);
  if (n_cube) s->cube.assign(cube_lights, cube_lights + n_cube);
  if (n_point) s->point.assign(point_lights, point_lights + n_point);
  if (n_env) s->env.assign(env_lights, env_lights + n_env);
  s->cfg = *config;
  uint32_t n_used = 0;
  for (uint32_t i = 0; i < n_inst; ++i) {
    n_used += instances[i].primitive_count > 0;
    if ((uint64_t)instances[i].first_primitive_index + instances[i].primitive_count > prim_capacity) {
      delete s;
      return nullptr;
    }
  }
  if (n_used == 0) {
    delete s;
    return nullptr;
  }
  build_scene_bvh(*s);
  return s;
}

int oc_scene_set_previous(oc_scene *s, const float *prev_transforms, const uint32_t *prev_first,
                          const uint32_t *remap_prims, uint32_t n_remap_prims,
                          const uint32_t *remap_lights, uint32_t n_remap_lights) {
  if (!s) return CT_E_INVALID;
  const size_t ni = s->inst.size();
  s->prev_transform.clear();
  s->prev_first.clear();
  if (prev_transforms) s->prev_transform.assign(prev_transforms, prev_transforms + 16 * ni);
  if (prev_first) s->prev_first.assign(prev_first, prev_first + ni);
  s->remap_prim.assign(remap_prims ? remap_prims : nullptr, remap_prims ? remap_prims + n_remap_prims : nullptr);
  s->remap_light.assign(remap_lights ? remap_lights : nullptr, remap_lights ? remap_lights + n_remap_lights : nullptr);
  return CT_OK;
}

void oc_scene_destroy(oc_scene *scene) { delete scene; }

int oc_scene_read_bvh(oc_scene *scene, uint32_t instance, ct_bvh_arrays *arrays) {
  if (!scene || !arrays) return CT_E_INVALID;
  if (instance == kInvalid) {
    export_bvh(scene->tlas, arrays);
    return CT_OK;
  }
  if (instance >= scene->blas.size()) return CT_E_INVALID;
  export_bvh(scene->blas[instance], arrays);
  return CT_OK;
}

static void brute_closest(const Scene &sc, V3 o, V3 d, ClosestState &st) {
  for (uint32_t i = 0; i < sc.inst.size(); ++i) {
    const float *m = &sc.w2o[16 * (size_t)i];
    V3 oo = xform_point(m, o), od = xform_dir(m, d);
    for (uint32_t k = 0; k < sc.inst[i].primitive_count; ++k) closest_leaf(sc, i, k, oo, od, st);
  }
}
static bool brute_any(const Scene &sc, V3 o, V3 d, float dist, uint32_t ti, uint32_t tp) {
  for (uint32_t i = 0; i < sc.inst.size(); ++i) {
    const float *m = &sc.w2o[16 * (size_t)i];
    V3 oo = xform_point(m, o), od = xform_dir(m, d);
    for (uint32_t k = 0; k < sc.inst[i].primitive_count; ++k)
      if (any_leaf(sc, i, k, oo, od, dist, ti, tp)) return true;
  }
  return false;
}

void oc_trace_closest(oc_scene *scene, uint32_t n, const float *origins, const float *directions,
                      float tmax, int brute, float *t_out, uint32_t *ids_out,
                      ct_trace_stats *stats) {
  uint64_t rays = 0, nodes = 0, leaves = 0, insts = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays, nodes, leaves, insts)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    V3 o = load3(&origins[3 * i]), d = load3(&directions[3 * i]);
    ClosestState st{tmax, kInvalid, kInvalid};
    TraceCount cnt;
    if (brute) {
      brute_closest(*scene, o, d, st);
      cnt.rays = 1;
    } else {
      st = trace_closest(*scene, o, d, tmax, cnt);
    }
    bool hit = st.inst != kInvalid;
    t_out[i] = hit ? st.t : -1.0f;
    ids_out[2 * i] = hit ? st.inst : kInvalid;
    ids_out[2 * i + 1] = hit ? st.prim : kInvalid;
    rays += cnt.rays; nodes += cnt.nodes; leaves += cnt.leaves; insts += cnt.instances;
  }
  if (stats) {
    stats->rays = rays; stats->nodes = nodes; stats->leaves = leaves; stats->instances = insts;
  }
}

void oc_trace_bounce(oc_scene *scene, uint32_t n, const float *origins, const float *directions, int brute,
                     uint32_t *ids_out, float *pos_out, float *nrm_out, float *dist_out, float *wo_out,
                     ct_trace_stats *stats) {
  uint64_t rays = 0, nodes = 0, leaves = 0, insts = 0;
  Ctx c;
  c.sc = scene;
  c.fr = nullptr;
  c.p = nullptr;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays, nodes, leaves, insts)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    V3 o = load3(&origins[3 * i]), d = load3(&directions[3 * i]);
    TraceCount cnt;
    Isect h;
    if (brute) {
      // every voxel, same box test and tie rule, then intersect()'s record
      V3 ro = compute_ray_origin(o, d);
      ClosestState st{kMaxDistance, kInvalid, kInvalid};
      for (uint32_t k = 0; k < scene->inst.size(); ++k) {
        const float *m = &scene->w2o[16 * (size_t)k];
        V3 oo = xform_point(m, ro), od = xform_dir(m, d);
        for (uint32_t q = 0; q < scene->inst[k].primitive_count; ++q) closest_leaf<true>(*scene, k, q, oo, od, st);
      }
      h.mat_idx = 0;
      V3 int_hit = v3(0, 0, 0);
      if (st.inst != kInvalid) {
        float t;
        V3 pos, nrm;
        hit_from_ray(*scene, st.inst, st.prim, ro, d, t, pos, nrm);
        int_hit = pos;
        h.is_hit = true;
        h.P = pos;
        h.n = nrm;
        h.inst = st.inst;
        h.prim = st.prim;
        h.distance = length(sub(o, int_hit));
      }
      h.wo = normalize(sub(o, int_hit));
      cnt.bounce = 1;
    } else {
      h = intersect_bounce(c, o, d, cnt);
    }
    ids_out[2 * i] = h.is_hit ? h.inst : kInvalid;
    ids_out[2 * i + 1] = h.is_hit ? h.prim : kInvalid;
    store3(&pos_out[3 * i], h.P);
    store3(&nrm_out[3 * i], h.n);
    dist_out[i] = h.distance;
    store3(&wo_out[3 * i], h.wo);
    rays += cnt.bounce; nodes += cnt.nodes; leaves += cnt.leaves; insts += cnt.instances;
  }
  if (stats) {
    stats->rays = rays; stats->nodes = nodes; stats->leaves = leaves; stats->instances = insts;
  }
}

void oc_sincos(float x, float *s, float *c) { sincos_ct(x, *s, *c); }

int oc_adjoint_render(oc_scene *scene, const ct_adjoint_desc *d, const ct_adjoint_view *views,
                      const float *albedo, const float *emission, float *image) {
  if (!scene || !d || !views || d->spp == 0 || d->max_depth > CT_ADJOINT_MAX_DEPTH) return CT_E_INVALID;
  uint64_t base = 0;
  for (uint32_t v = 0; v < d->n_views; ++v) {
    const ct_adjoint_view &V = views[v];
    const int64_t np = (int64_t)V.width * V.height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < np; ++p) {
      const uint32_t x = (uint32_t)(p % V.width), y = (uint32_t)(p / V.width), gpix = (uint32_t)(base + p);
      V3 sum = v3(0, 0, 0);
      for (uint32_t s = 0; s < d->spp; ++s)
        sum = add(sum, adj_trace(*scene, *d, V, x, y, tea(tea(gpix, d->seed), s), albedo, emission, nullptr));
      store3(&image[3 * (size_t)gpix], divs(sum, (float)d->spp));
    }
    base += (uint64_t)np;
  }
  return CT_OK;
}

int oc_path_trace(oc_scene *scene, const ct_adjoint_desc *d, const ct_adjoint_view *views, const ct_material *mats,
                  uint32_t n_mats, float *image) {
  if (!scene || !d || !views || d->spp == 0 || d->max_depth > CT_ADJOINT_MAX_DEPTH) return CT_E_INVALID;
  uint64_t base = 0;
  for (uint32_t v = 0; v < d->n_views; ++v) {
    const ct_adjoint_view &V = views[v];
    const int64_t np = (int64_t)V.width * V.height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < np; ++p) {
      const uint32_t x = (uint32_t)(p % V.width), y = (uint32_t)(p / V.width), gpix = (uint32_t)(base + p);
      V3 sum = v3(0, 0, 0);
      for (uint32_t s = 0; s < d->spp; ++s)
        sum = add(sum, path_trace_sample(*scene, *d, V, x, y, tea(tea(gpix, d->seed), s), mats, n_mats));
      store3(&image[3 * (size_t)gpix], divs(sum, (float)d->spp));
    }
    base += (uint64_t)np;
  }
  return CT_OK;
}

void oc_scatter_material(const ct_material *m, const float dir[3], const float nrm[3], uint32_t *seed, float sd[3],
                         float att[3], int *ok) {
  V3 s, a;
  *ok = scatter_material(*m, load3(dir), load3(nrm), *seed, s, a) ? 1 : 0;
  store3(sd, s);
  store3(att, a);
}

void oc_refraction(const float uv[3], const float n[3], float eta, float out[3]) {
  store3(out, refraction(load3(uv), load3(n), eta));
}

float oc_reflectance(float cosine, float ref_idx) { return reflectance(cosine, ref_idx); }

// Gradients accumulated in double, then added into the float outputs.
int oc_adjoint_backward(oc_scene *scene, const ct_adjoint_desc *d, const ct_adjoint_view *views,
                        const float *albedo, const float *emission, const float *grad_image,
                        float *grad_albedo, float *grad_emission) {
  if (!scene || !d || !views || d->spp == 0 || d->max_depth > CT_ADJOINT_MAX_DEPTH) return CT_E_INVALID;
  const size_t slots = scene->prim_capacity;
  std::vector<double> ga(3 * slots, 0.0), ge(slots, 0.0);
  uint64_t base = 0;
  for (uint32_t v = 0; v < d->n_views; ++v) {
    const ct_adjoint_view &V = views[v];
    const int64_t np = (int64_t)V.width * V.height;
    for (int64_t p = 0; p < np; ++p) {
      const uint32_t x = (uint32_t)(p % V.width), y = (uint32_t)(p / V.width), gpix = (uint32_t)(base + p);
      const double inv = 1.0 / (double)d->spp;
      const double G[3] = {grad_image[3 * (size_t)gpix] * inv, grad_image[3 * (size_t)gpix + 1] * inv,
                           grad_image[3 * (size_t)gpix + 2] * inv};
      for (uint32_t s = 0; s < d->spp; ++s) {
        AdjSample path;
        adj_trace(*scene, *d, V, x, y, tea(tea(gpix, d->seed), s), albedo, emission, &path);
        double S[3] = {0, 0, 0};
        for (int k = (int)path.nv - 1; k >= 0; --k) {
          const uint32_t g = path.g[k];
          const double b[3] = {path.beta[k].x, path.beta[k].y, path.beta[k].z};
          const double e = emission[g];
          ge[g] += G[0] * b[0] + G[1] * b[1] + G[2] * b[2];
          if ((uint32_t)k < d->max_depth) {
            for (int c = 0; c < 3; ++c) {
              ga[3 * g + c] += G[c] * b[c] * S[c];
              S[c] = e + (double)albedo[3 * (size_t)g + c] * S[c];
            }
          } else {
            S[0] = S[1] = S[2] = e;
          }
        }
      }
    }
    base += (uint64_t)np;
  }
  for (size_t i = 0; i < slots; ++i) {
    grad_emission[i] += (float)ge[i];
    for (int c = 0; c < 3; ++c) grad_albedo[3 * i + c] += (float)ga[3 * i + c];
  }
  return CT_OK;
}

void oc_random_cosine_direction(uint32_t *seed, const float n[3], float out[3]) {
  store3(out, random_cosine_direction(*seed, load3(n)));
}

void oc_trace_any(oc_scene *scene, uint32_t n, const float *origins, const float *directions,
                  const float *distances, const uint32_t *targets, int brute,
                  uint8_t *occluded_out, ct_trace_stats *stats) {
  uint64_t rays = 0, nodes = 0, leaves = 0, insts = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays, nodes, leaves, insts)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    V3 o = load3(&origins[3 * i]), d = load3(&directions[3 * i]);
    uint32_t ti = targets ? targets[2 * i] : kInvalid, tp = targets ? targets[2 * i + 1] : kInvalid;
    TraceCount cnt;
    bool occ;
    if (brute) {
      occ = brute_any(*scene, o, d, distances[i], ti, tp);
      cnt.rays = 1;
    } else {
      occ = trace_any(*scene, o, d, distances[i], ti, tp, cnt);
    }
    occluded_out[i] = occ ? 1 : 0;
    rays += cnt.rays; nodes += cnt.nodes; leaves += cnt.leaves; insts += cnt.instances;
  }
  if (stats) {
    stats->rays = rays; stats->nodes = nodes; stats->leaves = leaves; stats->instances = insts;
  }
}

int oc_frame_set_scene(oc_frame *frame, oc_scene *scene) {
  if (!frame || !scene) return CT_E_INVALID;
  frame->sc = scene;
  return CT_OK;
}

oc_frame *oc_frame_create(oc_scene *scene, uint32_t width, uint32_t height) {
  if (!scene || width == 0 || height == 0) return nullptr;
  oc_frame *f = new oc_frame();
  f->sc = scene;
  f->W = width;
  f->H = height;
  size_t np = (size_t)width * height;
  f->di_cur.assign(np, ct_di_info{});
  f->di_prev.assign(np, ct_di_info{});
  f->res_int.assign(np, ct_reservoir{});
  f->res_prev.assign(np, ct_reservoir{});
  f->velocity.assign(2 * np, 0.0f);
  f->img[0].assign(4 * np, 0.0f);
  f->img[1].assign(4 * np, 0.0f);
  f->taa.assign(4 * np, 0.0f);
  f->pres_o.assign(np, ct_path_reservoir{});
  f->pres_t.assign(np, ct_path_reservoir{});
  f->indirect.assign(3 * np, 0.0f);
  f->cur = 1;
  return f;
}

void oc_frame_destroy(oc_frame *frame) { delete frame; }

int oc_frame_render(oc_frame *frame, const ct_frame_params *params, int threads,
                    uint32_t row_begin, uint32_t row_end) {
  if (!frame || !params) return CT_E_INVALID;
  const ct_frame_params &p = *params;
  if (p.width != frame->W || p.height != frame->H) return CT_E_INVALID;
  if (p.restir.max_neighbors > CT_MAX_NEIGHBORS || p.restir.min_neighbors > CT_MAX_NEIGHBORS)
    return CT_E_INVALID;
  if ((p.is_active & CT_TAA_BIT) && !p.has_previous) return CT_E_INVALID;
  if (row_end > p.height || row_begin > row_end) return CT_E_INVALID;
  if (p.restir.indirect && (p.restir.max_depth == 0 || p.restir.max_depth > CT_MAX_PT_DEPTH ||
                            p.restir.pt_neighbor_count > CT_MAX_PT_NEIGHBORS))
    return CT_E_INVALID;
  Scene &sc = *frame->sc;
  Ctx c;
  c.sc = &sc;
  c.fr = frame;
  c.p = params;
  const size_t ni = sc.inst.size();
  c.pt.resize(16 * ni);
  for (size_t i = 0; i < ni; ++i) {
    float g[16];
    // geometry_T = previous_model * inverse(instance_model), motion_vectors.glsl:146
    const float *prev_model = sc.prev_transform.empty() ? sc.inst[i].transform : &sc.prev_transform[16 * i];
    mat4_mul(prev_model, &sc.w2o[16 * i], g);
    mat4_mul(p.prev_view_proj, g, &c.pt[16 * i]);
  }
  frame->cur ^= 1;
#ifdef _OPENMP
  int nt = threads > 0 ? threads : omp_get_max_threads();
#else
  int nt = 1;
  (void)threads;
#endif
  const uint32_t W = p.width;
  const int64_t rows = (int64_t)(row_end - row_begin);
  const int64_t np = rows * (int64_t)W;
  uint64_t nodes = 0, leaves = 0, insts = 0, cnodes = 0, cleaves = 0, cinsts = 0;
  uint64_t prim_rays = 0, sh_prepass = 0, sh_shade = 0, pt_nodes = 0, pt_leaves = 0, pt_insts = 0;
  uint64_t bounce = 0, pt_shadow = 0;
#pragma omp parallel num_threads(nt) reduction(+ : nodes, leaves, insts, cnodes, cleaves, cinsts, prim_rays, sh_prepass, bounce, pt_shadow, pt_nodes, pt_leaves, pt_insts)
  {
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < np; ++k) {
      uint32_t y = row_begin + (uint32_t)(k / W), x = (uint32_t)(k % W);
      TraceCount cnt, pc, ptc, pac;
      prepass_pixel(c, x, y, cnt, pc, ptc, pac);
      prim_rays += 1;
      sh_prepass += cnt.rays;
      nodes += cnt.nodes; leaves += cnt.leaves; insts += cnt.instances;
      cnodes += pc.nodes; cleaves += pc.leaves; cinsts += pc.instances;
      bounce += ptc.bounce; pt_shadow += pac.rays;
      pt_nodes += ptc.nodes; pt_leaves += ptc.leaves; pt_insts += ptc.instances;
    }
  }
#pragma omp parallel num_threads(nt) reduction(+ : nodes, leaves, insts, sh_shade, pt_shadow)
  {
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < np; ++k) {
      uint32_t y = row_begin + (uint32_t)(k / W), x = (uint32_t)(k % W);
      TraceCount cnt, ptc;
      shade_pixel(c, x, y, cnt, ptc);
      sh_shade += cnt.rays;
      nodes += cnt.nodes; leaves += cnt.leaves; insts += cnt.instances;
      pt_shadow += ptc.rays;
    }
  }
  if (p.is_active & CT_TAA_BIT) {
    uint32_t tw = p.restir.taa_truncate_dispatch ? (p.width / 8) * 8 : p.width;
    uint32_t th = p.restir.taa_truncate_dispatch ? (p.height / 8) * 8 : p.height;
    th = std::min(th, row_end);
    uint32_t tb = row_begin;
    const int64_t tn = (int64_t)(th > tb ? th - tb : 0) * tw;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
    for (int64_t k = 0; k < tn; ++k) taa_pixel(c, (uint32_t)(k % tw), tb + (uint32_t)(k / tw));
  }
  ct_frame_stats &s = frame->stats;
  s = ct_frame_stats{};
  s.primary_rays = prim_rays;
  s.shadow_rays = sh_prepass + sh_shade;
  s.shadow_rays_spatial = sh_shade;
  s.nodes = nodes;
  s.leaves = leaves;
  s.instances = insts;
  s.closest_nodes = cnodes;
  s.closest_leaves = cleaves;
  s.closest_instances = cinsts;
  s.bounce_rays = bounce;
  s.shadow_rays_pt = pt_shadow;
  s.shadow_rays += pt_shadow;
  s.pt_nodes = pt_nodes;
  s.pt_leaves = pt_leaves;
  s.pt_instances = pt_insts;
  return CT_OK;
}

// The frame's direct light with its final reservoirs held fixed, as a
// function of the materials' diffuse albedo d[m] and emission e[m]: pixel
// p with shading material m_p and final reservoir (y, W_y) shows
//   I_p = F_y(d[m_p], Le) * W_y + I_ind,p + e[m_p]   (e: with direct emittance)
// where F_y = brdf * Le * G replays the reservoir's light sample from its
// seed (reservoir_get_radiance, reservoir.glsl:98-112) with the visibility
// the frame found, Le = e[material of the light's voxel] for a cube light
// (a point or environment light keeps its own), and I_ind is the frame's
// indirect term, held fixed.  A channel whose F_y W_y + e is not finite
// (the frame's NaN scrub, rgen.glsl:445-450) shows 0 and has no gradient.
// The image is the shading output, before TAA (the frame's TAA input when
// TAA runs).
struct DiTerm {
  bool lit = false;          // a visible sample with W_y != 0
  uint32_t m_px = kInvalid;  // material of the shading voxel
  uint32_t m_light = kInvalid;
  V3 F{0, 0, 0}, dF_dd{0, 0, 0}, dF_de{0, 0, 0};  // F_y, dF_y/d d[m_px], dF_y/d Le (per channel)
  float W_y = 0.0f;
};
// The material of a cube light's voxel (its owner primitive), or kInvalid.
static uint32_t light_material(const Scene &sc, const ct_light &l) {
  if (l.type != CT_LIGHT_CUBE || l.instance_info.instance_id >= sc.inst.size()) return kInvalid;
  const ct_instance &in = sc.inst[l.instance_info.instance_id];
  if (l.instance_info.primitive_id >= in.primitive_count) return kInvalid;
  const uint64_t slot = (uint64_t)in.first_primitive_index + l.instance_info.primitive_id;
  const uint32_t m = slot < sc.prims.size() ? sc.prims[slot].material_index : kInvalid;
  return m < sc.mats.size() ? m : kInvalid;
}
// `emission` null: the light's own emissive as Le (the adjoint's point, the
// frame's values).
static DiTerm di_term(const Ctx &c, uint32_t idx, const float *diffuse, const float *emission) {
  const Scene &sc = *c.sc;
  const Frame &fr = *c.fr;
  DiTerm t;
  const ct_di_info &di = fr.di_prev[idx];
  t.m_px = di.mat_index;
  const Res r = res_from(fr.res_prev[idx]);
  t.W_y = r.W_y;
  const V3 Fs = r.F;
  if (r.W_y == 0.0f || (Fs.x == 0.0f && Fs.y == 0.0f && Fs.z == 0.0f) || !res_valid(r)) return t;
  const ct_light *l = light_by_type(c, r.l_type, r.Y);
  if (!l || t.m_px >= sc.mats.size()) return t;
  t.m_light = light_material(sc, *l);
  const V3 ray_d = normalize(sub(load3(di.position), load3(di.ray_origin)));
  auto eval = [&](V3 dv, bool own_le, V3 le) {
    ct_material m = sc.mats[t.m_px];
    store3(m.diffuse, dv);
    ct_light L = *l;
    if (!own_le) store3(L.emissive, le);
    Hit hit = hit_from_di(di);
    float pdf_out, G;
    uint32_t s = r.seed;
    TraceCount cnt;
    return calculate_sampled_light(c, ray_d, hit, m, L, 1.0f, pdf_out, G, s, false, false, false, cnt);
  };
  const bool own = t.m_light == kInvalid || !emission;
  const V3 d = load3(&diffuse[3 * (size_t)t.m_px]);
  const V3 le = own ? load3(l->emissive) : load3(&emission[3 * (size_t)t.m_light]);
  t.F = eval(d, own, le);
  t.dF_dd = eval(v3(1, 1, 1), own, le);
  t.dF_de = t.m_light == kInvalid ? v3(0, 0, 0) : eval(d, false, v3(1, 1, 1));
  t.lit = true;
  return t;
}
static Ctx reshade_ctx(oc_frame *frame, const ct_frame_params *params) {
  Ctx c;
  c.sc = frame->sc;
  c.fr = frame;
  c.p = params;
  return c;
}

int oc_frame_reshade(oc_frame *frame, const ct_frame_params *params, const float *diffuse, const float *emission,
                     uint32_t n_mats, float *image) {
  if (!frame || !params || !diffuse || !emission || !image) return CT_E_INVALID;
  if (params->width != frame->W || params->height != frame->H || n_mats != frame->sc->mats.size())
    return CT_E_INVALID;
  const Ctx c = reshade_ctx(frame, params);
  const ct_frame_params &p = *params;
  const size_t np = (size_t)frame->W * frame->H;
  for (size_t idx = 0; idx < np; ++idx) {
    const ct_di_info &di = frame->di_prev[idx];
    float *o = &image[4 * idx];
    if (!(di.distance > 0.0f)) {
      const float *src = (p.is_active & CT_TAA_BIT) ? &frame->taa[4 * idx] : &frame->img[frame->cur][4 * idx];
      for (int k = 0; k < 4; ++k) o[k] = src[k];
      continue;
    }
    const DiTerm t = di_term(c, (uint32_t)idx, diffuse, emission);
    V3 hv = v3(0, 0, 0);
    if (t.lit) hv = add(hv, muls(t.F, t.W_y));
    if (p.restir.indirect) {
      const ct_path_reservoir &fin = frame->pres_t[idx];
      hv = add(hv, p.restir.pt_spatial ? muls(load3(fin.F), fin.weight) : load3(&frame->indirect[3 * idx]));
    }
    if (p.restir.direct_emittance && t.m_px < n_mats) hv = add(hv, load3(&emission[3 * (size_t)t.m_px]));
    if (std::isinf(hv.x) || std::isnan(hv.x)) hv.x = 0.0f;
    if (std::isinf(hv.y) || std::isnan(hv.y)) hv.y = 0.0f;
    if (std::isinf(hv.z) || std::isnan(hv.z)) hv.z = 0.0f;
    store3(o, hv);
    o[3] = 1.0f;
  }
  return CT_OK;
}

int oc_frame_adjoint(oc_frame *frame, const ct_frame_params *params, const float *grad_image, float *grad_diffuse,
                     float *grad_emission) {
  if (!frame || !params || !grad_image || !grad_diffuse || !grad_emission) return CT_E_INVALID;
  if (params->width != frame->W || params->height != frame->H) return CT_E_INVALID;
  const Ctx c = reshade_ctx(frame, params);
  const Scene &sc = *frame->sc;
  const size_t nm = sc.mats.size();
  std::vector<float> d(3 * nm), e(3 * nm);
  for (size_t m = 0; m < nm; ++m)
    for (int k = 0; k < 3; ++k) {
      d[3 * m + k] = sc.mats[m].diffuse[k];
      e[3 * m + k] = sc.mats[m].emission[k];
    }
  const size_t np = (size_t)frame->W * frame->H;
  const bool emit = params->restir.direct_emittance != 0;
  for (size_t idx = 0; idx < np; ++idx) {
    if (!(frame->di_prev[idx].distance > 0.0f)) continue;
    const DiTerm t = di_term(c, (uint32_t)idx, d.data(), nullptr);
    if (t.m_px >= nm) continue;
    const float F[3] = {t.F.x, t.F.y, t.F.z}, Fd[3] = {t.dF_dd.x, t.dF_dd.y, t.dF_dd.z},
                Fe[3] = {t.dF_de.x, t.dF_de.y, t.dF_de.z};
    for (int k = 0; k < 3; ++k) {
      const float g = grad_image[4 * idx + k];
      const float v = (t.lit ? F[k] * t.W_y : 0.0f) + (emit ? e[3 * t.m_px + k] : 0.0f);
      if (std::isinf(v) || std::isnan(v)) continue;
      if (t.lit) {
        grad_diffuse[3 * (size_t)t.m_px + k] += g * (Fd[k] * t.W_y);
        if (t.m_light < nm) grad_emission[3 * (size_t)t.m_light + k] += g * (Fe[k] * t.W_y);
      }
      if (emit) grad_emission[3 * (size_t)t.m_px + k] += g;
    }
  }
  return CT_OK;
}

int oc_frame_read(oc_frame *frame, uint32_t which, void *dst, uint64_t bytes) {
  if (!frame || !dst) return CT_E_INVALID;
  size_t np = (size_t)frame->W * frame->H;
  const void *src = nullptr;
  size_t sz = 0;
  switch (which) {
    case CT_BUF_IMAGE: src = frame->img[frame->cur].data(); sz = 16 * np; break;
    case CT_BUF_PREV_IMAGE: src = frame->img[frame->cur ^ 1].data(); sz = 16 * np; break;
    case CT_BUF_TAA_FRAME: src = frame->taa.data(); sz = 16 * np; break;
    case CT_BUF_DI: src = frame->di_cur.data(); sz = sizeof(ct_di_info) * np; break;
    case CT_BUF_PREV_DI: src = frame->di_prev.data(); sz = sizeof(ct_di_info) * np; break;
    case CT_BUF_INT_RESERVOIR: src = frame->res_int.data(); sz = sizeof(ct_reservoir) * np; break;
    case CT_BUF_PREV_RESERVOIR: src = frame->res_prev.data(); sz = sizeof(ct_reservoir) * np; break;
    case CT_BUF_VELOCITY: src = frame->velocity.data(); sz = 8 * np; break;
    case CT_BUF_PATH_RESERVOIR: src = frame->pres_o.data(); sz = sizeof(ct_path_reservoir) * np; break;
    case CT_BUF_TEMPORAL_PATH_RESERVOIR: src = frame->pres_t.data(); sz = sizeof(ct_path_reservoir) * np; break;
    case CT_BUF_INDIRECT: src = frame->indirect.data(); sz = 12 * np; break;
    default: return CT_E_INVALID;
  }
  if (bytes < sz) return CT_E_INVALID;
  std::memcpy(dst, src, sz);
  return CT_OK;
}

int oc_frame_set_previous_scene(oc_frame *frame, oc_scene *scene) {
  if (!frame) return CT_E_INVALID;
  frame->prev_sc = scene;
  return CT_OK;
}

int oc_frame_get_stats(oc_frame *frame, ct_frame_stats *out) {
  if (!frame || !out) return CT_E_INVALID;
  *out = frame->stats;
  return CT_OK;
}

// ---------------------------------------------------------------- probes
// Single passes of the frame on caller-given inputs, for the tests that pin
// the oracle's ReSTIR arithmetic against independent models.

static Ctx probe_ctx(const oc_scene *scene, const ct_frame_params *p) {
  Ctx c;
  c.sc = scene;
  c.fr = nullptr;
  c.p = p;
  return c;
}

// RIS (restir_resampling.glsl:19-255, with its visibility pass unless the
// BRDF strategy's candidate won) at the shading point `di` describes, the
// camera ray from di.ray_origin through di.position, once per seed.
int oc_probe_ris(oc_scene *scene, const ct_frame_params *p, const ct_di_info *di, float confidence, uint32_t n,
                 const uint32_t *seeds, ct_reservoir *out) {
  if (!scene || !p || !di || (n && (!seeds || !out))) return CT_E_INVALID;
  Ctx c = probe_ctx(scene, p);
  const ct_material &mat = material_or_zero(*scene, di->mat_index);
  const V3 ro = load3(di->ray_origin);
  const V3 rd = normalize(sub(load3(di->position), ro));
  for (uint32_t k = 0; k < n; ++k) {
    Hit hit = hit_from_di(*di);
    uint32_t seed = seeds[k];
    TraceCount cnt, bcnt;
    out[k] = res_to(ris(c, confidence, ro, rd, hit, mat, seed, cnt, bcnt));
  }
  return CT_OK;
}

// Temporal reuse (restir_resampling.glsl:311-416) of the reservoir `r` at
// `di` with the previous frame's reservoir `prev` and DI `prev_di` at the
// predicted pixel, without remapping.
int oc_probe_temporal(oc_scene *scene, const ct_frame_params *p, const ct_di_info *di, const ct_reservoir *r,
                      const ct_di_info *prev_di, const ct_reservoir *prev, uint32_t seed, ct_reservoir *out) {
  if (!scene || !p || !di || !r || !prev_di || !prev || !out) return CT_E_INVALID;
  if (p->is_active & CT_REMAP_BIT) return CT_E_INVALID;
  Frame fr;
  fr.sc = scene;
  fr.W = fr.H = 1;
  fr.di_prev.assign(1, *prev_di);
  Ctx c = probe_ctx(scene, p);
  c.fr = &fr;
  const ct_material &mat = material_or_zero(*scene, di->mat_index);
  const V3 rd = normalize(sub(load3(di->position), load3(di->ray_origin)));
  Hit hit = hit_from_di(*di);
  Res rr = res_from(*r);
  TraceCount cnt;
  temporal_reuse(c, rr, res_from(*prev), 0, rd, hit, mat, seed, cnt);
  *out = res_to(rr);
  return CT_OK;
}

// The hybrid shift's retrace of reservoir `res` from the primary vertex
// `dst`: whether it is invertible, the vertex before the reconnection
// vertex (position, normal, wo, instance, primitive) and the throughput
// reaching it.
int oc_probe_retrace(oc_scene *scene, const ct_frame_params *p, const ct_di_info *dst, const ct_path_reservoir *res,
                     int *ok, float pos[3], float nrm[3], float wo[3], uint32_t ids[2], float tp[3]) {
  if (!scene || !p || !dst || !res || !ok || !pos || !nrm || !wo || !ids || !tp) return CT_E_INVALID;
  Ctx c = probe_ctx(scene, p);
  TraceCount cnt;
  const RcData rc = hybrid_retrace(c, isect_from_di(c, *dst, true), *res, cnt);
  *ok = rc.ok ? 1 : 0;
  store3(pos, rc.prev.P);
  store3(nrm, rc.prev.n);
  store3(wo, rc.prev.wo);
  ids[0] = rc.prev.inst;
  ids[1] = rc.prev.prim;
  store3(tp, rc.tp);
  return CT_OK;
}

// The pairwise MIS combination of spatial reuse (pairwise_mis.glsl:8-114)
// of the reservoir `r` at `center` with the given neighbours, as
// spatial_reuse runs it once its neighbours are chosen.
int oc_probe_pairwise(oc_scene *scene, const ct_frame_params *p, const ct_di_info *center, const ct_reservoir *r,
                      uint32_t num, const ct_di_info *ndi, const ct_reservoir *nres, uint32_t seed,
                      ct_reservoir *out) {
  if (!scene || !p || !center || !r || !out || num == 0 || num > CT_MAX_NEIGHBORS || !ndi || !nres)
    return CT_E_INVALID;
  Ctx c = probe_ctx(scene, p);
  const ct_material &mat = material_or_zero(*scene, center->mat_index);
  const V3 rd = normalize(sub(load3(center->position), load3(center->ray_origin)));
  Res rr = res_from(*r);
  Res nr[CT_MAX_NEIGHBORS];
  const ct_di_info *nd[CT_MAX_NEIGHBORS];
  for (uint32_t i = 0; i < num; ++i) {
    nr[i] = res_from(nres[i]);
    nd[i] = &ndi[i];
  }
  Hit hit = hit_from_di(*center);
  TraceCount cnt;
  pairwise_combine(c, rr, num, nr, nd, rd, hit, mat, seed, cnt);
  *out = res_to(rr);
  return CT_OK;
}

// The shift of path reservoir `res` from the primary vertex `src` to the
// primary vertex `dst` (compute_shifted_integrand_, shift.glsl:446-483),
// with visibility: the shifted integrand and its Jacobian.
int oc_probe_shift(oc_scene *scene, const ct_frame_params *p, const ct_di_info *dst, const ct_di_info *src,
                   int normalize_dst_wo, const ct_path_reservoir *res, float *jacobian, float integrand[3]) {
  if (!scene || !p || !dst || !src || !res || !jacobian || !integrand) return CT_E_INVALID;
  Ctx c = probe_ctx(scene, p);
  const Isect di = isect_from_di(c, *dst, normalize_dst_wo != 0), si = isect_from_di(c, *src, true);
  TraceCount cnt;
  V3 f = shifted_integrand_copy(c, *jacobian, di, si, *res, true, cnt);
  store3(integrand, f);
  return CT_OK;
}

}  // extern "C"
