// ct_device.h -- device-side math, RNG, ray-box test and BVH traversal.
//
// Every float expression is written as the sequence of single IEEE
// operations the reference GLSL denotes (dot = (x*x' + y*y') + z*z',
// normalize = v / length(v), mix = x*(1-a) + y*a, ...).  The library is
// compiled with -ffp-contract=off and correctly rounded division and
// square root, so nothing is fused or approximated.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ct_internal.h"

namespace ct {
namespace dev {

#define CT_D __device__ __forceinline__

struct F3 {
  float x, y, z;
};
CT_D F3 f3(float x, float y, float z) { return F3{x, y, z}; }
CT_D F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
CT_D F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
CT_D F3 mulv(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
CT_D F3 muls(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
CT_D F3 divs(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
CT_D F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
CT_D float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CT_D float length(F3 a) { return sqrtf(dot(a, a)); }
CT_D F3 normalize(F3 a) { return divs(a, length(a)); }
CT_D F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
CT_D F3 vabs(F3 a) { return f3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
CT_D float gmin(float a, float b) { return fminf(a, b); }
CT_D float gmax(float a, float b) { return fmaxf(a, b); }
CT_D float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
CT_D float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
CT_D F3 mix3(F3 x, F3 y, float a) { return f3(mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)); }
CT_D float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
CT_D F3 xyz(float4 v) { return f3(v.x, v.y, v.z); }

// Saturating conversions (NaN -> 0).
CT_D uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
CT_D int32_t f2i_sat(float x) {
  if (isnan(x)) return 0;
  if (x >= 2147483648.0f) return 0x7FFFFFFF;
  if (x <= -2147483648.0f) return (int32_t)0x80000000;
  return (int32_t)x;
}

// Column-major 4x4 (m[c*4 + r]): out_r = ((m0r*v0 + m1r*v1) + m2r*v2) + m3r*v3.
CT_D void mat4_vec(const float *m, float v0, float v1, float v2, float v3, float *o) {
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = m[0 * 4 + r] * v0 + m[1 * 4 + r] * v1 + m[2 * 4 + r] * v2 + m[3 * 4 + r] * v3;
}
// 3-row packing (m[c*3 + r]) of the upper three rows.
CT_D F3 xform3(const float *m, F3 p, float w) {
  return f3(m[0] * p.x + m[3] * p.y + m[6] * p.z + m[9] * w,
            m[1] * p.x + m[4] * p.y + m[7] * p.z + m[10] * w,
            m[2] * p.x + m[5] * p.y + m[8] * p.z + m[11] * w);
}

// ---------------------------------------------------------------- RNG
// random.glsl:26-40
CT_D uint32_t tea(uint32_t val0, uint32_t val1) {
  uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
  for (uint32_t n = 0; n < 16; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
// random.glsl:44-56
CT_D uint32_t lcg(uint32_t &prev) {
  prev = 1664525u * prev + 1013904223u;
  return prev & 0x00FFFFFFu;
}
CT_D float rnd(uint32_t &prev) { return (float)lcg(prev) / 16777216.0f; }
CT_D float rnd_interval(uint32_t &prev, float mn, float mx) { return mn + rnd(prev) * (mx - mn); }
// random.glsl:69-77
CT_D uint32_t urnd_interval(uint32_t &prev, uint32_t mn, uint32_t mx) {
  if (mn == mx) return mn;
  if (mn > mx) return f2u_sat(rnd_interval(prev, (float)mx, (float)mn));
  return f2u_sat(rnd_interval(prev, (float)mn, (float)mx));
}

// ------------------------------------------------------------ geometry
CT_D float safe_inverse(float x) { return (x == 0.0f) ? 1e12f : (1.0f / x); }

// primitives.glsl:233-257
CT_D F3 compute_ray_origin(F3 pos, F3 n) {
  const float origin = 1.f / 32.f, f_scale = 1.f / 65536.f, i_scale = 256.f;
  int32_t ix = f2i_sat(n.x * i_scale), iy = f2i_sat(n.y * i_scale), iz = f2i_sat(n.z * i_scale);
  float px = __uint_as_float(__float_as_uint(pos.x) + (uint32_t)(pos.x < 0.0f ? -ix : ix));
  float py = __uint_as_float(__float_as_uint(pos.y) + (uint32_t)(pos.y < 0.0f ? -iy : iy));
  float pz = __uint_as_float(__float_as_uint(pos.z) + (uint32_t)(pos.z < 0.0f ? -iz : iz));
  F3 fo = muls(n, f_scale);
  return f3(fabsf(pos.x) < origin ? pos.x + fo.x : px, fabsf(pos.y) < origin ? pos.y + fo.y : py,
            fabsf(pos.z) < origin ? pos.z + fo.z : pz);
}
// Box.rotation = mat3(1): rotation * v and v * rotation as written.
CT_D F3 rot_mul(F3 v) {
  return f3(1.0f * v.x + 0.0f * v.y + 0.0f * v.z, 0.0f * v.x + 1.0f * v.y + 0.0f * v.z,
            0.0f * v.x + 0.0f * v.y + 1.0f * v.z);
}
CT_D F3 mul_rot(F3 v) {
  return f3(v.x * 1.0f + v.y * 0.0f + v.z * 0.0f, v.x * 0.0f + v.y * 1.0f + v.z * 0.0f,
            v.x * 0.0f + v.y * 0.0f + v.z * 1.0f);
}

// primitives.glsl:273-355 with rotation = identity and radius = h on all
// axes.  `dr` is the (rotated) direction and `sg` = -sign(dr), both
// independent of the box.  oriented: divide by dr; else multiply by inv.
template <bool kCanStart, bool kOriented>
CT_D bool intersect_box(F3 c, F3 ro, F3 dr, F3 sg, F3 inv, float &dist, F3 *normal = nullptr) {
  const float r = kHalfVoxel, ir = 64.0f;  // safeInverse(HALF_VOXEL_EXTENT)
  F3 o = rot_mul(sub(ro, c));
  float winding = 1.0f;
  if (kCanStart) {
    F3 rb = muls(vabs(o), ir);
    winding = (gmax(rb.x, gmax(rb.y, rb.z)) < 1.0f) ? -1.0f : 1.0f;
  }
  float rw = r * winding;
  F3 dtp = sub(f3(rw * sg.x, rw * sg.y, rw * sg.z), o);
  if (kOriented)
    dtp = f3(dtp.x / dr.x, dtp.y / dr.y, dtp.z / dr.z);
  else
    dtp = mulv(dtp, inv);
  // Every term is evaluated (no short circuit): the tests are pure
  // arithmetic, so the result is the same and a wave takes no branches.
  const bool tx = (int)(dtp.x >= 0.0f) & (int)(fabsf(o.y + dr.y * dtp.x) < r) & (int)(fabsf(o.z + dr.z * dtp.x) < r);
  const bool ty = (int)(dtp.y >= 0.0f) & (int)(fabsf(o.z + dr.z * dtp.y) < r) & (int)(fabsf(o.x + dr.x * dtp.y) < r);
  const bool tz = (int)(dtp.z >= 0.0f) & (int)(fabsf(o.x + dr.x * dtp.z) < r) & (int)(fabsf(o.y + dr.y * dtp.z) < r);
  float sx = tx ? sg.x : 0.0f;
  float sy = tx ? 0.0f : (ty ? sg.y : 0.0f);
  float sz = (tx || ty) ? 0.0f : (tz ? sg.z : 0.0f);
  dist = (sx != 0.0f) ? dtp.x : ((sy != 0.0f) ? dtp.y : dtp.z);
  if (normal) *normal = kOriented ? rot_mul(f3(sx, sy, sz)) : f3(sx, sy, sz);
  return (sx != 0.0f) || (sy != 0.0f) || (sz != 0.0f);
}

// primitives.glsl:499-507
CT_D F3 cube_like_normal(F3 n) {
  F3 a = vabs(n);
  float mc = gmax(gmax(a.x, a.y), a.z);
  return (mc == a.x) ? f3(gsign(n.x), 0, 0) : (mc == a.y) ? f3(0, gsign(n.y), 0) : f3(0, 0, gsign(n.z));
}

// mat.glsl:111-202
CT_D F3 sky_grad(F3 dir, F3 a, F3 b) {
  F3 u = normalize(dir);
  float t = 0.5f * (u.y + 1.0f);
  return mix3(a, b, t);
}
CT_D F3 day_color(F3 d) { return sky_grad(d, f3(1.0f, 1.0f, 1.0f), f3(0.5f, 0.7f, 1.0f)); }
CT_D F3 night_color(F3 d) { return sky_grad(d, f3(0.2f, 0.2f, 0.3f), f3(0.2f, 0.2f, 0.5f)); }
CT_D F3 sunset_color(F3 d) { return sky_grad(d, f3(0.8f, 0.4f, 0.0f), f3(1.0f, 0.4f, 0.0f)); }
CT_D F3 sunrise_color(F3 d) { return sky_grad(d, f3(0.7f, 0.35f, 0.0f), f3(0.8f, 0.4f, 0.0f)); }
CT_D F3 sky_color(float t, bool afternoon, F3 dir) {
  const float night_duration = 0.5f, sunrise_duration = 0.10f, full_daylight_transition = 0.10f,
              full_night_transition = 0.10f;
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
// light.glsl:78-80
CT_D float luminance(F3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// sin and cos of x in [0, 2*pi).  GLSL leaves their precision to the
// implementation, so the library fixes one: Cody-Waite reduction by pi/2
// (three-part constant) and the single-precision minimax polynomials of
// the Cephes library (S. L. Moshier, sinf.c / cosf.c) on [-pi/4, pi/4].
// Only IEEE add/mul and floor are used, so the CPU restatement computes
// the same bits.
CT_D void sincos_ct(float x, float &s, float &c) {
  const float j = floorf(x * 0.636619772367581343f + 0.5f);
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

constexpr float kPi = 3.1415926535897932384626433832795f;  // DAXA_PI, shared.inl:49

CT_D bool normal_near_zero(F3 v) {  // prng.glsl:17-21
  const float s = 1e-8f;
  return fabsf(v.x) < s && fabsf(v.y) < s && fabsf(v.z) < s;
}
// random_cosine_direction, prng.glsl:84-108
CT_D F3 random_cosine_direction(uint32_t &seed, F3 n) {
  const float r1 = rnd(seed);
  const float r2 = rnd(seed);
  const float z = sqrtf(1.0f - r2);
  const float phi = 2.0f * kPi * r1;
  float sp, cp;
  sincos_ct(phi, sp, cp);
  const float x = cp * sqrtf(r2);
  const float y = sp * sqrtf(r2);
  F3 u = fabsf(n.x) > 0.1f ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
  u = normalize(cross(n, u));
  F3 v = cross(n, u);
  return add(add(muls(u, x), muls(v, y)), muls(n, z));
}


// ------------------------------------------------------------ traversal
constexpr uint32_t kMaxTraversalSteps = 1u << 22;

struct TraceCnt {
  uint32_t nodes = 0, leaves = 0, insts = 0;
};

CT_D bool slab(float lx, float ly, float lz, float hx, float hy, float hz, F3 o, F3 inv,
               float limit, float &tn) {
  float t1x = (lx - o.x) * inv.x, t2x = (hx - o.x) * inv.x;
  float t1y = (ly - o.y) * inv.y, t2y = (hy - o.y) * inv.y;
  float t1z = (lz - o.z) * inv.z, t2z = (hz - o.z) * inv.z;
  float n0 = gmax(gmax(gmin(t1x, t2x), gmin(t1y, t2y)), gmin(t1z, t2z));
  float f0 = gmin(gmin(gmax(t1x, t2x), gmax(t1y, t2y)), gmax(t1z, t2z));
  tn = n0;
  return (n0 <= f0) && (f0 >= 0.0f) && (n0 <= limit);
}

// Per-lane traversal stack: the newest kStackReg entries in registers
// (r[0] is the top), older ones in this lane's LDS column.
struct Stack {
  uint32_t r[kStackReg > 0 ? kStackReg : 1];
  uint32_t *lds;
  uint32_t stride;
  int n;
  CT_D void push(uint32_t v, bool &overflow) {
    if (kStackReg == 0) {
      if (n >= kStackLds) { overflow = true; return; }
      lds[(uint32_t)n * stride] = v;
      ++n;
      return;
    }
    if (n >= kStackReg) {
      if (n - kStackReg >= kStackLds) { overflow = true; return; }
      lds[(uint32_t)(n - kStackReg) * stride] = r[kStackReg - 1];
    }
#pragma unroll
    for (int i = kStackReg - 1; i > 0; --i) r[i] = r[i - 1];
    r[0] = v;
    ++n;
  }
  CT_D uint32_t pop() {
    if (kStackReg == 0) {
      --n;
      return lds[(uint32_t)n * stride];
    }
    uint32_t v = r[0];
#pragma unroll
    for (int i = 0; i < kStackReg - 1; ++i) r[i] = r[i + 1];
    --n;
    if (n >= kStackReg) r[kStackReg - 1] = lds[(uint32_t)(n - kStackReg) * stride];
    return v;
  }
};

// A load through a pointer known to address device (global) memory: a
// global load instead of a flat one, which would also count against the
// LDS wait counter that the traversal stack uses.
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef uint32_t gu4 __attribute__((ext_vector_type(4)));
CT_D float gld(const float *p) { return *(const __attribute__((address_space(1))) float *)p; }
CT_D uint32_t gld(const uint32_t *p) { return *(const __attribute__((address_space(1))) uint32_t *)p; }
CT_D float4 gld(const float4 *p) {
  const gf4 v = *(const __attribute__((address_space(1))) gf4 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
CT_D uint4 gld(const uint4 *p) {
  const gu4 v = *(const __attribute__((address_space(1))) gu4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Test both children of a node; returns the next ref (or kInvalid to pop)
// and pushes the other child when both are hit.  Closest-hit walks take the
// near child first (ties to child 0), so the best hit prunes early.  Any-hit
// walks take the far child first (ties to child 1): a shadow ray's
// occluders cluster at its target end, where lights sit among other
// voxels, and the outcome does not depend on the order.
template <bool kFarFirst>
CT_D uint32_t visit_node(const DNode *nodes, uint32_t idx, F3 o, F3 inv, float limit, Stack &stk, bool &overflow) {
  const float4 *p = reinterpret_cast<const float4 *>(nodes + idx);
  float4 a = gld(p), b = gld(p + 1), c = gld(p + 2);
  uint4 d = gld(reinterpret_cast<const uint4 *>(nodes + idx) + 3);
  float tn0, tn1;
  bool h0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, o, inv, limit, tn0);
  bool h1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, o, inv, limit, tn1);
  if (h0 && h1) {
    const bool swap = kFarFirst ? !(tn1 < tn0) : (tn1 < tn0);
    stk.push(swap ? d.x : d.y, overflow);
    return swap ? d.y : d.x;
  }
  if (h0) return d.x;
  if (h1) return d.y;
  return kInvalid;
}

// Leaf semantics of a traversal.
//  kTraceAny:     is_vertex_visible (light.glsl:82-142): is_hit_from_ray
//                 with rayCanStartInBox = false; occluded when a non-target
//                 voxel has 0 <= t <= limit.
//  kTraceClosest: traceRayEXT + rint.glsl:10-43 (rayCanStartInBox = true,
//                 oriented): closest t > 0.
//  kTraceBounce:  intersect() of bounce.glsl:9-130 with its is_hit_from_ray
//                 box test (rayCanStartInBox = true, non-oriented): closest
//                 t >= 0.  The reference commits the first generated hit
//                 (bounce.glsl:118); the closest one is taken here.
// Closest modes break equal t by the smaller (instance, primitive).
constexpr int kTraceAny = 0, kTraceClosest = 1, kTraceBounce = 2;

// Two-level traversal (TLAS in world space, each BLAS in object space) as
// a resumable walker: start() takes a ray, each step() visits one TLAS
// node, BLAS node or voxel leaf (a step that reaches an instance also
// enters it and visits the BLAS root), and `done` is set once the ray is
// finished (for an any-hit query `occluded` tells the outcome).  `stk` is
// the lane's LDS stack column.
// The ray a node test uses (cur_o, cur_inv) is the world ray in the TLAS
// and the object-space ray in a BLAS; it is switched on entering and on
// leaving an instance.  Instance entry comes first in a step and the
// stack update last, so in a wave whose lanes take different branches each
// state register is written after every read of its old value: the
// branches' results merge without register copies.
template <int kMode, bool kCount>
struct Walker {
  static constexpr bool kAny = kMode == kTraceAny;
  F3 o, d, winv, cur_o, cur_inv, dr, sg, sinv;
  float limit, best_t;
  uint32_t ti, tp, best_i, best_p, first, inst, ref, steps;
  int base;
  bool in_blas, done, occluded, overflow;
  const DNode *tlas_nodes, *blas_nodes;  // read once per ray, not per step
  Stack st;

  CT_D void start(const DScene &S, F3 o_, F3 d_, float limit_, uint32_t ti_, uint32_t tp_, uint32_t *stk,
                  uint32_t stride) {
    o = o_;
    d = d_;
    limit = limit_;
    ti = ti_;
    tp = tp_;
    winv = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    cur_o = o; cur_inv = winv; dr = d; sg = d; sinv = d;
    best_t = limit_;
    best_i = kInvalid;
    best_p = kInvalid;
    first = 0;
    inst = 0;
    base = 0;
    steps = 0;
    in_blas = false;
    occluded = false;
    overflow = false;
    st.lds = stk;
    st.stride = stride;
    st.n = 0;
    tlas_nodes = S.tlas_nodes;
    blas_nodes = S.blas_nodes;
    ref = S.tlas_root;
    done = ref == kInvalid;
  }

  CT_D void step(const DScene &S, TraceCnt &cnt) {
    if (++steps > kMaxTraversalSteps) {  // only a malformed tree can get here
      overflow = true;
      done = true;
      return;
    }
    if ((ref & kLeaf) && !in_blas) {
      // Enter the instance: the object-space ray, then its BLAS root below.
      inst = gld(S.tlas_inst + (ref & ~kLeaf));
      if (kCount) cnt.insts++;
      const DInst *I = S.inst + inst;
      float m[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) m[k] = gld(I->w2o + k);
      const F3 oo = xform3(m, o, 1.0f), od = xform3(m, d, 0.0f);
      // An instance without rotation or scale gives back the direction's
      // own bits (its zero terms add signed zeros), and then 1 / od is winv:
      // the three divisions run only for lanes whose direction changed.
      F3 oinv = winv;
      if (__float_as_uint(od.x) != __float_as_uint(d.x) || __float_as_uint(od.y) != __float_as_uint(d.y) ||
          __float_as_uint(od.z) != __float_as_uint(d.z))
        oinv = f3(1.0f / od.x, 1.0f / od.y, 1.0f / od.z);
      if (kMode != kTraceClosest) {
        dr = od;
        // safe_inverse(od) with the quotient above
        sinv = f3(od.x == 0.0f ? 1e12f : oinv.x, od.y == 0.0f ? 1e12f : oinv.y, od.z == 0.0f ? 1e12f : oinv.z);
      } else {
        dr = mul_rot(od);
      }
      sg = f3(-gsign(dr.x), -gsign(dr.y), -gsign(dr.z));
      cur_o = oo;
      cur_inv = oinv;
      first = gld(&I->first);
      in_blas = true;
      base = st.n;
      ref = gld(&I->root);
    }
    uint32_t next = kInvalid;
    bool finished = false;
    if (!(ref & kLeaf)) {
      // TLAS and BLAS nodes take the same code.
      if (kCount) cnt.nodes++;
      const DNode *nodes = in_blas ? blas_nodes + first : tlas_nodes;
      next = visit_node<kAny>(nodes, ref, cur_o, cur_inv, kAny ? limit : best_t, st, overflow);
    } else {
      if (kCount) cnt.leaves++;
      const float4 L = gld(S.blas_leaves + first + (ref & ~kLeaf));
      const uint32_t pid = __float_as_uint(L.w);
      float t;
      if (kAny) {
        const bool hit = intersect_box<false, false>(f3(L.x, L.y, L.z), cur_o, dr, sg, sinv, t);
        if (hit && (t >= 0.0f) && (t <= limit) && !(inst == ti && pid == tp)) {
          occluded = true;
          finished = true;
        }
      } else {
        const bool hit = (kMode == kTraceClosest)
                             ? intersect_box<true, true>(f3(L.x, L.y, L.z), cur_o, dr, sg, sinv, t)
                             : intersect_box<true, false>(f3(L.x, L.y, L.z), cur_o, dr, sg, sinv, t);
        if (hit && ((kMode == kTraceClosest) ? (t > 0.0f) : (t >= 0.0f))) {
          const bool better =
              (t < best_t) || (t == best_t && ((inst < best_i) || (inst == best_i && pid < best_p)));
          if (better) {
            best_t = t;
            best_i = inst;
            best_p = pid;
          }
        }
      }
    }
    if (finished || overflow) {
      done = true;
    } else if (next != kInvalid) {
      ref = next;
    } else {
      if (in_blas && st.n == base) {
        in_blas = false;
        cur_o = o;
        cur_inv = winv;
      }
      if (st.n == 0)
        done = true;
      else
        ref = st.pop();
    }
  }
};

// The whole walk of one ray.  Returns true when an any-hit query is
// occluded; a closest query leaves its hit in best_t/best_i/best_p (which
// enter as the current limit and an invalid hit).
template <int kMode, bool kCount>
CT_D bool traverse_mode(const DScene &S, F3 o, F3 d, float limit, uint32_t ti, uint32_t tp,
                        float &best_t, uint32_t &best_i, uint32_t &best_p, uint32_t *stk,
                        uint32_t stride, TraceCnt &cnt, bool &overflow) {
  Walker<kMode, kCount> w;
  w.start(S, o, d, limit, ti, tp, stk, stride);
  w.best_t = best_t;
  w.best_i = best_i;
  w.best_p = best_p;
  while (!w.done) w.step(S, cnt);
  best_t = w.best_t;
  best_i = w.best_i;
  best_p = w.best_p;
  overflow = overflow || w.overflow;
  return w.occluded && !w.overflow;
}

template <bool kAny, bool kCount>
CT_D bool traverse(const DScene &S, F3 o, F3 d, float limit, uint32_t ti, uint32_t tp,
                   float &best_t, uint32_t &best_i, uint32_t &best_p, uint32_t *stk,
                   uint32_t stride, TraceCnt &cnt, bool &overflow) {
  return traverse_mode<kAny ? kTraceAny : kTraceClosest, kCount>(S, o, d, limit, ti, tp, best_t, best_i, best_p, stk,
                                                                  stride, cnt, overflow);
}

}  // namespace dev
}  // namespace ct
