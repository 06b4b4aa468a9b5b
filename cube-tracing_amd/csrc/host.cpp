// host.cpp -- host-only parts of the library: error strings, camera
// matrices (src/camera.h:75-93, main.cpp:1452-1460), the synthetic voxel
// scene generator (the map loader's role, src/gvox/map_loader.cpp:224-300)
// and load_scene (src/main.cpp:1045-1126).

#include <algorithm>
#include <omp.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ct_internal.h"

namespace ct {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

void mat4_inverse_host(const float *m, float *out);

}  // namespace ct

using namespace ct;

namespace {

void mat4_mul_host(const float *a, const float *b, float *out) {
  float r[16];
  for (int c = 0; c < 4; ++c)
    for (int k = 0; k < 4; ++k)
      r[c * 4 + k] = a[0 * 4 + k] * b[c * 4 + 0] + a[1 * 4 + k] * b[c * 4 + 1] + a[2 * 4 + k] * b[c * 4 + 2] +
                     a[3 * 4 + k] * b[c * 4 + 3];
  std::memcpy(out, r, sizeof(r));
}

// ------------------------------------------------------------ noise
uint32_t tea_host(uint32_t v0, uint32_t v1) {
  uint32_t s0 = 0;
  for (uint32_t n = 0; n < 16; n++) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}

struct Rng {  // NR LCG stream seeded by TEA
  uint32_t s;
  explicit Rng(uint32_t seed) : s(seed) {}
  uint32_t next() {
    s = 1664525u * s + 1013904223u;
    return s;
  }
  double uniform() {  // [0, 1) with 32 random bits from two draws
    uint32_t hi = next() >> 8, lo = next() >> 8;
    return ((double)hi * 16777216.0 + (double)lo) / 281474976710656.0;
  }
  uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % n; }
};

// Lattice of one octave: values tea(seed', hash(ix,iy,iz)) / 2^32 on the
// integer points covering the grid.
struct Lattice {
  int nx, ny, nz;
  std::vector<float> v;
  void build(uint32_t seed, int oct, int gx, int gy, int gz, double freq) {
    nx = (int)std::floor((gx + 0.5) * freq) + 2;
    ny = (int)std::floor((gy + 0.5) * freq) + 2;
    nz = (int)std::floor((gz + 0.5) * freq) + 2;
    v.resize((size_t)nx * ny * nz);
    uint32_t s = tea_host(seed, 0x0C7A0E00u + (uint32_t)oct);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
          v[((size_t)z * ny + y) * nx + x] = (float)((double)tea_host(s, h) / 4294967296.0);
        }
  }
  float at(int x, int y, int z) const { return v[((size_t)z * ny + y) * nx + x]; }
  double eval(double px, double py, double pz) const {
    int ix = (int)std::floor(px), iy = (int)std::floor(py), iz = (int)std::floor(pz);
    double fx = px - ix, fy = py - iy, fz = pz - iz;
    double sx = fx * fx * (3 - 2 * fx), sy = fy * fy * (3 - 2 * fy), sz = fz * fz * (3 - 2 * fz);
    double c00 = at(ix, iy, iz) * (1 - sx) + at(ix + 1, iy, iz) * sx;
    double c10 = at(ix, iy + 1, iz) * (1 - sx) + at(ix + 1, iy + 1, iz) * sx;
    double c01 = at(ix, iy, iz + 1) * (1 - sx) + at(ix + 1, iy, iz + 1) * sx;
    double c11 = at(ix, iy + 1, iz + 1) * (1 - sx) + at(ix + 1, iy + 1, iz + 1) * sx;
    double c0 = c00 * (1 - sy) + c10 * sy, c1 = c01 * (1 - sy) + c11 * sy;
    return c0 * (1 - sz) + c1 * sz;
  }
};

inline uint32_t ord_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

}  // namespace

namespace {
int add_cube_lights(ct_ctx *ctx, uint32_t nl);
}  // namespace

extern "C" {

const char *ct_last_error_message(void) { return ct::g_error.c_str(); }

void ct_restir_params_reference(ct_restir_params *o) {
  if (!o) return;
  *o = ct_restir_params{};
  o->min_ris_point_samples = 1;   // defines.glsl:34
  o->max_ris_point_samples = 4;   // :35
  o->min_ris_env_samples = 1;     // :36
  o->max_ris_env_samples = 4;     // :37
  o->min_ris_cube_samples = 4;    // :38
  o->max_ris_cube_samples = 32;   // :39
  o->max_influence_from_the_past = 20.0f;  // :41
  o->min_neighbors = 3;           // :42
  o->max_neighbors = 4;           // :43
  o->min_neighbor_radius = 2.0f;  // :44
  o->max_neighbor_radius = 50.0f; // :45
  o->max_distance_to_hit = 1e3f;  // :48
  o->taa_alpha = 0.8f;            // taa.glsl:13
  o->taa_truncate_dispatch = 1;   // main.cpp:1642-1643
  o->direct_emittance = 1;        // defines.glsl:19
  o->indirect = 0;                // INDIRECT_ILLUMINATION_ON, defines.glsl:20
  o->max_depth = 3;               // MAX_DEPTH, shared.inl:61 (main.cpp:1330)
  o->pt_temporal = 1;             // RESTIR_PT_TEMPORAL_ON, defines.glsl:15
  o->pt_spatial = 1;              // RESTIR_PT_SPATIAL_ON, defines.glsl:16
  o->pt_shift_mapping = CT_SHIFT_RECONNECTION;  // rgen.glsl:282
  o->pt_neighbor_count = 3;       // NEIGHBOR_COUNT, defines.glsl:61
  o->pt_neighbor_radius = 1;      // NEIGHBOR_RADIUS, defines.glsl:62
  o->pt_m_cap = 20.0f;            // MAX_INFLUENCE_FROM_THE_PAST_THRESHOLD_PT, defines.glsl:58
  o->near_field_distance = 0.015625f * 0.5f;  // NEAR_FIELD_DISTANCE, defines.glsl:53
  o->roughness_threshold = 0.2f;  // SPECULAR_ROUGHNESS_THRESHOLD, defines.glsl:55
  o->pt_russian_roulette = 0;     // USE_RUSSIAN_ROULETTE, defines.glsl:59
}

int ct_mat4_inverse(const float m[16], float out[16]) {
  if (!m || !out) return fail(CT_E_INVALID, "ct_mat4_inverse: null argument");
  mat4_inverse_host(m, out);
  return CT_OK;
}

int ct_camera_matrices(const float pos[3], const float fwd[3], const float up[3], float fov, float aspect,
                       float zn, float zf, float inv_view[16], float inv_proj[16]) {
  if (!pos || !fwd || !up || !inv_view || !inv_proj) return fail(CT_E_INVALID, "ct_camera_matrices: null argument");
  if (!(aspect > 0.0f) || !(zf > zn)) return fail(CT_E_INVALID, "ct_camera_matrices: bad projection");
  // glm::lookAt (right-handed) with center = position + forward.
  float c[3] = {pos[0] + fwd[0], pos[1] + fwd[1], pos[2] + fwd[2]};
  float f[3] = {c[0] - pos[0], c[1] - pos[1], c[2] - pos[2]};
  float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (float &v : f) v = v / fl;
  float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  float sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (float &v : s) v = v / sl;
  float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  float view[16] = {0};
  view[0] = s[0]; view[4] = s[1]; view[8] = s[2];
  view[1] = u[0]; view[5] = u[1]; view[9] = u[2];
  view[2] = -f[0]; view[6] = -f[1]; view[10] = -f[2];
  view[12] = -(s[0] * pos[0] + s[1] * pos[1] + s[2] * pos[2]);
  view[13] = -(u[0] * pos[0] + u[1] * pos[1] + u[2] * pos[2]);
  view[14] = (f[0] * pos[0] + f[1] * pos[1] + f[2] * pos[2]);
  view[15] = 1.0f;
  // glm::perspectiveRH_ZO
  float th = std::tan(fov / 2.0f);
  float proj[16] = {0};
  proj[0] = 1.0f / (aspect * th);
  proj[5] = 1.0f / th;
  proj[10] = zf / (zn - zf);
  proj[11] = -1.0f;
  proj[14] = -(zf * zn) / (zf - zn);
  mat4_inverse_host(view, inv_view);
  mat4_inverse_host(proj, inv_proj);
  inv_proj[5] *= -1.0f;  // Vulkan NDC y flip, main.cpp:1460
  return CT_OK;
}

int ct_camera_prev_view_proj(const float piv[16], const float pip[16], float out[16]) {
  if (!piv || !pip || !out) return fail(CT_E_INVALID, "ct_camera_prev_view_proj: null argument");
  float a[16], b[16];
  mat4_inverse_host(pip, a);
  mat4_inverse_host(piv, b);
  mat4_mul_host(a, b, out);  // inverse(prev_inv_proj) * inverse(prev_inv_view)
  return CT_OK;
}

int ct_scene_generate(const ct_scene_desc *d, ct_scene **out) {
  if (!d || !out) return fail(CT_E_INVALID, "ct_scene_generate: null argument");
  *out = nullptr;
  const uint32_t gx = d->grid[0], gy = d->grid[1], gz = d->grid[2];
  if (!gx || !gy || !gz || !d->chunk[0] || !d->chunk[1] || !d->chunk[2] || gx > 4096 || gy > 4096 || gz > 4096)
    return fail(CT_E_INVALID, "ct_scene_generate: bad grid or chunk size");
  const uint64_t ncell = (uint64_t)gx * gy * gz;
  if (d->occupied_count == 0 || d->occupied_count > ncell || d->occupied_count > CT_MAX_PRIMITIVES)
    return fail(CT_E_INVALID, "ct_scene_generate: occupied_count out of range");
  if (d->emissive_count > d->occupied_count) return fail(CT_E_INVALID, "ct_scene_generate: emissive_count > occupied_count");
  if (d->palette_size > 256 || d->palette_emissive > d->palette_size)
    return fail(CT_E_INVALID, "ct_scene_generate: palette of more than 256 entries or more emissive entries than entries");
  const float e = d->voxel_extent > 0.0f ? d->voxel_extent : CT_VOXEL_EXTENT;
  std::unique_ptr<ct_scene> sc(new ct_scene());
  sc->desc = *d;
  // 3-octave value noise: frequencies 1/16, 1/8, 1/4 per voxel, amplitudes 1, 1/2, 1/4.
  const double freq[3] = {1.0 / 16.0, 1.0 / 8.0, 1.0 / 4.0}, amp[3] = {1.0, 0.5, 0.25};
  Lattice lat[3];
  for (int o = 0; o < 3; ++o) lat[o].build(d->seed, o, (int)gx, (int)gy, (int)gz, freq[o]);
  auto noise_key = [&](uint32_t x, uint32_t y, uint32_t z) {
    double v = 0.0;
    for (int o = 0; o < 3; ++o)
      v += amp[o] * lat[o].eval((x + 0.5) * freq[o], (y + 0.5) * freq[o], (z + 0.5) * freq[o]);
    return ord_bits((float)v);
  };
  // Exactly occupied_count cells: the largest noise values, ties to the
  // smaller linear index.  The cell key is (noise bits << 32) | ~linear
  // index; the threshold key is the occupied_count-th largest, found from a
  // histogram of the top 16 noise bits and an exact selection inside the
  // bin that holds it, so no per-cell array is kept.
  const int nthr = std::max(1, omp_get_max_threads());
  std::vector<uint64_t> hist((size_t)nthr << 16, 0);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < (int64_t)gz; ++z) {
    uint64_t *h = &hist[(size_t)omp_get_thread_num() << 16];
    for (uint32_t y = 0; y < gy; ++y)
      for (uint32_t x = 0; x < gx; ++x) ++h[noise_key(x, y, (uint32_t)z) >> 16];
  }
  uint64_t above = 0;
  uint32_t bin = 0xFFFFu;
  for (;; --bin) {
    uint64_t c = 0;
    for (int t = 0; t < nthr; ++t) c += hist[((size_t)t << 16) + bin];
    if (above + c >= d->occupied_count || bin == 0) break;
    above += c;
  }
  std::vector<std::vector<uint64_t>> cand(nthr);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < (int64_t)gz; ++z) {
    std::vector<uint64_t> &cd = cand[omp_get_thread_num()];
    for (uint32_t y = 0; y < gy; ++y)
      for (uint32_t x = 0; x < gx; ++x) {
        uint32_t k = noise_key(x, y, (uint32_t)z);
        if ((k >> 16) != bin) continue;
        uint64_t lin = ((uint64_t)z * gy + y) * gx + x;
        cd.push_back(((uint64_t)k << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)lin));
      }
  }
  std::vector<uint64_t> sel;
  for (auto &cd : cand) sel.insert(sel.end(), cd.begin(), cd.end());
  std::vector<std::vector<uint64_t>>().swap(cand);
  const uint64_t need = d->occupied_count - above;
  std::nth_element(sel.begin(), sel.begin() + (need - 1), sel.end(), std::greater<uint64_t>());
  const uint64_t thr = sel[need - 1];
  std::vector<uint64_t>().swap(sel);
  // Chunks in z-major order become instances; voxels inside a chunk in
  // z, y, x order become the instance's primitives.  Chunks are filled in
  // parallel and concatenated in chunk order.
  const uint32_t cx = d->chunk[0], cy = d->chunk[1], cz = d->chunk[2];
  const uint32_t ncx = (gx + cx - 1) / cx, ncy = (gy + cy - 1) / cy, ncz = (gz + cz - 1) / cz;
  const int hx = (int)gx / 2, hy = (int)gy / 2, hz = (int)gz / 2;
  const int64_t nchunk = (int64_t)ncx * ncy * ncz;
  std::vector<std::vector<ct_aabb>> chunk_boxes(nchunk);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nchunk; ++c) {
    const uint32_t bx = (uint32_t)(c % ncx), by = (uint32_t)((c / ncx) % ncy), bz = (uint32_t)(c / ((int64_t)ncx * ncy));
    std::vector<ct_aabb> &out_boxes = chunk_boxes[c];
    for (uint32_t z = bz * cz; z < std::min(gz, (bz + 1) * cz); ++z)
      for (uint32_t y = by * cy; y < std::min(gy, (by + 1) * cy); ++y)
        for (uint32_t x = bx * cx; x < std::min(gx, (bx + 1) * cx); ++x) {
          uint64_t lin = ((uint64_t)z * gy + y) * gx + x;
          uint64_t key = ((uint64_t)noise_key(x, y, z) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)lin);
          if (key < thr) continue;
          ct_aabb b;
          b.minimum[0] = (float)((int)x - hx) * e;
          b.minimum[1] = (float)((int)y - hy) * e;
          b.minimum[2] = (float)((int)z - hz) * e;
          b.maximum[0] = (float)((int)x - hx + 1) * e;
          b.maximum[1] = (float)((int)y - hy + 1) * e;
          b.maximum[2] = (float)((int)z - hz + 1) * e;
          out_boxes.push_back(b);
        }
  }
  uint64_t total = 0;
  for (auto &cb : chunk_boxes) total += cb.size();
  if (total != d->occupied_count) return fail(CT_E_STATE, "ct_scene_generate: occupancy selection miscounted");
  sc->aabbs.reserve(total);
  for (auto &cb : chunk_boxes) {
    if (cb.empty()) continue;
    ct_instance in{};
    for (int k = 0; k < 16; ++k) in.transform[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    in.first_primitive_index = (uint32_t)sc->aabbs.size();
    in.primitive_count = (uint32_t)cb.size();
    sc->instances.push_back(in);
    sc->aabbs.insert(sc->aabbs.end(), cb.begin(), cb.end());
    std::vector<ct_aabb>().swap(cb);
  }
  const uint32_t np = (uint32_t)sc->aabbs.size();
  const bool palette = d->palette_size > 0;
  // Palette mode: one uint8 entry per voxel (a hash of the seed and the voxel
  // index), and the palette's emissive entries (a seeded draw).
  std::vector<uint8_t> entry;
  std::vector<uint8_t> entry_emits(palette ? d->palette_size : 0, 0);
  if (palette) {
    entry.resize(np);
    for (uint32_t v = 0; v < np; ++v) entry[v] = (uint8_t)(tea_host(d->seed ^ 0x9A1u, v) % d->palette_size);
    Rng re(tea_host(d->seed, 0xE3Au));
    std::vector<uint32_t> perm(d->palette_size);
    for (uint32_t i = 0; i < d->palette_size; ++i) perm[i] = i;
    for (uint32_t i = 0; i < d->palette_emissive; ++i) {
      uint32_t j = i + (uint32_t)re.below(d->palette_size - i);
      std::swap(perm[i], perm[j]);
      entry_emits[perm[i]] = 1;
    }
  }
  // Emissive voxels: those of an emissive palette entry, or a seeded partial
  // Fisher-Yates draw; kept in voxel order.
  std::vector<uint32_t> lightsel;
  if (palette) {
    for (uint32_t v = 0; v < np; ++v)
      if (entry_emits[entry[v]]) lightsel.push_back(v);
  } else {
    Rng r(tea_host(d->seed, 0xE1u));
    std::vector<uint32_t> perm(np);
    for (uint32_t i = 0; i < np; ++i) perm[i] = i;
    for (uint32_t i = 0; i < d->emissive_count; ++i) {
      uint32_t j = i + (uint32_t)r.below(np - i);
      std::swap(perm[i], perm[j]);
    }
    lightsel.assign(perm.begin(), perm.begin() + d->emissive_count);
    std::sort(lightsel.begin(), lightsel.end());
  }
  // Materials: one per voxel (Lambertian, map_loader.cpp:209-220 fields)
  // or a shared palette.
  Rng ra(tea_host(d->seed, 0xA1Bu));
  auto make_mat = [&](float r, float g, float b) {
    ct_material m{};
    m.type = CT_MATERIAL_LAMBERTIAN;
    m.diffuse[0] = r; m.diffuse[1] = g; m.diffuse[2] = b;
    m.shininess = 1.0f; m.roughness = 1.0f; m.ior = 1.0f; m.dissolve = 1.0f; m.illum = 3;
    return m;
  };
  const float amin = d->albedo_min, aspan = d->albedo_max - d->albedo_min;
  if (palette) {
    sc->mats.resize(d->palette_size);
    Rng rp(tea_host(d->seed, 0xE3Bu));
    for (uint32_t p = 0; p < d->palette_size; ++p) {
      float r = (float)(amin + aspan * ra.uniform()), g = (float)(amin + aspan * ra.uniform()),
            b = (float)(amin + aspan * ra.uniform());
      sc->mats[p] = make_mat(r, g, b);
      if (entry_emits[p])
        for (int a = 0; a < 3; ++a) sc->mats[p].emission[a] = (float)(d->le_min + (d->le_max - d->le_min) * rp.uniform());
    }
  }
  std::vector<uint32_t> light_of(np, CT_INVALID_INDEX);
  Rng rl(tea_host(d->seed, 0x1E11u));
  for (uint32_t l = 0; l < (uint32_t)lightsel.size(); ++l) {
    uint32_t v = lightsel[l];
    light_of[v] = l;
    ct_light L{};
    const ct_aabb &b = sc->aabbs[v];
    for (int a = 0; a < 3; ++a) L.position[a] = (b.minimum[a] + b.maximum[a]) * 0.5f;
    if (palette) {
      for (int a = 0; a < 3; ++a) L.emissive[a] = sc->mats[entry[v]].emission[a];
    } else {
      float le = (float)(d->le_min + (d->le_max - d->le_min) * rl.uniform());
      L.emissive[0] = L.emissive[1] = L.emissive[2] = le;
    }
    // the instance holding voxel v: the last one whose first primitive is <= v
    uint32_t lo = 0, hi = (uint32_t)sc->instances.size();
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) / 2;
      if (sc->instances[mid].first_primitive_index <= v) lo = mid; else hi = mid;
    }
    L.instance_info.instance_id = lo;
    L.instance_info.primitive_id = v - sc->instances[lo].first_primitive_index;
    L.size = e;
    L.type = CT_LIGHT_CUBE;
    sc->lights.push_back(L);
  }
  sc->prims.resize(np);
  if (!palette) {
    sc->mats.resize(np);
    for (uint32_t v = 0; v < np; ++v) {
      float r = (float)(amin + aspan * ra.uniform()), g = (float)(amin + aspan * ra.uniform()),
            b = (float)(amin + aspan * ra.uniform());
      sc->mats[v] = make_mat(r, g, b);
      if (light_of[v] != CT_INVALID_INDEX)
        for (int a = 0; a < 3; ++a) sc->mats[v].emission[a] = sc->lights[light_of[v]].emissive[a];
      sc->prims[v].material_index = v;
      sc->prims[v].light_index = light_of[v];
    }
  } else {
    for (uint32_t v = 0; v < np; ++v) {
      sc->prims[v].material_index = entry[v];
      sc->prims[v].light_index = light_of[v];
    }
  }
  for (int a = 0; a < 3; ++a) {
    sc->gmin[a] = (float)(-(int)d->grid[a] / 2) * e;
    sc->gmax[a] = (float)((int)d->grid[a] - (int)d->grid[a] / 2) * e;
  }
  *out = sc.release();
  return CT_OK;
}

int ct_scene_get_info(const ct_scene *s, ct_scene_info *o) {
  if (!s || !o) return fail(CT_E_INVALID, "ct_scene_get_info: null argument");
  o->instance_count = (uint32_t)s->instances.size();
  o->primitive_count = (uint32_t)s->aabbs.size();
  o->material_count = (uint32_t)s->mats.size();
  o->cube_light_count = (uint32_t)s->lights.size();
  o->instances = s->instances.data();
  o->aabbs = s->aabbs.data();
  o->primitives = s->prims.data();
  o->materials = s->mats.data();
  o->cube_lights = s->lights.data();
  for (int a = 0; a < 3; ++a) {
    o->grid_min[a] = s->gmin[a];
    o->grid_max[a] = s->gmax[a];
  }
  return CT_OK;
}

int ct_scene_destroy(ct_scene *s) {
  if (!s) return fail(CT_E_INVALID, "ct_scene_destroy: null handle");
  delete s;
  return CT_OK;
}

int ct_ctx_append_materials(ct_ctx *ctx, const ct_material *m, uint32_t count, uint32_t *base);
int ct_ctx_get_light_config(ct_ctx *ctx, ct_light_config *out);

int ct_scene_load(const ct_scene *s, ct_accel *a, ct_ctx *ctx, void *stream) {
  return ct_scene_load_transformed(s, a, ctx, nullptr, stream);
}

int ct_scene_load_transformed(const ct_scene *s, ct_accel *a, ct_ctx *ctx, const float transform[16], void *stream) {
  if (!s || !a || !ctx) return fail(CT_E_INVALID, "ct_scene_load: null argument");
  if (ct::deletions_pending(*a))
    return fail(CT_E_STATE, "ct_scene_load: primitive deletions are queued; run update_scene first");
  ct_staging st;
  int rc = ct_accel_get_staging(a, &st);
  if (rc) return rc;
  const uint32_t ni = (uint32_t)s->instances.size(), np = (uint32_t)s->aabbs.size(), nl = (uint32_t)s->lights.size();
  if ((uint64_t)st.host_instance_count + ni > st.max_instance_count) return fail(CT_E_CAPACITY, "ct_scene_load: too many instances");
  if ((uint64_t)st.host_primitive_count + np > st.max_primitive_count) return fail(CT_E_CAPACITY, "ct_scene_load: too many primitives");
  if ((uint64_t)st.cube_light_count + nl > st.max_cube_light_count) return fail(CT_E_CAPACITY, "ct_scene_load: too many cube lights");
  // Instance ids come from the lowest free slots; a fresh or hole-free
  // manager hands out the next consecutive ids.
  uint32_t id_base = 0;
  while (id_base < a->instance_used.size() && a->instance_used[id_base]) ++id_base;
  for (uint32_t i = id_base; i < a->instance_used.size(); ++i)
    if (a->instance_used[i]) return fail(CT_E_UNSUPPORTED, "ct_scene_load: instance id space has holes");
  uint32_t mat_base = 0;
  if ((rc = ct_ctx_append_materials(ctx, s->mats.data(), (uint32_t)s->mats.size(), &mat_base))) return rc;
  const uint32_t ib = st.host_instance_count, pb = st.host_primitive_count, lb = st.cube_light_count;
  for (uint32_t q = 0; q < ni; ++q) {
    ct_instance in = s->instances[q];
    in.first_primitive_index += pb;
    st.instances[ib + q] = in;
  }
  std::memcpy(st.aabbs + pb, s->aabbs.data(), sizeof(ct_aabb) * np);
  for (uint32_t v = 0; v < np; ++v) {
    ct_primitive p = s->prims[v];
    p.material_index += mat_base;
    if (p.light_index != CT_INVALID_INDEX) p.light_index += lb;
    st.primitives[pb + v] = p;
  }
  for (uint32_t l = 0; l < nl; ++l) {
    ct_light L = s->lights[l];
    L.instance_info.instance_id += id_base;
    st.cube_lights[lb + l] = L;
  }
  if ((rc = ct_accel_set_cube_light_count(a, lb + nl))) return rc;
  ct_task t{};
  t.type = CT_TASK_BUILD_BLAS_FROM_CPU;
  t.u.blas_build_from_cpu.instance_count = ni;
  t.u.blas_build_from_cpu.primitive_count = np;
  for (int k = 0; k < 16; ++k)
    t.u.blas_build_from_cpu.transform[k] = transform ? transform[k] : ((k % 5 == 0) ? 1.0f : 0.0f);
  if ((rc = ct_accel_task_queue_add(a, &t))) return rc;
  if ((rc = ct_accel_update_scene(a, 1, stream, nullptr))) return rc;
  return add_cube_lights(ctx, nl);
}

int ct_scene_bind(const ct_scene *s, ct_ctx *ctx) {
  if (!s || !ctx) return fail(CT_E_INVALID, "ct_scene_bind: null argument");
  if (!ctx->mats_host.empty()) return fail(CT_E_STATE, "ct_scene_bind: the context already holds materials");
  uint32_t base = 0;
  int rc;
  if ((rc = ct_ctx_append_materials(ctx, s->mats.data(), (uint32_t)s->mats.size(), &base))) return rc;
  return add_cube_lights(ctx, (uint32_t)s->lights.size());
}

}  // extern "C"

namespace {
// load_lights, main.cpp:418-438 (integer-division pdfs) with
// brdf_count = BRDF_SAMPLING_COUNT when cube lights exist (main.cpp:1317-1320).
int add_cube_lights(ct_ctx *ctx, uint32_t nl) {
  int rc;
  ct_light_config cfg;
  if ((rc = ct_ctx_get_light_config(ctx, &cfg))) return rc;
  cfg.cube_light_count += nl;
  if (cfg.cube_light_count > 0) cfg.brdf_count = 1;
  cfg.light_count = cfg.point_light_count + cfg.cube_light_count + cfg.sphere_light_count + cfg.analytic_light_count +
                    cfg.env_map_count;
  auto ipdf = [&](uint32_t c) { return c == 0 ? 0.0f : (float)(c / cfg.light_count); };
  cfg.point_light_pdf = ipdf(cfg.point_light_count);
  cfg.cube_light_pdf = ipdf(cfg.cube_light_count);
  cfg.sphere_light_pdf = ipdf(cfg.sphere_light_count);
  cfg.analytic_light_pdf = ipdf(cfg.analytic_light_count);
  cfg.env_map_pdf = ipdf(cfg.env_map_count);
  uint32_t geo = cfg.cube_light_count + cfg.sphere_light_count;
  cfg.brdf_pdf = cfg.brdf_count == 0 || geo == 0 ? 0.0f : (float)(cfg.brdf_count / geo);
  return ct_ctx_set_light_config(ctx, &cfg);
}
}  // namespace
