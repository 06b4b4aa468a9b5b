// vox.cpp -- MagicaVoxel .vox scenes: the map loader's voxel -> AABB /
// PRIMITIVE / MATERIAL / LIGHT rules (src/gvox/map_loader.cpp:97-300) over
// this library's own parser of the public VOX chunk format.  The reference
// reads the file through gvox's "magicavoxel" parse adapter
// (map_loader.cpp:404-407, 479-493); gvox is not part of the reference, so
// the parts it decides (region partitioning, placement of models, which
// voxels are emissive) are fixed here as follows:
//
//  * Chunks: "VOX " + version, then MAIN whose children are read in file
//    order.  SIZE + XYZI pairs are models; RGBA is the palette (entry i is
//    color index i + 1); without RGBA the format's default palette is used.
//    MATL (version 200 material dictionaries) and MATT (version 150
//    material records) mark color indices as emissive: MATL `_type` =
//    `_emit` with `_emit` > 0, or MATT type 3 with weight > 0.
//  * Placement: with a scene graph (nTRN / nGRP / nSHP, walked depth-first
//    from node 0, child order kept) a model of size s under the summed
//    nTRN translations t (frame 0's `_t`) has its voxel v at
//    v - floor(s / 2) + t.  Rotations other than the identity are rejected.
//    Without a scene graph, or when one model is asked for, voxel v sits at
//    v.  Within one model a cell written twice keeps the later XYZI entry;
//    color index 0 is empty.
//  * Regions: each model is one region, or it is cut into chunk-aligned
//    boxes (floor division of the placed coordinates) visited in z, y, x
//    order.  Each non-empty region becomes one instance with an identity
//    transform (map_loader.cpp:317-325).  Empty regions make no instance.
//
// Inside a region the loader's own rules apply (map_loader.cpp:106-311):
// cells in z, y, x order; the palette is deduplicated per region, the
// first voxel of a color index adding a Lambertian material with diffuse
// rgb / 255 and emission (rgb / 255) * flux for emissive colors; the axis
// remap maps the placed coordinates to the grid; AABB = [g e, (g + 1) e];
// an emissive voxel adds a cube light at g e + e / 2 with size e whose
// instance_info is (instance, primitive within the instance).  Capacity
// overflows, which the reference drops silently (map_loader.cpp:200,
// 284-289, 299-305), are errors here.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "ct_internal.h"

using namespace ct;

namespace {

struct Reader {
  const uint8_t *p;
  size_t n, off = 0;
  bool ok = true;
  Reader(const uint8_t *p_, size_t n_) : p(p_), n(n_) {}
  bool have(size_t k) const { return ok && off <= n && n - off >= k; }
  int32_t i32() {
    if (!have(4)) {
      ok = false;
      return 0;
    }
    int32_t v;
    std::memcpy(&v, p + off, 4);
    off += 4;
    return v;
  }
  uint8_t u8() {
    if (!have(1)) {
      ok = false;
      return 0;
    }
    return p[off++];
  }
  std::string str() {
    const int32_t len = i32();
    if (len < 0 || !have((size_t)len)) {
      ok = false;
      return std::string();
    }
    std::string s(reinterpret_cast<const char *>(p + off), (size_t)len);
    off += (size_t)len;
    return s;
  }
  std::map<std::string, std::string> dict() {
    std::map<std::string, std::string> d;
    const int32_t count = i32();
    if (count < 0 || count > 4096) ok = false;
    for (int32_t k = 0; ok && k < count; ++k) {
      std::string key = str();
      std::string val = str();
      d[key] = val;
    }
    return d;
  }
};

struct Model {
  int32_t size[3] = {0, 0, 0};
  std::vector<uint8_t> cells;  // [z][y][x] color index, 0 = empty
};

struct Node {
  int kind = 0;  // 1 transform, 2 group, 3 shape
  std::vector<int32_t> children;  // transform: 1 child; group: n; shape: model ids
  int32_t t[3] = {0, 0, 0};
  bool rotated = false;
};

// The format's default palette (used when a file has no RGBA chunk): entry
// 0 is empty, entries 1..215 run over the 6 levels ff, cc, 99, 66, 33, 00
// of red (outermost), green, then blue (innermost), leaving out black;
// entries 216..255 are 10-level ramps ee dd bb aa 88 77 55 44 22 11 of red,
// green, blue and grey.
void default_palette(uint8_t pal[256][3]) {
  static const uint8_t six[6] = {0xff, 0xcc, 0x99, 0x66, 0x33, 0x00};
  static const uint8_t ten[10] = {0xee, 0xdd, 0xbb, 0xaa, 0x88, 0x77, 0x55, 0x44, 0x22, 0x11};
  int k = 0;
  pal[k][0] = pal[k][1] = pal[k][2] = 0;
  ++k;
  for (int r = 0; r < 6; ++r)
    for (int g = 0; g < 6; ++g)
      for (int b = 0; b < 6; ++b) {
        if (r == 5 && g == 5 && b == 5) continue;
        pal[k][0] = six[r];
        pal[k][1] = six[g];
        pal[k][2] = six[b];
        ++k;
      }
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < 10; ++i) {
      pal[k][0] = (c == 0 || c == 3) ? ten[i] : 0;
      pal[k][1] = (c == 1 || c == 3) ? ten[i] : 0;
      pal[k][2] = (c == 2 || c == 3) ? ten[i] : 0;
      ++k;
    }
}

// Grid coordinates of a placed voxel, map_loader.cpp:226-262.
inline void axis_remap(uint32_t axis, const int32_t p[3], int32_t g[3]) {
  switch (axis) {
    case CT_AXIS_X_BOTTOM_TOP:
    case CT_AXIS_X_TOP_BOTTOM:
      g[0] = p[1], g[1] = p[2], g[2] = p[0];
      break;
    case CT_AXIS_Y_BOTTOM_TOP:
    case CT_AXIS_Y_TOP_BOTTOM:
      g[0] = p[2], g[1] = p[0], g[2] = p[1];
      break;
    default:  // Z_BOTTOM_TOP, Z_TOP_BOTTOM
      g[0] = p[0], g[1] = p[1], g[2] = p[2];
      break;
  }
}

inline int32_t floor_div(int32_t a, int32_t b) {
  int32_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

struct Placement {
  int32_t model;
  int32_t offset[3];
};

// Largest summed translation accepted; placements stay far inside int32.
constexpr int64_t kMaxTranslation = int64_t(1) << 24;

// Walks the scene graph from `id` with the summed translation `t`.  Every
// node may be reached once: a graph whose nodes share children (a DAG, or a
// cycle) is refused, which bounds the walk by the node count.
int walk_graph(const std::map<int32_t, Node> &nodes, int32_t id, const int64_t t[3], int depth,
               const std::vector<Model> &models, std::vector<Placement> &out, std::set<int32_t> &seen) {
  if (depth > 64) return fail(CT_E_INVALID, "ct_scene_from_vox: scene graph too deep");
  auto it = nodes.find(id);
  if (it == nodes.end()) return fail(CT_E_INVALID, "ct_scene_from_vox: scene graph names a missing node");
  if (!seen.insert(id).second) return fail(CT_E_INVALID, "ct_scene_from_vox: scene graph reaches a node twice");
  const Node &n = it->second;
  int rc;
  if (n.kind == 1) {
    if (n.rotated) return fail(CT_E_UNSUPPORTED, "ct_scene_from_vox: rotated nTRN nodes are not supported");
    const int64_t tt[3] = {t[0] + n.t[0], t[1] + n.t[1], t[2] + n.t[2]};
    for (int a = 0; a < 3; ++a)
      if (tt[a] > kMaxTranslation || tt[a] < -kMaxTranslation)
        return fail(CT_E_INVALID, "ct_scene_from_vox: scene graph translation out of range");
    for (int32_t c : n.children)
      if ((rc = walk_graph(nodes, c, tt, depth + 1, models, out, seen))) return rc;
  } else if (n.kind == 2) {
    for (int32_t c : n.children)
      if ((rc = walk_graph(nodes, c, t, depth + 1, models, out, seen))) return rc;
  } else {
    for (int32_t m : n.children) {
      if (m < 0 || (size_t)m >= models.size()) return fail(CT_E_INVALID, "ct_scene_from_vox: nSHP names a missing model");
      Placement pl;
      pl.model = m;
      for (int a = 0; a < 3; ++a) pl.offset[a] = (int32_t)(t[a] - models[m].size[a] / 2);
      out.push_back(pl);
    }
  }
  return CT_OK;
}

}  // namespace

extern "C" int ct_scene_from_vox(const void *data, size_t bytes, const ct_vox_desc *desc, ct_scene **out) {
  if (!data || !desc || !out) return fail(CT_E_INVALID, "ct_scene_from_vox: null argument");
  *out = nullptr;
  if (desc->axis > CT_AXIS_Z_TOP_BOTTOM) return fail(CT_E_INVALID, "ct_scene_from_vox: bad axis");
  const float e = desc->voxel_extent > 0.0f ? desc->voxel_extent : CT_VOXEL_EXTENT;
  const float flux = desc->flux > 0.0f ? desc->flux : 1.0f;
  Reader r(static_cast<const uint8_t *>(data), bytes);
  if (bytes < 8 || std::memcmp(data, "VOX ", 4) != 0) return fail(CT_E_INVALID, "ct_scene_from_vox: not a VOX file");
  r.off = 8;  // magic + version
  if (!r.have(12)) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated MAIN chunk");
  if (std::memcmp(r.p + r.off, "MAIN", 4) != 0) return fail(CT_E_INVALID, "ct_scene_from_vox: MAIN chunk missing");
  r.off += 4;
  const int32_t main_content = r.i32(), main_children = r.i32();
  if (main_content < 0 || main_children < 0 || !r.have((size_t)main_content + (size_t)main_children))
    return fail(CT_E_INVALID, "ct_scene_from_vox: truncated MAIN chunk");
  r.off += (size_t)main_content;
  const size_t end = r.off + (size_t)main_children;

  std::vector<Model> models;
  uint8_t pal[256][3];
  default_palette(pal);
  bool emissive[256] = {false};
  std::map<int32_t, Node> nodes;
  int32_t pending_size[3] = {-1, -1, -1};
  while (r.off < end) {
    if (!r.have(12)) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated chunk header");
    char id[5] = {0};
    std::memcpy(id, r.p + r.off, 4);
    r.off += 4;
    const int32_t nc = r.i32(), nk = r.i32();
    if (nc < 0 || nk < 0 || r.off + (size_t)nc + (size_t)nk > end)
      return fail(CT_E_INVALID, std::string("ct_scene_from_vox: chunk ") + id + " overruns MAIN");
    Reader c(r.p, r.off + (size_t)nc);  // this chunk's content only
    c.off = r.off;
    r.off += (size_t)nc + (size_t)nk;
    if (!std::strcmp(id, "SIZE")) {
      for (int a = 0; a < 3; ++a) pending_size[a] = c.i32();
      if (!c.ok) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated SIZE");
      for (int a = 0; a < 3; ++a)
        if (pending_size[a] <= 0 || pending_size[a] > 2048) return fail(CT_E_INVALID, "ct_scene_from_vox: bad model size");
      if ((int64_t)pending_size[0] * pending_size[1] * pending_size[2] > (int64_t)1 << 28)
        return fail(CT_E_INVALID, "ct_scene_from_vox: model too large");
    } else if (!std::strcmp(id, "XYZI")) {
      if (pending_size[0] < 0) return fail(CT_E_INVALID, "ct_scene_from_vox: XYZI without SIZE");
      Model m;
      std::memcpy(m.size, pending_size, sizeof(m.size));
      pending_size[0] = -1;
      m.cells.assign((size_t)m.size[0] * m.size[1] * m.size[2], 0);
      const int32_t nv = c.i32();
      if (!c.ok || nv < 0 || !c.have(4 * (size_t)nv)) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated XYZI");
      for (int32_t k = 0; k < nv; ++k) {
        const uint8_t x = c.u8(), y = c.u8(), z = c.u8(), ci = c.u8();
        if (x >= m.size[0] || y >= m.size[1] || z >= m.size[2])
          return fail(CT_E_INVALID, "ct_scene_from_vox: voxel outside its model");
        m.cells[((size_t)z * m.size[1] + y) * m.size[0] + x] = ci;  // a later entry wins
      }
      models.push_back(std::move(m));
    } else if (!std::strcmp(id, "RGBA")) {
      if (!c.have(1024)) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated RGBA");
      pal[0][0] = pal[0][1] = pal[0][2] = 0;
      for (int i = 0; i < 255; ++i)
        for (int k = 0; k < 3; ++k) pal[i + 1][k] = c.p[c.off + 4 * i + k];
    } else if (!std::strcmp(id, "MATL")) {
      const int32_t mid = c.i32();
      auto d = c.dict();
      if (!c.ok) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated MATL");
      auto ty = d.find("_type"), em = d.find("_emit");
      if (mid > 0 && mid < 256 && ty != d.end() && ty->second == "_emit" && em != d.end())
        emissive[mid] = std::strtod(em->second.c_str(), nullptr) > 0.0;
    } else if (!std::strcmp(id, "MATT")) {
      const int32_t mid = c.i32(), type = c.i32();
      float weight = 0.0f;
      const int32_t wbits = c.i32();
      std::memcpy(&weight, &wbits, 4);
      if (!c.ok) return fail(CT_E_INVALID, "ct_scene_from_vox: truncated MATT");
      if (mid > 0 && mid < 256) emissive[mid] = (type == 3 && weight > 0.0f);
    } else if (!std::strcmp(id, "nTRN") || !std::strcmp(id, "nGRP") || !std::strcmp(id, "nSHP")) {
      Node n;
      const int32_t nid = c.i32();
      (void)c.dict();  // node attributes (name, hidden)
      if (id[1] == 'T') {
        n.kind = 1;
        n.children.push_back(c.i32());
        (void)c.i32();  // reserved
        (void)c.i32();  // layer
        const int32_t frames = c.i32();
        if (!c.ok || frames < 0 || frames > 65536) return fail(CT_E_INVALID, "ct_scene_from_vox: bad nTRN");
        for (int32_t f = 0; f < frames; ++f) {
          auto d = c.dict();
          if (f != 0) continue;
          auto t = d.find("_t");
          if (t != d.end()) {
            long v[3] = {0, 0, 0};
            const char *s = t->second.c_str();
            char *endp = nullptr;
            for (int a = 0; a < 3; ++a) {
              v[a] = std::strtol(s, &endp, 10);
              if (endp == s) return fail(CT_E_INVALID, "ct_scene_from_vox: bad nTRN translation");
              s = endp;
            }
            for (int a = 0; a < 3; ++a) {
              if (v[a] > kMaxTranslation || v[a] < -kMaxTranslation)
                return fail(CT_E_INVALID, "ct_scene_from_vox: nTRN translation out of range");
              n.t[a] = (int32_t)v[a];
            }
          }
          auto rr = d.find("_r");
          if (rr != d.end() && std::strtol(rr->second.c_str(), nullptr, 10) != 4) n.rotated = true;  // 4 = identity
        }
      } else if (id[1] == 'G') {
        n.kind = 2;
        const int32_t k = c.i32();
        if (!c.ok || k < 0 || !c.have(4 * (size_t)k)) return fail(CT_E_INVALID, "ct_scene_from_vox: bad nGRP");
        for (int32_t i = 0; i < k; ++i) n.children.push_back(c.i32());
      } else {
        n.kind = 3;
        const int32_t k = c.i32();
        if (!c.ok || k < 0 || k > 65536) return fail(CT_E_INVALID, "ct_scene_from_vox: bad nSHP");
        for (int32_t i = 0; i < k; ++i) {
          n.children.push_back(c.i32());
          (void)c.dict();
        }
      }
      if (!c.ok) return fail(CT_E_INVALID, std::string("ct_scene_from_vox: truncated ") + id);
      nodes[nid] = n;
    }
    // other chunks (PACK, LAYR, IMAP, rOBJ, rCAM, NOTE, ...) carry nothing the loader uses
  }
  if (models.empty()) return fail(CT_E_INVALID, "ct_scene_from_vox: no models");

  std::vector<Placement> places;
  if (desc->model >= 0) {
    if ((size_t)desc->model >= models.size()) return fail(CT_E_INVALID, "ct_scene_from_vox: model index out of range");
    places.push_back(Placement{desc->model, {0, 0, 0}});
  } else if (!nodes.empty()) {
    const int64_t zero[3] = {0, 0, 0};
    std::set<int32_t> seen;
    int rc = walk_graph(nodes, 0, zero, 0, models, places, seen);
    if (rc) return rc;
  } else {
    for (size_t m = 0; m < models.size(); ++m) places.push_back(Placement{(int32_t)m, {0, 0, 0}});
  }

  std::unique_ptr<ct_scene> sc(new ct_scene());
  std::vector<int32_t> mat_of(256);
  for (const Placement &pl : places) {
    const Model &m = models[pl.model];
    int32_t lo[3], hi[3], csz[3], c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = pl.offset[a];
      hi[a] = pl.offset[a] + m.size[a];  // exclusive
      csz[a] = desc->chunk[a] ? (int32_t)std::min<uint32_t>(desc->chunk[a], 1u << 20) : m.size[a];
      c0[a] = desc->chunk[a] ? floor_div(lo[a], csz[a]) : 0;
      c1[a] = desc->chunk[a] ? floor_div(hi[a] - 1, csz[a]) : 0;
    }
    for (int32_t bz = c0[2]; bz <= c1[2]; ++bz)
      for (int32_t by = c0[1]; by <= c1[1]; ++by)
        for (int32_t bx = c0[0]; bx <= c1[0]; ++bx) {
          const int32_t b[3] = {bx, by, bz};
          int32_t rlo[3], rhi[3];
          for (int a = 0; a < 3; ++a) {
            rlo[a] = desc->chunk[a] ? std::max(lo[a], b[a] * csz[a]) : lo[a];
            rhi[a] = desc->chunk[a] ? std::min(hi[a], (b[a] + 1) * csz[a]) : hi[a];
          }
          const uint32_t inst = (uint32_t)sc->instances.size();
          const uint32_t first = (uint32_t)sc->aabbs.size();
          uint32_t count = 0;
          std::fill(mat_of.begin(), mat_of.end(), -1);
          for (int32_t z = rlo[2]; z < rhi[2]; ++z)
            for (int32_t y = rlo[1]; y < rhi[1]; ++y)
              for (int32_t x = rlo[0]; x < rhi[0]; ++x) {
                const uint8_t ci = m.cells[((size_t)(z - lo[2]) * m.size[1] + (y - lo[1])) * m.size[0] + (x - lo[0])];
                if (!ci) continue;
                if (sc->aabbs.size() >= CT_MAX_PRIMITIVES) return fail(CT_E_CAPACITY, "ct_scene_from_vox: too many voxels");
                const float cr = pal[ci][0] / 255.0f, cg = pal[ci][1] / 255.0f, cb = pal[ci][2] / 255.0f;
                const bool light = emissive[ci] && (pal[ci][0] | pal[ci][1] | pal[ci][2]);
                const float er = light ? cr * flux : 0.0f, eg = light ? cg * flux : 0.0f, eb = light ? cb * flux : 0.0f;
                if (mat_of[ci] < 0) {
                  mat_of[ci] = (int32_t)sc->mats.size();
                  ct_material mt{};
                  mt.type = CT_MATERIAL_LAMBERTIAN;
                  mt.diffuse[0] = cr, mt.diffuse[1] = cg, mt.diffuse[2] = cb;
                  mt.emission[0] = er, mt.emission[1] = eg, mt.emission[2] = eb;
                  mt.shininess = 1.0f, mt.roughness = 1.0f, mt.ior = 1.0f, mt.dissolve = 1.0f, mt.illum = 3;
                  sc->mats.push_back(mt);
                }
                const int32_t p[3] = {x, y, z};
                int32_t g[3];
                axis_remap(desc->axis, p, g);
                ct_aabb bb;
                for (int a = 0; a < 3; ++a) {
                  bb.minimum[a] = (float)g[a] * e;
                  bb.maximum[a] = (float)(g[a] + 1) * e;
                }
                sc->aabbs.push_back(bb);
                ct_primitive pr;
                pr.material_index = (uint32_t)mat_of[ci];
                pr.light_index = CT_INVALID_INDEX;
                if (light) {
                  ct_light L{};
                  for (int a = 0; a < 3; ++a) L.position[a] = (float)g[a] * e + e * 0.5f;
                  L.emissive[0] = er, L.emissive[1] = eg, L.emissive[2] = eb;
                  L.instance_info.instance_id = inst;
                  L.instance_info.primitive_id = count;
                  L.size = e;
                  L.type = CT_LIGHT_CUBE;
                  pr.light_index = (uint32_t)sc->lights.size();
                  sc->lights.push_back(L);
                }
                sc->prims.push_back(pr);
                ++count;
              }
          if (!count) continue;
          ct_instance in{};
          for (int k = 0; k < 16; ++k) in.transform[k] = (k % 5 == 0) ? 1.0f : 0.0f;
          in.first_primitive_index = first;
          in.primitive_count = count;
          sc->instances.push_back(in);
        }
  }
  if (sc->aabbs.empty()) return fail(CT_E_INVALID, "ct_scene_from_vox: the file holds no voxels");
  for (int a = 0; a < 3; ++a) {
    sc->gmin[a] = sc->aabbs[0].minimum[a];
    sc->gmax[a] = sc->aabbs[0].maximum[a];
  }
  for (const ct_aabb &b : sc->aabbs)
    for (int a = 0; a < 3; ++a) {
      sc->gmin[a] = std::min(sc->gmin[a], b.minimum[a]);
      sc->gmax[a] = std::max(sc->gmax[a], b.maximum[a]);
    }
  *out = sc.release();
  return CT_OK;
}
