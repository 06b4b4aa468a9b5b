// ct_internal.h -- host-side internals shared by the library's translation
// units: error reporting, HIP checks, the device-side data layouts and the
// accel/ctx objects behind the opaque handles of include/ct_abi.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <deque>
#include <queue>
#include <string>
#include <vector>

#include "../../include/ct_abi.h"

namespace ct {

// ------------------------------------------------------------ errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define CT_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::ct::fail(CT_E_HIP, std::string(#call) + ": " +                \
                                      hipGetErrorString(e_));                 \
  } while (0)

// ------------------------------------------------------ device layouts
constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kHalfVoxel = 0.015625f;      // HALF_VOXEL_EXTENT
constexpr float kLeafPad = 0.000244140625f;  // 2^-12 node padding
// Traversal stack per lane: the top CT_STACK_REG entries live in registers
// (a shift register), deeper ones spill to CT_STACK_LDS LDS slots.
#ifndef CT_STACK_REG
#define CT_STACK_REG 4
#endif
#ifndef CT_STACK_LDS
#define CT_STACK_LDS 28
#endif
constexpr int kStackReg = CT_STACK_REG;
constexpr int kStackLds = CT_STACK_LDS;
constexpr int kStackSize = kStackReg + kStackLds;  // capacity the build's depth must fit
constexpr uint32_t kRaySlots = 16;           // shadow-ray slots per pixel and pass

// BVH2 node, 64 B: child0 box (lo xyz, hi xyz), child1 box, two refs.
// A ref with bit 31 set is a leaf (sorted position), else an internal node.
struct alignas(16) DNode {
  float4 a;  // c0.lo.xyz, c0.hi.x
  float4 b;  // c0.hi.yz, c1.lo.xy
  float4 c;  // c1.lo.z, c1.hi.xyz
  uint4 d;   // ref0, ref1, -, -
};
static_assert(sizeof(DNode) == 64, "node layout");

// Per-instance record the kernels read.
struct alignas(16) DInst {
  float w2o[12];    // world-to-object, column c row r at [c*3 + r] (rows 0..2)
  float model[12];  // rchit.glsl:29-33 model matrix (same packing)
  uint32_t first;   // first_primitive_index: base of its nodes and leaves
  uint32_t count;   // primitive_count
  uint32_t root;    // BLAS root ref (relative to `first`)
  uint32_t pad;
};

struct DScene {
  const DNode *blas_nodes;
  const float4 *blas_leaves;  // (centre.xyz, prim id bits), Morton order
  const DInst *inst;
  const DNode *tlas_nodes;
  const uint32_t *tlas_inst;  // instance id per TLAS leaf
  uint32_t tlas_root;
  uint32_t n_inst;            // entries of `inst` (instance slots)
};

// Raw device buffer with RAII.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grow to at least `n` bytes (contents are not preserved).
  int reserve(size_t n) {
    if (n <= bytes && p) return CT_OK;
    release();
    size_t m = n ? n : 16;
    hipError_t e = hipMalloc(&p, m);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(CT_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = m;
    return CT_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Pinned host memory (hipHostMalloc) with RAII: device-to-host copies into
// it run at the link's rate, where pageable memory is staged by the runtime.
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grow to at least `n` bytes (contents are not preserved).
  int reserve(size_t n) {
    if (n <= bytes && p) return CT_OK;
    release();
    size_t m = n ? n : 16;
    hipError_t e = hipHostMalloc(&p, m, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(CT_E_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    bytes = m;
    return CT_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Build scratch for the LBVH kernels (reused between builds).
struct BuildScratch {
  DevBuf keys[2], vals[2], centers, lo, hi, parent_int, parent_leaf, flags, heights, hist, seg,
      root_box, depth, sorted_keys32;
};

// ---------------------------------------------------------- accel
struct BvhSet {
  DevBuf nodes, leaves, instances_d, prims, aabbs, cube_lights, inst_aos;
  DevBuf tlas_nodes, tlas_inst, tlas_keys, tlas_parent;
  DevBuf blas_keys, blas_ids, blas_parent;  // kept for ct_accel_read_bvh
  DevBuf geo_t;                             // per instance: previous transform * world-to-object (16 floats)
  DevBuf prev_first;                        // per instance: first_primitive_index before the last update
  DevBuf remap_prim, remap_light;           // remapping tables (u32 per slot / per cube light)
  uint32_t remap_prim_cap = 0, remap_light_cap = 0;
  uint32_t tlas_root = kLeaf;
  uint32_t tlas_count = 0;
  uint32_t tlas_depth = 0;
  uint32_t blas_depth_max = 0;
  uint32_t generation = 0;
  uint32_t prim_capacity = 0;
  uint32_t cube_light_count = 0;
  uint32_t cube_light_hw = 0;   // entries held: deleted lights leave stale records past the count
};

// The scene as the frames before an update saw it: the traversable state
// (BLAS nodes and leaves, instance records, TLAS) and the WORLD instance,
// primitive and AABB buffers, copied on the device when an update starts.
// While the accel is SETTLING these are get_previous_tlas and the
// get_previous_*_buffer set (accel_struct_mngr.hpp:216-250); the cube
// lights and the remapping tables have one copy, as in the reference.
struct PrevSet {
  DevBuf nodes, leaves, instances_d, tlas_nodes, tlas_inst, inst_aos, aabbs, prims;
  uint32_t tlas_root = kLeaf, tlas_count = 0, tlas_depth = 0, generation = 0, n_inst = 0, prim_capacity = 0;
  bool valid = false;
};

struct FreeList {  // first-fit over [0, capacity) (containers/free_list.hpp:103-278)
  struct Range { uint32_t off, size; };
  std::vector<Range> free_ranges;
  void init(uint32_t capacity) { free_ranges.assign(1, Range{0, capacity}); }
  bool allocate(uint32_t size, uint32_t &off);
  void deallocate(uint32_t off, uint32_t size);
};

// A task completed by update_scene, kept for UNDO_OP_CPU.  For a primitive
// deletion it holds what the undo needs: the instance's last index before
// the deletion, the deleted light and the light moved into its place, and
// the deleted records.
struct DoneTask {
  ct_task task{};
  uint32_t last = 0, light = 0xFFFFFFFFu, light_moved = 0xFFFFFFFFu;
  ct_aabb aabb{};
  ct_primitive prim{};
  ct_light light_rec{};
  // The light list a primitive deletion left: its length and the record the
  // tail light became at `light`.  An undo applies only to that list.
  uint32_t light_count_after = 0;
  ct_light moved_rec{};
};

struct Accel {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t max_instances = 0, max_primitives = 0, max_cube_lights = 0;
  // host staging (per batch, reset after processing)
  std::vector<ct_instance> temp_instances;
  std::vector<ct_primitive> primitives_host;
  std::vector<ct_aabb> aabb_host;
  std::vector<ct_light> cube_lights_host;
  std::vector<uint32_t> prim_index_host;    // UPDATE_BLAS primitive indices
  uint32_t temp_instance_count = 0, temp_primitive_count = 0, temp_prim_index_count = 0;
  uint32_t cube_light_count = 0;
  uint32_t light_high_water = 0;            // most cube lights ever held
  std::queue<ct_task> tasks;
  // committed host mirror of the device scene
  std::vector<ct_instance> instances;       // by instance id
  std::vector<uint8_t> instance_used;       // free_uuid_list state
  std::vector<float> w2o;                   // [id][16]
  std::vector<float> prev_transform;        // [id][16]: transforms before the last update
  std::vector<uint32_t> prev_first;         // [id]: first_primitive_index before the last update
  std::deque<DoneTask> done;                // completed tasks, most recent last
  std::vector<std::pair<uint32_t, uint32_t>> settle_prims;  // remap slot ranges to clear
  std::vector<uint32_t> settle_lights;                      // remap light entries to clear
  std::vector<ct_aabb> aabbs;               // by global primitive index
  std::vector<ct_primitive> prims;
  std::vector<float> blas_root_box;         // [id][6]
  std::vector<uint32_t> blas_depth;
  FreeList prim_alloc;
  uint32_t prim_high_water = 0;
  BvhSet set;
  PrevSet prev;
  // voxel-deletion marks (squared_brush.slang): instance bits, slot bits, counters
  DevBuf brush_imask, brush_pmask, brush_counters;
  uint32_t brush_imask_words = 0, brush_pmask_words = 0;
  // The library's own compaction of an instance (k_rearrange): its
  // DELETE_PRIMITIVE_BLAS_FROM_GPU task's serial, the instance's slot range,
  // the deleted local indices (ascending) and the survivors at or above the
  // kept count (ascending) that took the deleted slots below it, pairwise.
  // update_scene applies the moves to the host mirror instead of reading the
  // range back.  task_serial counts every task ever added.
  struct Compaction {
    uint64_t serial;
    uint32_t id, first, count;
    std::vector<uint32_t> del, sur;
  };
  std::vector<Compaction> compactions;
  uint64_t task_serial = 0;
  // materials of ct_path_trace, packed, kept between calls
  DevBuf pt_mats;
  uint32_t pt_mat_count = 0;
  BuildScratch scratch;
  PinnedBuf readback;                       // GPU deletions' slots on their way to the host mirror
  DevBuf upload;                            // packed slot records and remapping writes on their way in
  uint32_t status = CT_AS_IDLE;
  uint32_t generation = 0;
};

// Whether the task queue holds primitive deletions (CPU or GPU).  Their
// light moves are computed on the light list as update_scene finds it, so
// nothing may append lights before they run.
inline bool deletions_pending(const Accel &a) {
  std::queue<ct_task> q = a.tasks;
  for (; !q.empty(); q.pop())
    if (q.front().type == CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU || q.front().type == CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU)
      return true;
  return false;
}



// ------------------------------------------------------------- ctx
struct FrameBufs {
  DevBuf di_pos[2], di_nrm[2], di_org[2], di_hit[2];  // DI, current/previous
  DevBuf resA[2], resB[2], resC[2];                   // intermediate / previous
  DevBuf velocity, img[2], taa, hit, pidx, prev_res_copy;
  DevBuf mask, gather, spatial, temporal;             // per-pixel pass state
  DevBuf ray_o, ray_d, ray_t, queue, counters, counters_host_staging;
  DevBuf pres[2], indirect;  // ReSTIR PT: output / temporal path reservoirs, indirect colour
  DevBuf pstate, ptq[2];     // ReSTIR PT wavefront: path state, segment queues
  DevBuf retrace;            // hybrid shift: retrace records of temporal path reuse (2 per pixel)
  bool pt_ready = false;     // allocated (and zeroed) on the first frame with indirect light
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t max_w = 0, max_h = 0;
  FrameBufs fb;
  DevBuf materials, point_lights, env_lights, pt, inst_pt;
  std::vector<ct_material> mats_host;
  std::vector<ct_light> point_host, env_host;
  ct_light_config cfg{};
  int cur = 1;
  uint32_t instrument = 0;
  hipEvent_t ev[48] = {};
  int ev_seg[2 * 16] = {};  // event pairs around each path-segment traversal
  int n_seg = 0;
  int n_events = 0;
  ct_frame_stats stats{};
  uint64_t *counters_pinned = nullptr;
  bool stats_pending = false;
  hipStream_t stream_last = nullptr;
  uint32_t last_w = 0, last_h = 0;
  DevBuf frame_dev, blk_dev;    // the frame record and its block tables on the device
  uint32_t blk_key[4] = {0, 0, 0, 0};
  uint32_t blk_counts[3] = {0, 0, 0};
  const void *lights_accel = nullptr;  // accel and build whose cube lights were uploaded
  uint32_t lights_generation = 0;
  // A frame can run as kSplitParts interleaved parts of the image on as
  // many streams (the render stream and part_stream[1..]). Each part has
  // its own frame record, block table, ray and segment queues, counters and
  // segment buffers; the per-pixel buffers are shared.
  static constexpr int kMaxParts = 4;
  struct PartBufs {
    DevBuf queue, counters, ptq[2], seg_o, seg_d, seg_hit, frame, blk;
  };
  PartBufs part[kMaxParts];
  hipStream_t part_stream[kMaxParts] = {};
  hipEvent_t ev_fork = nullptr, ev_p1[kMaxParts] = {}, ev_p2[kMaxParts] = {};
  uint32_t split_key[6] = {0, 0, 0, 0, 0, 0};
  uint32_t split_counts[kMaxParts][3] = {};
  uint32_t split_px[kMaxParts] = {};
  uint32_t nparts = 0;  // parts of the last frame (0: it ran whole)
  // The last frame's record (its final DI and reservoirs stay in fb until
  // the next frame), for ct_frame_adjoint: valid while the accel keeps the
  // build it was rendered with.
  std::vector<unsigned char> last_frame;
  const void *last_frame_accel = nullptr;
  uint32_t last_frame_generation = 0;
};

// Kernel-side views used by trace.hip and frame.hip.
DScene scene_view(const Accel &a);
// The previous set's traversal view while SETTLING (else the current one).
DScene previous_scene_view(const Accel &a);
bool previous_distinct(const Accel &a);

// Launch the any-hit traversal over a ray queue (frame.hip / trace.hip).
struct TraceAnyArgs {
  const float4 *ray_o;   // (origin, distance)
  const float4 *ray_d;   // (direction, target instance bits)
  const uint32_t *ray_t; // target primitive
  const uint32_t *queue; // ray indices, or null: index = queue position
  const uint32_t *count; // device count (queue length)
  uint32_t *head;        // device work counter, zeroed by the launch
  uint32_t *mask;        // per-pixel occlusion bits (index / 16, bit index % 16)
  uint8_t *occ;          // alternatively one byte per ray (index), may be null
  unsigned long long *stats;  // nodes, leaves, instances (instrumented launch) or null
  uint32_t max_rays;
};
int launch_trace_any(const DScene &s, const TraceAnyArgs &args, hipStream_t stream);

struct TraceClosestArgs {
  const float4 *ray_o;   // (origin, -)
  const float4 *ray_d;   // (direction, -)
  float tmax;
  uint32_t n;
  float4 *hit;           // (t, instance bits, primitive bits, -)
  unsigned long long *stats;
};
int launch_trace_closest(const DScene &s, const TraceClosestArgs &args, hipStream_t stream);

int device_count_gfx950();

}  // namespace ct

struct ct_accel : ct::Accel {};
struct ct_ctx : ct::Ctx {};
// A host-side voxel scene (generated or loaded from a .vox file): the
// arrays the map loader writes into the staging buffers.
struct ct_scene {
  ct_scene_desc desc{};
  std::vector<ct_instance> instances;  // first_primitive_index: host order, contiguous
  std::vector<ct_aabb> aabbs;
  std::vector<ct_primitive> prims;
  std::vector<ct_material> mats;
  std::vector<ct_light> lights;
  float gmin[3] = {0, 0, 0}, gmax[3] = {0, 0, 0};
};
