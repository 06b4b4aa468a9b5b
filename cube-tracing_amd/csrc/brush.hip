// brush.hip -- voxel editing on the GPU: the squared deletion brush
// (src/shaders/slang/compute/squared_brush.slang:37-133), the swap-with-tail
// rearrangement of every marked instance (rearrangement.slang:8-204) and the
// host step that turns the marks into DELETE_PRIMITIVE_BLAS_FROM_GPU tasks
// (check_voxel_modifications / process_voxel_modifications,
// accel_struct_mngr.cpp:2717-2871).
//
// Marks live in the accel: one bit per instance id, one bit per primitive
// slot (first_primitive_index + primitive id), and two counters (newly
// marked instances, newly marked primitives), all set with atomics, so any
// number of brush strokes accumulate until the marks are processed.
//
// The rearrangement keeps the reference's result -- for an instance of n
// primitives with d marked, the k-th marked local index below n - d takes
// the primitive and AABB of the k-th unmarked local index at or above
// n - d, the remapping table gets remap[first + moved] = new local index and
// remap[first + deleted] = ~0u -- but is organised for the GPU as one
// workgroup per marked instance: a block scan over the instance's mask words
// ranks every deleted bit and every tail survivor, and a binary search over
// the survivors' word prefix pairs them.  The reference's per-thread version
// scans every instance's words for each deleted primitive and has three
// faults this code does not copy: it finds the instance with `>=` where the
// k-th deletion needs `>` (rearrangement.slang:142), it scans only
// ceil(marked instances / 32) words of the instance mask, so instances with
// ids beyond that are never compacted (:102-106), and it tests the tail's
// mask bits at local rather than slot indices (:35-46 called from :180).

#include <algorithm>
#include <vector>

#include "ct_internal.h"

using namespace ct;

namespace {

constexpr int kBrushBlock = 256;
constexpr uint32_t kInvalidIdx = 0xFFFFFFFFu;

__global__ void k_brush_fill(uint32_t *p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// delete_primitive_from_instance (squared_brush.slang:37-74): sets the
// instance and slot bits; returns bit 0 when the instance bit is new and
// bit 1 when the slot bit is new.
__device__ uint32_t mark_voxel(uint32_t iid, uint32_t pid, const ct_instance *inst, uint32_t n_inst,
                               uint32_t slot_bits, uint32_t *imask, uint32_t *pmask) {
  if (iid >= n_inst) return 0u;
  const uint32_t slot = inst[iid].first_primitive_index + pid;
  if (pid >= inst[iid].primitive_count || slot >= slot_bits) return 0u;
  const uint32_t ib = 1u << (iid & 31u), pb = 1u << (slot & 31u);
  // Bits are only ever set, so a plain read that already shows the bit is
  // final and saves the atomic: every voxel of an instance shares its
  // instance word, and the atomics on it would otherwise serialise.
  uint32_t old_i = imask[iid >> 5];
  if (!(old_i & ib)) old_i = atomicOr(&imask[iid >> 5], ib);
  uint32_t old_p = pmask[slot >> 5];
  if (!(old_p & pb)) old_p = atomicOr(&pmask[slot >> 5], pb);
  return ((old_i & ib) ? 0u : 1u) | ((old_p & pb) ? 0u : 2u);
}

// BRUSH_COUNTER increments of a whole wave in one atomic per counter
// (increment_brush_counter_*, squared_brush.slang:7-19); every lane of the
// wave calls it.
__device__ void count_marks(uint32_t res, uint32_t *counters) {
  const unsigned long long ni = __ballot((res & 1u) != 0u), np = __ballot((res & 2u) != 0u);
  if ((threadIdx.x & 63u) == 0u) {
    if (ni) atomicAdd(&counters[0], (uint32_t)__popcll(ni));
    if (np) atomicAdd(&counters[1], (uint32_t)__popcll(np));
  }
}

// entry_brush (squared_brush.slang:79-133) over the pixels of the window
// only: [x0, x0 + nx) x [y0, y0 + ny), which the host clipped as the shader
// does.  The DI records are the context's stored (previous) ones.
__global__ void k_brush(const float4 *di_pos, const uint4 *di_hit, uint32_t W, uint32_t px, uint32_t py, uint32_t x0,
                        uint32_t y0, uint32_t nx, uint32_t ny, const ct_instance *inst, uint32_t n_inst,
                        uint32_t slot_bits, uint32_t *imask, uint32_t *pmask, uint32_t *counters) {
  const uint32_t tx = blockIdx.x * 16u + (threadIdx.x & 15u), ty = blockIdx.y * 16u + (threadIdx.x >> 4);
  uint32_t res = 0u;
  const size_t centre = (size_t)py * W + px;
  if (tx < nx && ty < ny && di_pos[centre].w > 0.0f) {
    const size_t q = (size_t)(y0 + ty) * W + (x0 + tx);
    const uint4 h = di_hit[q];
    if (di_pos[q].w > 0.0f && h.x == di_hit[centre].x)  // only the centre pixel's instance
      res = mark_voxel(h.x, h.y, inst, n_inst, slot_bits, imask, pmask);
  }
  count_marks(res, counters);
}

__global__ void k_brush_mark(const uint2 *hits, uint32_t n, const ct_instance *inst, uint32_t n_inst,
                             uint32_t slot_bits, uint32_t *imask, uint32_t *pmask, uint32_t *counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t res = i < n ? mark_voxel(hits[i].x, hits[i].y, inst, n_inst, slot_bits, imask, pmask) : 0u;
  count_marks(res, counters);
}

struct Marked {
  uint32_t first, count, word_base, pad;
};

// Exclusive scan of one u32 per thread over the block; returns the total.
__device__ uint32_t block_scan(uint32_t v, uint32_t &excl, uint32_t *lds) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  uint32_t x = v;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63u) lds[wave] = x;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < kBrushBlock / 64; ++w) {
    const uint32_t s = lds[w];
    if (w < wave) before += s;
    total += s;
  }
  __syncthreads();
  excl = before + x - v;
  return total;
}

__device__ inline uint32_t word_bits(uint32_t w, uint32_t n) {  // valid local bits of word w
  const uint32_t lo = 32u * w;
  return n - lo >= 32u ? 0xFFFFFFFFu : ((1u << (n - lo)) - 1u);
}

__device__ inline uint32_t tail_bits(uint32_t w, uint32_t keep) {  // bits with local index >= keep
  const uint32_t lo = 32u * w;
  if (keep <= lo) return 0xFFFFFFFFu;
  if (keep - lo >= 32u) return 0u;
  return ~((1u << (keep - lo)) - 1u);
}

// rearrangement.slang:62-204, one workgroup per marked instance.
__global__ void __launch_bounds__(kBrushBlock) k_rearrange(const Marked *mk, const uint32_t *pmask, ct_primitive *prims,
                                                           ct_aabb *aabbs, uint32_t *remap, uint32_t *delp,
                                                           uint32_t *surp, uint32_t *deleted) {
  __shared__ uint32_t lds[kBrushBlock / 64];
  const Marked m = mk[blockIdx.x];
  const uint32_t n = m.count, f = m.first, nw = (n + 31u) / 32u, w0 = f >> 5;
  uint32_t *dp = delp + m.word_base, *sp = surp + m.word_base;
  // marked primitives of the instance
  uint32_t d = 0;
  for (uint32_t w = threadIdx.x; w < nw; w += kBrushBlock) d += __popc(pmask[w0 + w] & word_bits(w, n));
  uint32_t ex;
  d = block_scan(d, ex, lds);
  const uint32_t keep = n - d;
  // per-word prefixes of marked bits and of unmarked bits at or above keep
  uint32_t carry_d = 0, carry_s = 0;
  for (uint32_t base = 0; base < nw; base += kBrushBlock) {
    const uint32_t w = base + threadIdx.x;
    uint32_t nd = 0, ns = 0;
    if (w < nw) {
      const uint32_t valid = word_bits(w, n), del = pmask[w0 + w] & valid;
      nd = __popc(del);
      ns = __popc(~del & valid & tail_bits(w, keep));
    }
    uint32_t ed, es;
    const uint32_t td = block_scan(nd, ed, lds);
    const uint32_t ts = block_scan(ns, es, lds);
    if (w < nw) {
      dp[w] = carry_d + ed;
      sp[w] = carry_s + es;
    }
    carry_d += td;
    carry_s += ts;
  }
  __syncthreads();
  // pair the k-th deleted index below keep with the k-th tail survivor
  for (uint32_t w = threadIdx.x; w < nw; w += kBrushBlock) {
    const uint32_t valid = word_bits(w, n);
    uint32_t del = pmask[w0 + w] & valid;
    uint32_t k = dp[w];
    while (del) {
      const uint32_t b = __ffs(del) - 1u;
      del &= del - 1u;
      const uint32_t i = 32u * w + b;
      if (i < keep) {
        // last word ws with sp[ws] <= k
        uint32_t lo = 0, hi = nw;  // invariant: sp[lo] <= k, answer in [lo, hi)
        while (hi - lo > 1u) {
          const uint32_t mid = (lo + hi) >> 1;
          if (sp[mid] <= k) lo = mid;
          else hi = mid;
        }
        const uint32_t wv = word_bits(lo, n);
        uint32_t sur = ~(pmask[w0 + lo] & wv) & wv & tail_bits(lo, keep);
        for (uint32_t r = k - sp[lo]; r; --r) sur &= sur - 1u;
        const uint32_t s = 32u * lo + (__ffs(sur) - 1u);
        prims[f + i] = prims[f + s];
        aabbs[f + i] = aabbs[f + s];
        remap[f + s] = i;
      }
      remap[f + i] = kInvalidIdx;
      ++k;
    }
  }
  if (threadIdx.x == 0) deleted[blockIdx.x] = d;
}

int ensure_brush(Accel &a) {
  if (a.brush_pmask.p) return CT_OK;
  const uint32_t iw = (a.max_instances + 31u) / 32u + 1u;
  const uint32_t pw = (a.max_primitives + 63u) / 32u + 1u;
  int rc;
  if ((rc = a.brush_imask.reserve(4ull * iw)) || (rc = a.brush_pmask.reserve(4ull * pw)) ||
      (rc = a.brush_counters.reserve(16)))
    return rc;
  a.brush_imask_words = iw;
  a.brush_pmask_words = pw;
  hipStream_t st = a.stream;
  hipLaunchKernelGGL(k_brush_fill, dim3((iw + 255) / 256), dim3(256), 0, st, a.brush_imask.as<uint32_t>(), 0u, iw);
  hipLaunchKernelGGL(k_brush_fill, dim3((pw + 255) / 256), dim3(256), 0, st, a.brush_pmask.as<uint32_t>(), 0u, pw);
  hipLaunchKernelGGL(k_brush_fill, dim3(1), dim3(256), 0, st, a.brush_counters.as<uint32_t>(), 0u, 4u);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));
  return CT_OK;
}

}  // namespace

extern "C" {

int ct_brush_apply(ct_ctx *c, ct_accel *a, uint32_t px, uint32_t py, uint32_t window, void *stream) {
  if (!c || !a) return fail(CT_E_INVALID, "ct_brush_apply: null argument");
  const uint32_t W = c->last_w, H = c->last_h;
  if (!W || !H) return fail(CT_E_STATE, "ct_brush_apply: the context has rendered no frame");
  if (px >= W || py >= H) return fail(CT_E_INVALID, "ct_brush_apply: pixel outside the last frame");
  if (!a->set.inst_aos.p || a->instances.empty()) return fail(CT_E_STATE, "ct_brush_apply: the accel holds no scene");
  CT_HIP(hipSetDevice(a->device));
  int rc;
  if ((rc = ensure_brush(*a))) return rc;
  (void)stream;
  CT_HIP(hipDeviceSynchronize());  // the frame's DI and the accel's tables are complete
  hipStream_t st = a->stream;
  // the window as squared_brush.slang:108-109 clips it, in u32 arithmetic
  // (a cursor within `window` of the left or top edge wraps and selects
  // nothing, as in the shader)
  const uint32_t xmin = std::max(px - window, 0u), ymin = std::max(py - window, 0u);
  const uint32_t xmax = std::min(px + window, W - 1u), ymax = std::min(py + window, H - 1u);
  if (xmin > xmax || ymin > ymax) return CT_OK;
  const uint32_t nx = xmax - xmin + 1u, ny = ymax - ymin + 1u;
  hipLaunchKernelGGL(k_brush, dim3((nx + 15) / 16, (ny + 15) / 16), dim3(256), 0, st, c->fb.di_pos[1].as<float4>(),
                     c->fb.di_hit[1].as<uint4>(), W, px, py, xmin, ymin, nx, ny, a->set.inst_aos.as<ct_instance>(),
                     (uint32_t)a->instances.size(), 32u * a->brush_pmask_words, a->brush_imask.as<uint32_t>(),
                     a->brush_pmask.as<uint32_t>(), a->brush_counters.as<uint32_t>());
  CT_HIP(hipGetLastError());
  return CT_OK;
}

int ct_brush_mark(ct_accel *a, const ct_object_info *voxels, uint32_t n, void *stream) {
  if (!a || (n && !voxels)) return fail(CT_E_INVALID, "ct_brush_mark: null argument");
  if (!a->set.inst_aos.p || a->instances.empty()) return fail(CT_E_STATE, "ct_brush_mark: the accel holds no scene");
  if (!n) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  int rc;
  if ((rc = ensure_brush(*a))) return rc;
  (void)stream;
  CT_HIP(hipDeviceSynchronize());
  hipStream_t st = a->stream;
  DevBuf buf;
  if ((rc = buf.reserve(8ull * n))) return rc;
  CT_HIP(hipMemcpyAsync(buf.p, voxels, 8ull * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_brush_mark, dim3((n + 255) / 256), dim3(256), 0, st, buf.as<uint2>(), n,
                     a->set.inst_aos.as<ct_instance>(), (uint32_t)a->instances.size(), 32u * a->brush_pmask_words,
                     a->brush_imask.as<uint32_t>(), a->brush_pmask.as<uint32_t>(), a->brush_counters.as<uint32_t>());
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));  // `buf` is freed on return
  return CT_OK;
}

int ct_brush_get_counters(ct_accel *a, uint32_t *instance_count, uint32_t *primitive_count) {
  if (!a) return fail(CT_E_INVALID, "ct_brush_get_counters: null handle");
  uint32_t cnt[2] = {0, 0};
  if (a->brush_counters.p) {
    CT_HIP(hipSetDevice(a->device));
    CT_HIP(hipDeviceSynchronize());
    CT_HIP(hipMemcpy(cnt, a->brush_counters.p, 8, hipMemcpyDeviceToHost));
  }
  if (instance_count) *instance_count = cnt[0];
  if (primitive_count) *primitive_count = cnt[1];
  return CT_OK;
}

int ct_accel_process_voxel_modifications(ct_accel *a, void *stream, uint32_t *instances, uint32_t *primitives) {
  if (!a) return fail(CT_E_INVALID, "ct_accel_process_voxel_modifications: null handle");
  if (instances) *instances = 0;
  if (primitives) *primitives = 0;
  if (!a->brush_pmask.p || a->status != CT_AS_IDLE) return CT_OK;  // only when idle (:2855)
  CT_HIP(hipSetDevice(a->device));
  (void)stream;
  CT_HIP(hipDeviceSynchronize());  // marks of every stroke and every frame reading the tables are done
  hipStream_t st = a->stream;
  uint32_t cnt[2];
  CT_HIP(hipMemcpy(cnt, a->brush_counters.p, 8, hipMemcpyDeviceToHost));
  if (cnt[0] == 0u) return CT_OK;
  std::vector<uint32_t> im(a->brush_imask_words);
  CT_HIP(hipMemcpy(im.data(), a->brush_imask.p, 4ull * im.size(), hipMemcpyDeviceToHost));
  std::vector<Marked> mk;
  std::vector<uint32_t> ids;
  uint32_t words = 0;
  for (uint32_t w = 0; w < im.size(); ++w)
    for (uint32_t bits = im[w]; bits; bits &= bits - 1u) {
      const uint32_t id = 32u * w + (uint32_t)(__builtin_ctz(bits));
      if (id >= a->instances.size() || !a->instance_used[id]) continue;  // deleted since it was marked
      const ct_instance &in = a->instances[id];
      if (!in.primitive_count) continue;
      if (in.first_primitive_index & 31u)
        return fail(CT_E_STATE, "ct_accel_process_voxel_modifications: unaligned primitive range");
      if ((uint64_t)in.first_primitive_index + in.primitive_count > a->set.prim_capacity ||
          (uint64_t)in.first_primitive_index + in.primitive_count > a->set.remap_prim_cap)
        return fail(CT_E_STATE, "ct_accel_process_voxel_modifications: instance outside the committed buffers");
      mk.push_back(Marked{in.first_primitive_index, in.primitive_count, words, 0u});
      ids.push_back(id);
      words += (in.primitive_count + 31u) / 32u;
    }
  int rc;
  if (!mk.empty()) {
    DevBuf dmk, scratch, del;
    if ((rc = dmk.reserve(sizeof(Marked) * mk.size())) || (rc = scratch.reserve(8ull * words)) ||
        (rc = del.reserve(4ull * mk.size())))
      return rc;
    CT_HIP(hipMemcpyAsync(dmk.p, mk.data(), sizeof(Marked) * mk.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_rearrange, dim3((uint32_t)mk.size()), dim3(kBrushBlock), 0, st, dmk.as<Marked>(),
                       a->brush_pmask.as<uint32_t>(), a->set.prims.as<ct_primitive>(), a->set.aabbs.as<ct_aabb>(),
                       a->set.remap_prim.as<uint32_t>(), scratch.as<uint32_t>(), scratch.as<uint32_t>() + words,
                       del.as<uint32_t>());
    CT_HIP(hipGetLastError());
    std::vector<uint32_t> dc(mk.size());
    CT_HIP(hipMemcpyAsync(dc.data(), del.p, 4ull * dc.size(), hipMemcpyDeviceToHost, st));
    // the compacted instances' marks (one span of words), from which the
    // host mirror repeats the kernel's moves
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for (const Marked &m : mk) {
      lo = std::min(lo, m.first >> 5);
      hi = std::max(hi, (m.first >> 5) + (m.count + 31u) / 32u);
    }
    std::vector<uint32_t> pm(hi - lo);
    CT_HIP(hipMemcpyAsync(pm.data(), a->brush_pmask.as<uint32_t>() + lo, 4ull * pm.size(), hipMemcpyDeviceToHost, st));
    CT_HIP(hipStreamSynchronize(st));
    // k_rearrange pairs the k-th deleted local index below the kept count
    // with the k-th survivor at or above it, both ascending
    std::vector<Accel::Compaction> cp(mk.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < (int64_t)mk.size(); ++k) {
      const uint32_t n = mk[k].count, nw = (n + 31u) / 32u, w0 = (mk[k].first >> 5) - lo;
      Accel::Compaction &c = cp[k];
      for (uint32_t w = 0; w < nw; ++w) {
        const uint32_t valid = (w + 1u) * 32u <= n ? 0xFFFFFFFFu : (1u << (n & 31u)) - 1u;
        for (uint32_t b = pm[w0 + w] & valid; b; b &= b - 1u) c.del.push_back(32u * w + (uint32_t)__builtin_ctz(b));
      }
      const uint32_t keep = n - (uint32_t)c.del.size();
      size_t d = std::lower_bound(c.del.begin(), c.del.end(), keep) - c.del.begin();
      for (uint32_t i = keep; i < n && c.sur.size() < d; ++i)
        if (!((pm[w0 + (i >> 5)] >> (i & 31u)) & 1u)) c.sur.push_back(i);
    }
    // one DELETE_PRIMITIVE_BLAS_FROM_GPU task per instance, in id order (:2786-2838)
    for (size_t k = 0; k < mk.size(); ++k) {
      if (!dc[k]) continue;
      ct_task t{};
      t.type = CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU;
      t.u.blas_del_prim_gpu.instance_index = ids[k];
      t.u.blas_del_prim_gpu.del_prim_count = dc[k];
      const uint64_t serial = a->task_serial;
      if ((rc = ct_accel_task_queue_add(a, &t))) return rc;
      if (cp[k].del.size() == dc[k]) {
        cp[k].serial = serial;
        cp[k].id = ids[k];
        cp[k].first = mk[k].first;
        cp[k].count = mk[k].count;
        a->compactions.push_back(std::move(cp[k]));
      }
      if (instances) ++*instances;
      if (primitives) *primitives += dc[k];
    }
  }
  // restore_bitmask_buffers + zeroed counters (:2717-2733, 2863-2865)
  const uint32_t iw = a->brush_imask_words, pw = a->brush_pmask_words;
  hipLaunchKernelGGL(k_brush_fill, dim3((iw + 255) / 256), dim3(256), 0, st, a->brush_imask.as<uint32_t>(), 0u, iw);
  hipLaunchKernelGGL(k_brush_fill, dim3((pw + 255) / 256), dim3(256), 0, st, a->brush_pmask.as<uint32_t>(), 0u, pw);
  hipLaunchKernelGGL(k_brush_fill, dim3(1), dim3(256), 0, st, a->brush_counters.as<uint32_t>(), 0u, 4u);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));
  return CT_OK;
}

}  // extern "C"
