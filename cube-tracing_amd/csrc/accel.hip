// accel.hip -- the acceleration-structure manager behind ct_accel_* and the
// GPU LBVH build (Morton codes, stable LDS radix sort, Karras topology,
// bottom-up refit) for every BLAS and the TLAS.
//
// Host contract follows ACCEL_STRUCT_MNGR (src/accel_struct_mngr.hpp:18-664,
// src/accel_struct_mngr.cpp): a staging batch is queued with task_queue_add
// and committed by update_scene; instance ids are the lowest free slot
// (containers/uuid.hpp:22-37) and each instance gets a primitive range
// aligned to 32 from a first-fit free list (accel_struct_mngr.cpp:2175-2205).
// The Vulkan driver's opaque BVH is replaced by the LBVH built here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ct_device.h"
#include "ct_internal.h"

namespace ct {

// ------------------------------------------------------------ free list
bool FreeList::allocate(uint32_t size, uint32_t &off) {
  for (size_t i = 0; i < free_ranges.size(); ++i) {
    Range &r = free_ranges[i];
    if (r.size >= size) {
      off = r.off;
      r.off += size;
      r.size -= size;
      if (r.size == 0) free_ranges.erase(free_ranges.begin() + (long)i);
      return true;
    }
  }
  return false;
}
void FreeList::deallocate(uint32_t off, uint32_t size) {
  free_ranges.push_back(Range{off, size});
  std::sort(free_ranges.begin(), free_ranges.end(),
            [](const Range &a, const Range &b) { return a.off < b.off; });
  std::vector<Range> merged;
  for (const Range &r : free_ranges) {
    if (!merged.empty() && merged.back().off + merged.back().size == r.off)
      merged.back().size += r.size;
    else
      merged.push_back(r);
  }
  free_ranges.swap(merged);
}

namespace {

constexpr int kBlock = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;

__device__ __forceinline__ uint32_t fenc(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float fdec(uint32_t u) {
  uint32_t v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

__device__ __forceinline__ uint32_t find_seg(const uint32_t *off, uint32_t nseg, uint32_t k) {
  uint32_t lo = 0, hi = nseg;  // off[lo] <= k < off[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

struct SegView {
  const uint32_t *off;   // [nseg + 1] item offsets
  const uint32_t *base;  // [nseg] node/leaf base (first_primitive_index)
  const uint32_t *src;   // [nseg] source primitive base in the AABB array
  uint32_t nseg;
  uint32_t n;            // total items
};

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per-segment centre bounds: one set of 6 atomics per block when the
// block's items belong to one segment (the common case), else per item.
__device__ void bounds_accumulate(SegView sv, uint32_t k, bool valid, uint32_t s, float cx, float cy, float cz,
                                  uint32_t *bounds) {
  __shared__ uint32_t red[6][kBlock / 64];
  const uint32_t first = blockIdx.x * blockDim.x;
  const uint32_t last = min(first + blockDim.x, sv.n) - 1u;
  const bool uniform = find_seg(sv.off, sv.nseg, first) == find_seg(sv.off, sv.nseg, last);
  if (!uniform) {
    if (valid) {
      uint32_t *bb = bounds + 6 * s;
      atomicMin(bb + 0, fenc(cx)); atomicMin(bb + 1, fenc(cy)); atomicMin(bb + 2, fenc(cz));
      atomicMax(bb + 3, fenc(cx)); atomicMax(bb + 4, fenc(cy)); atomicMax(bb + 5, fenc(cz));
    }
    return;
  }
  uint32_t v[6] = {valid ? fenc(cx) : 0xFFFFFFFFu, valid ? fenc(cy) : 0xFFFFFFFFu, valid ? fenc(cz) : 0xFFFFFFFFu,
                   valid ? fenc(cx) : 0u, valid ? fenc(cy) : 0u, valid ? fenc(cz) : 0u};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      uint32_t o = __shfl_xor(v[a], off);
      v[a] = a < 3 ? min(v[a], o) : max(v[a], o);
    }
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 6; ++a) red[a][w] = v[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sfirst = find_seg(sv.off, sv.nseg, first);
    uint32_t *bb = bounds + 6 * sfirst;
    for (int a = 0; a < 6; ++a) {
      uint32_t r = red[a][0];
      for (int q = 1; q < (int)(blockDim.x / 64); ++q) r = a < 3 ? min(r, red[a][q]) : max(r, red[a][q]);
      if (a < 3) atomicMin(bb + a, r); else atomicMax(bb + a, r);
    }
  }
}

// Item centres of voxel AABBs and per-segment centre bounds.
__global__ void __launch_bounds__(kBlock) k_blas_items(SegView sv, const ct_aabb *aabbs, float4 *centers,
                                                        uint32_t *bounds) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = k < sv.n;
  uint32_t s = 0;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  if (valid) {
    s = find_seg(sv.off, sv.nseg, k);
    uint32_t g = sv.src[s] + (k - sv.off[s]);
    const ct_aabb &b = aabbs[g];
    cx = (b.minimum[0] + b.maximum[0]) * 0.5f;
    cy = (b.minimum[1] + b.maximum[1]) * 0.5f;
    cz = (b.minimum[2] + b.maximum[2]) * 0.5f;
    centers[k] = make_float4(cx, cy, cz, 0.0f);
  }
  bounds_accumulate(sv, k, valid, s, cx, cy, cz, bounds);
}

__global__ void __launch_bounds__(kBlock) k_item_bounds(SegView sv, const float4 *centers, uint32_t *bounds) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = k < sv.n;
  uint32_t s = 0;
  float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (valid) {
    s = find_seg(sv.off, sv.nseg, k);
    c = centers[k];
  }
  bounds_accumulate(sv, k, valid, s, c.x, c.y, c.z, bounds);
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// 30-bit Morton key of the centre quantised to 1024 cells over the
// segment's centre bounds; key64 = segment << 32 | morton, value = local index.
__global__ void k_morton(SegView sv, const float4 *centers, const uint32_t *bounds,
                         unsigned long long *keys, uint32_t *vals) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  const uint32_t *bb = bounds + 6 * s;
  float4 c = centers[k];
  float cc[3] = {c.x, c.y, c.z};
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float lo = fdec(bb[a]), hi = fdec(bb[3 + a]);
    float ext = hi - lo;
    float scale = ext > 0.0f ? 1024.0f / ext : 0.0f;
    float f = (cc[a] - lo) * scale;
    f = fminf(fmaxf(f, 0.0f), 1023.0f);
    q[a] = (uint32_t)f;
  }
  uint32_t m = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
  keys[k] = ((unsigned long long)s << 32) | m;
  vals[k] = k - sv.off[s];
}

// ------------------------------------------------------ LDS radix sort
__global__ void __launch_bounds__(kBlock) k_rs_hist(const unsigned long long *keys, uint32_t n,
                                                     int shift, uint32_t *hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kRsTile;
#pragma unroll 4
  for (int i = 0; i < kRsItems; ++i) {
    uint32_t idx = base + i * kBlock + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of hist[256 * nb] in digit-major order (one block).
__global__ void __launch_bounds__(1024) k_rs_scan(uint32_t *hist, uint32_t total) {
  __shared__ uint32_t part[1024];
  uint32_t t = threadIdx.x;
  uint32_t chunk = (total + 1023) / 1024;
  uint32_t b = t * chunk, e = min(b + chunk, total);
  uint32_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += hist[i];
  part[t] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    uint32_t v = (t >= off) ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (uint32_t i = b; i < e; ++i) {
    uint32_t h = hist[i];
    hist[i] = run;
    run += h;
  }
}

// Stable scatter: rank of an item among equal digits = items of earlier
// rounds, of earlier waves in this round, then of lower lanes (ballots).
__global__ void __launch_bounds__(kBlock) k_rs_scatter(const unsigned long long *kin, const uint32_t *vin,
                                                        unsigned long long *kout, uint32_t *vout,
                                                        uint32_t n, int shift, const uint32_t *hist) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kBlock / 64][256];
  __shared__ uint32_t woff[kBlock / 64][256];
  __shared__ uint32_t gbase[256];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  run[tid] = 0;
  for (int ww = 0; ww < kBlock / 64; ++ww) wcnt[ww][tid] = 0;
  gbase[tid] = hist[tid * gridDim.x + blockIdx.x];
  __syncthreads();
  const unsigned long long lt = (1ull << lane) - 1ull;
  uint32_t base = blockIdx.x * kRsTile;
  for (int i = 0; i < kRsItems; ++i) {
    uint32_t idx = base + i * kBlock + tid;
    bool valid = idx < n;
    unsigned long long key = valid ? kin[idx] : 0ull;
    uint32_t val = valid ? vin[idx] : 0u;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      bool set = (d >> bit) & 1u;
      unsigned long long bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    uint32_t rank = __popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    {
      uint32_t acc = run[tid];
      for (int ww = 0; ww < kBlock / 64; ++ww) {
        woff[ww][tid] = acc;
        acc += wcnt[ww][tid];
        wcnt[ww][tid] = 0;
      }
      run[tid] = acc;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = gbase[d] + woff[w][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
  }
}

// ------------------------------------------------------ Karras topology
__device__ __forceinline__ int delta(const uint32_t *keys, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j > n - 1) return -1;
  uint32_t ki = keys[i], kj = keys[j];
  if (ki == kj) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(ki ^ kj);
}

__global__ void k_split_keys(const unsigned long long *keys, uint32_t *keys32, uint32_t n) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) keys32[k] = (uint32_t)keys[k];
}

__global__ void k_karras(SegView sv, const uint32_t *skeys, DNode *nodes, uint32_t *parent_int,
                         uint32_t *parent_leaf) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  const int64_t n = sv.off[s + 1] - sv.off[s];
  const int64_t i = k - sv.off[s];
  if (i >= n - 1) return;
  const uint32_t *keys = skeys + sv.off[s];
  const uint32_t nb = sv.base[s];
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int64_t sp = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (sp + t) * d) > dnode) sp += t;
  } while (t > 1);
  int64_t gamma = i + sp * d + (d < 0 ? d : 0);
  int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  uint32_t left = (lo == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
  uint32_t right = (hi == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  nodes[nb + i].d = make_uint4(left, right, 0u, 0u);
  if (left & kLeaf) parent_leaf[nb + (left & ~kLeaf)] = (uint32_t)i; else parent_int[nb + left] = (uint32_t)i;
  if (right & kLeaf) parent_leaf[nb + (right & ~kLeaf)] = (uint32_t)i; else parent_int[nb + right] = (uint32_t)i;
}

// Leaf records (Morton order) and the exported sorted key/id arrays.
__global__ void k_leaves(SegView sv, const unsigned long long *skeys, const uint32_t *svals,
                         const float4 *centers, float4 *leaves, uint32_t *out_keys,
                         uint32_t *out_ids, const uint32_t *payload) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  uint32_t j = k - sv.off[s];
  uint32_t p = svals[k];
  float4 c = centers[sv.off[s] + p];
  uint32_t id = payload ? payload[sv.off[s] + p] : p;
  leaves[sv.base[s] + j] = make_float4(c.x, c.y, c.z, __uint_as_float(id));
  out_keys[sv.base[s] + j] = (uint32_t)skeys[k];
  out_ids[sv.base[s] + j] = p;
}

__device__ __forceinline__ float ld_relaxed(const float *p) {
  uint32_t u = __hip_atomic_load(reinterpret_cast<const uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __uint_as_float(u);
}

// Bottom-up refit: the second child to arrive at a node merges both
// child boxes (min/max are exact, so the result is order independent).
__global__ void k_refit(SegView sv, const uint32_t *svals, const float4 *centers, const float4 *lo_in,
                        const float4 *hi_in, float pad_r, DNode *nodes, const uint32_t *parent_int,
                        const uint32_t *parent_leaf, uint32_t *flags, uint32_t *heights, float *root_box,
                        uint32_t *depth) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  uint32_t j = k - sv.off[s];
  uint32_t n = sv.off[s + 1] - sv.off[s];
  uint32_t nb = sv.base[s];
  uint32_t p = svals[k];
  float b[6];
  if (lo_in) {
    float4 l = lo_in[sv.off[s] + p], h = hi_in[sv.off[s] + p];
    b[0] = l.x; b[1] = l.y; b[2] = l.z; b[3] = h.x; b[4] = h.y; b[5] = h.z;
  } else {
    float4 c = centers[sv.off[s] + p];
    b[0] = c.x - pad_r; b[1] = c.y - pad_r; b[2] = c.z - pad_r;
    b[3] = c.x + pad_r; b[4] = c.y + pad_r; b[5] = c.z + pad_r;
  }
  if (n == 1) {
    for (int a = 0; a < 6; ++a) root_box[6 * s + a] = b[a];
    return;
  }
  uint32_t ref = kLeaf | j;
  uint32_t node = parent_leaf[nb + j];
  uint32_t h = 0;
  for (int level = 0; node != kInvalid && node + 1 < n && level < 256; ++level) {
    DNode *nd = nodes + nb + node;
    float *f = reinterpret_cast<float *>(nd);
    uint32_t r0 = nd->d.x;
    int slot = (r0 == ref) ? 0 : 6;
    for (int a = 0; a < 6; ++a) f[slot + a] = b[a];
    __threadfence();
    uint32_t old = atomicAdd(&flags[nb + node], 1u);
    if (old != 1u) return;  // first arrival (0) waits for its sibling; >1 cannot happen in a tree
    __threadfence();
    float c0[6], c1[6];
    for (int a = 0; a < 6; ++a) {
      c0[a] = ld_relaxed(f + a);
      c1[a] = ld_relaxed(f + 6 + a);
    }
    for (int a = 0; a < 3; ++a) {
      b[a] = fminf(c0[a], c1[a]);
      b[3 + a] = fmaxf(c0[3 + a], c1[3 + a]);
    }
    // height = 1 + the taller child (leaves have height 0)
    uint32_t r0c = nd->d.x, r1c = nd->d.y;
    uint32_t h0 = (r0c & kLeaf) ? 0u : __hip_atomic_load(heights + nb + r0c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t h1 = (r1c & kLeaf) ? 0u : __hip_atomic_load(heights + nb + r1c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    h = max(h0, h1) + 1u;
    heights[nb + node] = h;
    ref = node;
    node = parent_int[nb + node];
  }
  for (int a = 0; a < 6; ++a) root_box[6 * s + a] = b[a];
  depth[s] = h;
}

// Structural check of the built topology: every child ref in range, the
// root never a child, and parent links consistent.  Any violation sets
// *err, and the host refuses the build, so traversal only ever sees a
// well-formed tree.
__global__ void k_validate(SegView sv, const DNode *nodes, const uint32_t *parent_int, const uint32_t *parent_leaf,
                           uint32_t *err) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  const uint32_t n = sv.off[s + 1] - sv.off[s];
  const uint32_t i = k - sv.off[s];
  if (i + 1 >= n) return;
  const uint32_t nb = sv.base[s];
  uint4 d = nodes[nb + i].d;
  uint32_t refs[2] = {d.x, d.y};
  for (int c = 0; c < 2; ++c) {
    uint32_t r = refs[c];
    if (r & kLeaf) {
      uint32_t j = r & ~kLeaf;
      if (j >= n || parent_leaf[nb + j] != i) atomicOr(err, 1u);
    } else {
      if (r == 0 || r + 1 >= n || parent_int[nb + r] != i) atomicOr(err, 2u);
    }
  }
}

inline uint32_t blocks(uint64_t n, uint32_t b = kBlock) { return (uint32_t)((n + b - 1) / b); }

// Host copy of a column-major mat4 * (x, y, z, 1), summed as the kernels do.
void host_xform_point(const float *m, const float p[3], float out[3]) {
  for (int r = 0; r < 3; ++r) out[r] = m[0 * 4 + r] * p[0] + m[1 * 4 + r] * p[1] + m[2 * 4 + r] * p[2] + m[3 * 4 + r] * 1.0f;
}

struct BuildOut {
  std::vector<float> root_box;   // [nseg][6]
  std::vector<uint32_t> depth;   // [nseg]
};

// Generic LBVH build over nseg segments of items.  centers/lo/hi are
// device float4 arrays in item order (lo/hi may be null: leaf box =
// centre +- pad_r).  Writes nodes/leaves/keys/ids/parents at each
// segment's base.
int build_segments(Accel &a, hipStream_t st, const std::vector<uint32_t> &off,
                   const std::vector<uint32_t> &base, const std::vector<uint32_t> &src, bool voxel_items,
                   const ct_aabb *d_aabbs, const float4 *d_centers_in, const float4 *d_lo,
                   const float4 *d_hi, const uint32_t *d_payload, float pad_r, uint32_t node_capacity,
                   DNode *d_nodes, float4 *d_leaves, uint32_t *d_keys_out, uint32_t *d_ids_out,
                   uint32_t *d_parent_int, uint32_t *d_parent_leaf, BuildOut &out) {
  BuildScratch &S = a.scratch;
  const uint32_t nseg = (uint32_t)base.size();
  const uint32_t n = off.back();
  int rc;
  if ((rc = S.seg.reserve(sizeof(uint32_t) * (3 * (size_t)nseg + 1)))) return rc;
  uint32_t *d_off = S.seg.as<uint32_t>(), *d_base = d_off + nseg + 1, *d_src = d_base + nseg;
  CT_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(uint32_t) * (nseg + 1), hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(d_base, base.data(), sizeof(uint32_t) * nseg, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(d_src, src.data(), sizeof(uint32_t) * nseg, hipMemcpyHostToDevice, st));
  SegView sv{d_off, d_base, d_src, nseg, n};
  if ((rc = S.keys[0].reserve(8 * (size_t)n)) || (rc = S.keys[1].reserve(8 * (size_t)n)) ||
      (rc = S.vals[0].reserve(4 * (size_t)n)) || (rc = S.vals[1].reserve(4 * (size_t)n)) ||
      (rc = S.centers.reserve(16 * (size_t)n)) || (rc = S.flags.reserve(4 * (size_t)node_capacity)) ||
      (rc = S.heights.reserve(4 * (size_t)node_capacity)) ||
      (rc = S.root_box.reserve(24 * (size_t)nseg)) || (rc = S.depth.reserve(4 * (size_t)nseg)) ||
      (rc = S.sorted_keys32.reserve(4 * (size_t)n)) || (rc = S.hist.reserve(4 * 256 * (size_t)(blocks(n, kRsTile) + 1))))
    return rc;
  uint32_t *d_bounds = S.hist.as<uint32_t>();  // reused before the sort needs it
  std::vector<uint32_t> binit(6 * (size_t)nseg);
  for (uint32_t s = 0; s < nseg; ++s)
    for (int q = 0; q < 6; ++q) binit[6 * s + q] = q < 3 ? 0xFFFFFFFFu : 0u;
  if ((rc = S.lo.reserve(24 * (size_t)nseg))) return rc;
  d_bounds = S.lo.as<uint32_t>();
  CT_HIP(hipMemcpyAsync(d_bounds, binit.data(), 24 * (size_t)nseg, hipMemcpyHostToDevice, st));
  float4 *d_centers = S.centers.as<float4>();
  if (voxel_items) {
    hipLaunchKernelGGL(k_blas_items, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_aabbs, d_centers, d_bounds);
  } else {
    CT_HIP(hipMemcpyAsync(d_centers, d_centers_in, 16 * (size_t)n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_item_bounds, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_centers, d_bounds);
  }
  unsigned long long *k0 = S.keys[0].as<unsigned long long>(), *k1 = S.keys[1].as<unsigned long long>();
  uint32_t *v0 = S.vals[0].as<uint32_t>(), *v1 = S.vals[1].as<uint32_t>();
  hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_centers, d_bounds, k0, v0);
  // LSD passes: 4 over the 30-bit Morton code, then the segment id bytes.
  int passes = 4;
  for (uint32_t m = nseg - 1; m > 0; m >>= 8) ++passes;
  uint32_t nb = blocks(n, kRsTile);
  uint32_t *d_hist = S.hist.as<uint32_t>();
  for (int p = 0; p < passes; ++p) {
    int shift = p < 4 ? 8 * p : 32 + 8 * (p - 4);
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kBlock), 0, st, k0, n, shift, d_hist);
    hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, st, d_hist, 256u * nb);
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kBlock), 0, st, k0, v0, k1, v1, n, shift, d_hist);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  uint32_t *d_k32 = S.sorted_keys32.as<uint32_t>();
  hipLaunchKernelGGL(k_split_keys, dim3(blocks(n)), dim3(kBlock), 0, st, k0, d_k32, n);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(node_capacity)), dim3(kBlock), 0, st, d_parent_int, kInvalid, node_capacity);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(node_capacity)), dim3(kBlock), 0, st, d_parent_leaf, kInvalid, node_capacity);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(node_capacity)), dim3(kBlock), 0, st, S.flags.as<uint32_t>(), 0u, node_capacity);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(nseg)), dim3(kBlock), 0, st, S.depth.as<uint32_t>(), 0u, nseg);
  hipLaunchKernelGGL(k_karras, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_k32, d_nodes, d_parent_int, d_parent_leaf);
  hipLaunchKernelGGL(k_leaves, dim3(blocks(n)), dim3(kBlock), 0, st, sv, k0, v0, d_centers, d_leaves, d_keys_out, d_ids_out, d_payload);
  hipLaunchKernelGGL(k_refit, dim3(blocks(n)), dim3(kBlock), 0, st, sv, v0, d_centers, d_lo, d_hi, pad_r, d_nodes,
                     d_parent_int, d_parent_leaf, S.flags.as<uint32_t>(), S.heights.as<uint32_t>(),
                     S.root_box.as<float>(), S.depth.as<uint32_t>());
  uint32_t *d_err = S.flags.as<uint32_t>();  // flags are free again after the refit
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, st, d_err, 0u, 1u);
  hipLaunchKernelGGL(k_validate, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_nodes, d_parent_int, d_parent_leaf, d_err);
  CT_HIP(hipGetLastError());
  uint32_t herr = 0;
  CT_HIP(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, st));
  out.root_box.resize(6 * (size_t)nseg);
  out.depth.resize(nseg);
  CT_HIP(hipMemcpyAsync(out.root_box.data(), S.root_box.p, 24 * (size_t)nseg, hipMemcpyDeviceToHost, st));
  CT_HIP(hipMemcpyAsync(out.depth.data(), S.depth.p, 4 * (size_t)nseg, hipMemcpyDeviceToHost, st));
  CT_HIP(hipStreamSynchronize(st));
  if (herr) return fail(CT_E_HIP, "LBVH validation failed (code " + std::to_string(herr) + ")");
  return CT_OK;
}

}  // namespace

DScene scene_view(const Accel &a) {
  DScene s;
  s.blas_nodes = a.set.nodes.as<DNode>();
  s.blas_leaves = a.set.leaves.as<float4>();
  s.inst = a.set.instances_d.as<DInst>();
  s.tlas_nodes = a.set.tlas_nodes.as<DNode>();
  s.tlas_inst = a.set.tlas_inst.as<uint32_t>();
  s.tlas_root = a.set.tlas_root;
  s.n_inst = (uint32_t)a.instances.size();
  return s;
}

// Rebuild every BLAS and the TLAS of the committed scene.
int rebuild_all(Accel &a) {
  hipStream_t st = a.stream;
  BvhSet &S = a.set;
  int rc;
  std::vector<uint32_t> ids;
  for (uint32_t i = 0; i < a.max_instances && i < a.instance_used.size(); ++i)
    if (a.instance_used[i]) ids.push_back(i);
  const uint32_t cap = std::max<uint32_t>(a.prim_high_water, 1);
  S.prim_capacity = cap;
  if ((rc = S.aabbs.reserve(sizeof(ct_aabb) * (size_t)cap)) || (rc = S.prims.reserve(sizeof(ct_primitive) * (size_t)cap)) ||
      (rc = S.nodes.reserve(sizeof(DNode) * (size_t)cap)) || (rc = S.leaves.reserve(16 * (size_t)cap)) ||
      (rc = S.blas_keys.reserve(4 * (size_t)cap)) || (rc = S.blas_ids.reserve(4 * (size_t)cap)) ||
      (rc = S.blas_parent.reserve(8 * (size_t)cap)) ||
      (rc = S.inst_aos.reserve(sizeof(ct_instance) * (size_t)std::max<size_t>(a.instances.size(), 1))) ||
      (rc = S.instances_d.reserve(sizeof(DInst) * (size_t)std::max<size_t>(a.instances.size(), 1))) ||
      (rc = S.cube_lights.reserve(sizeof(ct_light) * (size_t)std::max<uint32_t>(a.cube_light_count, 1))))
    return rc;
  CT_HIP(hipMemcpyAsync(S.aabbs.p, a.aabbs.data(), sizeof(ct_aabb) * (size_t)a.prim_high_water, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(S.prims.p, a.prims.data(), sizeof(ct_primitive) * (size_t)a.prim_high_water, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(S.inst_aos.p, a.instances.data(), sizeof(ct_instance) * a.instances.size(), hipMemcpyHostToDevice, st));
  if (a.cube_light_count)
    CT_HIP(hipMemcpyAsync(S.cube_lights.p, a.cube_lights_host.data(), sizeof(ct_light) * a.cube_light_count,
                          hipMemcpyHostToDevice, st));
  S.cube_light_count = a.cube_light_count;
  a.blas_root_box.assign(6 * a.instances.size(), 0.0f);
  a.blas_depth.assign(a.instances.size(), 0);
  std::vector<DInst> dinst(std::max<size_t>(a.instances.size(), 1));
  std::memset(dinst.data(), 0, sizeof(DInst) * dinst.size());
  S.tlas_count = (uint32_t)ids.size();
  if (ids.empty()) {
    S.tlas_root = kInvalid;
    S.tlas_depth = 0;
    S.blas_depth_max = 0;
    S.generation = ++a.generation;
    CT_HIP(hipStreamSynchronize(st));
    return CT_OK;
  }
  // ---- BLASes: one segment per instance.
  std::vector<uint32_t> off(1, 0), base, src;
  for (uint32_t id : ids) {
    base.push_back(a.instances[id].first_primitive_index);
    src.push_back(a.instances[id].first_primitive_index);
    off.push_back(off.back() + a.instances[id].primitive_count);
  }
  BuildOut bo;
  uint32_t *par = S.blas_parent.as<uint32_t>();
  rc = build_segments(a, st, off, base, src, true, S.aabbs.as<ct_aabb>(), nullptr, nullptr, nullptr, nullptr,
                      kHalfVoxel + kLeafPad, cap, S.nodes.as<DNode>(), S.leaves.as<float4>(),
                      S.blas_keys.as<uint32_t>(), S.blas_ids.as<uint32_t>(), par, par + cap, bo);
  if (rc) return rc;
  uint32_t dmax = 0;
  for (size_t q = 0; q < ids.size(); ++q) {
    uint32_t id = ids[q];
    for (int k = 0; k < 6; ++k) a.blas_root_box[6 * id + k] = bo.root_box[6 * q + k];
    a.blas_depth[id] = bo.depth[q];
    dmax = std::max(dmax, bo.depth[q]);
    DInst &d = dinst[id];
    const float *w = &a.w2o[16 * (size_t)id];
    const float *t = a.instances[id].transform;
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 3; ++r) d.w2o[c * 3 + r] = w[c * 4 + r];
    // rchit.glsl:29-33 puts gl_ObjectToWorldEXT[0][1] where [1][0] belongs.
    const float m[12] = {t[0], t[1], t[2], t[1], t[5], t[6], t[8], t[9], t[10], t[12], t[13], t[14]};
    std::memcpy(d.model, m, sizeof(m));
    d.first = a.instances[id].first_primitive_index;
    d.count = a.instances[id].primitive_count;
    d.root = d.count == 1 ? kLeaf : 0u;
  }
  S.blas_depth_max = dmax;
  CT_HIP(hipMemcpyAsync(S.instances_d.p, dinst.data(), sizeof(DInst) * a.instances.size(), hipMemcpyHostToDevice, st));
  // geometry_T = previous_model * inverse(model) per instance
  // (motion_vectors.glsl:146); the scene is static between builds, so
  // previous_model = model.  Frames multiply it by their prev_view_proj.
  {
    std::vector<float> geo(16 * std::max<size_t>(a.instances.size(), 1), 0.0f);
    for (size_t i = 0; i < a.instances.size(); ++i) {
      if (i >= a.instance_used.size() || !a.instance_used[i]) continue;
      const float *t = a.instances[i].transform, *w = &a.w2o[16 * i];
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
          geo[16 * i + c * 4 + r] = t[0 * 4 + r] * w[c * 4 + 0] + t[1 * 4 + r] * w[c * 4 + 1] +
                                    t[2 * 4 + r] * w[c * 4 + 2] + t[3 * 4 + r] * w[c * 4 + 3];
    }
    if ((rc = S.geo_t.reserve(sizeof(float) * geo.size()))) return rc;
    CT_HIP(hipMemcpyAsync(S.geo_t.p, geo.data(), sizeof(float) * geo.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipStreamSynchronize(st));
  }
  // ---- TLAS over the instances' padded world boxes, in slot order.
  const uint32_t ni = (uint32_t)ids.size();
  std::vector<float4> tc(ni), tlo(ni), thi(ni);
  for (uint32_t q = 0; q < ni; ++q) {
    uint32_t id = ids[q];
    const float *rb = &a.blas_root_box[6 * id];
    float wl[3] = {INFINITY, INFINITY, INFINITY}, wh[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
      float p[3] = {(k & 1) ? rb[3] : rb[0], (k & 2) ? rb[4] : rb[1], (k & 4) ? rb[5] : rb[2]}, w[3];
      host_xform_point(a.instances[id].transform, p, w);
      for (int c = 0; c < 3; ++c) {
        wl[c] = std::fmin(wl[c], w[c]);
        wh[c] = std::fmax(wh[c], w[c]);
      }
    }
    float l[3], h[3], c[3];
    for (int k = 0; k < 3; ++k) {
      l[k] = wl[k] - kLeafPad;
      h[k] = wh[k] + kLeafPad;
      c[k] = (l[k] + h[k]) * 0.5f;
    }
    tc[q] = make_float4(c[0], c[1], c[2], 0.0f);
    tlo[q] = make_float4(l[0], l[1], l[2], 0.0f);
    thi[q] = make_float4(h[0], h[1], h[2], 0.0f);
  }
  DevBuf tbuf;
  if ((rc = tbuf.reserve(16 * 3 * (size_t)ni + 4 * (size_t)ni)) || (rc = S.tlas_nodes.reserve(sizeof(DNode) * ni)) ||
      (rc = S.tlas_inst.reserve(16 * (size_t)ni)) || (rc = S.tlas_keys.reserve(8 * (size_t)ni)) ||
      (rc = S.tlas_parent.reserve(8 * (size_t)ni)))
    return rc;
  float4 *dc = tbuf.as<float4>(), *dl = dc + ni, *dh = dl + ni;
  uint32_t *dpay = reinterpret_cast<uint32_t *>(dh + ni);
  CT_HIP(hipMemcpyAsync(dc, tc.data(), 16 * (size_t)ni, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(dl, tlo.data(), 16 * (size_t)ni, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(dh, thi.data(), 16 * (size_t)ni, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(dpay, ids.data(), 4 * (size_t)ni, hipMemcpyHostToDevice, st));
  // TLAS leaves hold (centre, instance id); tlas_inst is the id column.
  DevBuf tleaves;
  if ((rc = tleaves.reserve(16 * (size_t)ni))) return rc;
  std::vector<uint32_t> toff{0, ni}, tbase{0}, tsrc{0};
  uint32_t *tk = S.tlas_keys.as<uint32_t>();
  uint32_t *tpar = S.tlas_parent.as<uint32_t>();
  BuildOut to;
  rc = build_segments(a, st, toff, tbase, tsrc, false, nullptr, dc, dl, dh, dpay, 0.0f, ni,
                      S.tlas_nodes.as<DNode>(), tleaves.as<float4>(), tk, tk + ni, tpar, tpar + ni, to);
  if (rc) return rc;
  // Extract the instance-id column of the TLAS leaves.
  std::vector<float4> hl(ni);
  CT_HIP(hipMemcpy(hl.data(), tleaves.p, 16 * (size_t)ni, hipMemcpyDeviceToHost));
  std::vector<uint32_t> tinst(ni);
  for (uint32_t q = 0; q < ni; ++q) std::memcpy(&tinst[q], &hl[q].w, 4);
  CT_HIP(hipMemcpy(S.tlas_inst.p, tinst.data(), 4 * (size_t)ni, hipMemcpyHostToDevice));
  S.tlas_root = ni == 1 ? kLeaf : 0u;
  S.tlas_depth = to.depth[0];
  S.generation = ++a.generation;
  if (S.tlas_depth + S.blas_depth_max > (uint32_t)kStackSize)
    return fail(CT_E_UNSUPPORTED, "BVH depth " + std::to_string(S.tlas_depth + S.blas_depth_max) +
                                      " exceeds the traversal stack (" + std::to_string(kStackSize) + ")");
  return CT_OK;
}

int device_count_gfx950() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  int ok = 0;
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, i) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) ++ok;
  }
  return ok;
}

}  // namespace ct

using namespace ct;

namespace {
void host_inverse(const float *m, float *out) {
  // Cofactor inverse in double, rounded to float.
  double a[16], inv[16];
  for (int i = 0; i < 16; ++i) a[i] = m[i];
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  double id = det != 0.0 ? 1.0 / det : 0.0;
  for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] * id);
}
inline uint32_t align32(uint32_t v) { return (v + CT_PRIMITIVE_ALIGNMENT - 1) & ~(CT_PRIMITIVE_ALIGNMENT - 1); }
}  // namespace

namespace ct {
void mat4_inverse_host(const float *m, float *out) { host_inverse(m, out); }
}

extern "C" {

int ct_device_available(void) { return ct::device_count_gfx950() > 0 ? 1 : 0; }

int ct_accel_create(const ct_accel_desc *desc, ct_accel **out) {
  if (!desc || !out) return fail(CT_E_INVALID, "ct_accel_create: null argument");
  *out = nullptr;
  if (desc->max_instance_count == 0 || desc->max_instance_count > CT_MAX_INSTANCES || desc->max_primitive_count == 0 ||
      desc->max_primitive_count > CT_MAX_PRIMITIVES)
    return fail(CT_E_INVALID, "ct_accel_create: capacities out of range");
  if (ct::device_count_gfx950() == 0) return fail(CT_E_NODEVICE, "ct_accel_create: no gfx950 device");
  CT_HIP(hipSetDevice(desc->device));
  ct_accel *a = new ct_accel();
  a->device = desc->device;
  a->max_instances = desc->max_instance_count;
  a->max_primitives = desc->max_primitive_count;
  a->max_cube_lights = desc->max_cube_light_count;
  a->temp_instances.assign(a->max_instances, ct_instance{});
  a->primitives_host.assign(a->max_primitives, ct_primitive{});
  a->aabb_host.assign(a->max_primitives, ct_aabb{});
  a->cube_lights_host.assign(std::max<uint32_t>(a->max_cube_lights, 1), ct_light{});
  a->prim_alloc.init(a->max_primitives);
  hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete a;
    return fail(CT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = a;
  return CT_OK;
}

int ct_accel_destroy(ct_accel *a) {
  if (!a) return fail(CT_E_INVALID, "ct_accel_destroy: null handle");
  (void)hipSetDevice(a->device);
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
  return CT_OK;
}

int ct_accel_get_staging(ct_accel *a, ct_staging *o) {
  if (!a || !o) return fail(CT_E_INVALID, "ct_accel_get_staging: null argument");
  o->instances = a->temp_instances.data();
  o->primitives = a->primitives_host.data();
  o->aabbs = a->aabb_host.data();
  o->cube_lights = a->cube_lights_host.data();
  o->host_instance_count = a->temp_instance_count;
  o->host_primitive_count = a->temp_primitive_count;
  o->cube_light_count = a->cube_light_count;
  o->max_instance_count = a->max_instances;
  o->max_primitive_count = a->max_primitives;
  o->max_cube_light_count = a->max_cube_lights;
  return CT_OK;
}

int ct_accel_set_cube_light_count(ct_accel *a, uint32_t count) {
  if (!a) return fail(CT_E_INVALID, "ct_accel_set_cube_light_count: null handle");
  if (count > a->max_cube_lights) return fail(CT_E_CAPACITY, "cube light count exceeds max_cube_light_count");
  a->cube_light_count = count;
  return CT_OK;
}

int ct_accel_request_aabb_host_buffer_count(ct_accel *a, uint32_t count, ct_aabb **base, uint32_t *offset) {
  if (!a || !base || !offset) return fail(CT_E_INVALID, "request_aabb_host_buffer_count: null argument");
  if ((uint64_t)a->temp_primitive_count + count > a->max_primitives)
    return fail(CT_E_CAPACITY, "request_aabb_host_buffer_count: exceeds max_primitive_count");
  *offset = a->temp_primitive_count;
  *base = a->aabb_host.data();
  a->temp_primitive_count += count;
  return CT_OK;
}

int ct_accel_task_queue_add(ct_accel *a, const ct_task *task) {
  if (!a || !task) return fail(CT_E_INVALID, "task_queue_add: null argument");
  if (task->type > CT_TASK_UNDO_OP_CPU) return fail(CT_E_INVALID, "task_queue_add: unknown task type");
  if (task->type == CT_TASK_BUILD_BLAS_FROM_CPU) {
    if ((uint64_t)a->temp_instance_count + task->u.blas_build_from_cpu.instance_count > a->max_instances)
      return fail(CT_E_CAPACITY, "task_queue_add: instance staging full");
    a->temp_instance_count += task->u.blas_build_from_cpu.instance_count;
    a->temp_primitive_count += task->u.blas_build_from_cpu.primitive_count;
  }
  a->tasks.push(*task);
  return CT_OK;
}

int ct_accel_update_scene(ct_accel *a, int synchronize, void *stream, int *did_work) {
  if (!a) return fail(CT_E_INVALID, "update_scene: null handle");
  if (did_work) *did_work = 0;
  if (a->tasks.empty()) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  if (stream) CT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  // process_task_queue, accel_struct_mngr.cpp:2115-2443
  uint32_t cursor = 0;
  while (!a->tasks.empty()) {
    ct_task t = a->tasks.front();
    a->tasks.pop();
    switch (t.type) {
      case CT_TASK_BUILD_BLAS_FROM_CPU: {
        for (uint32_t q = 0; q < t.u.blas_build_from_cpu.instance_count; ++q, ++cursor) {
          if (cursor >= a->max_instances) return fail(CT_E_CAPACITY, "update_scene: staging cursor out of range");
          const ct_instance &ti = a->temp_instances[cursor];
          uint32_t id = 0;
          while (id < a->instance_used.size() && a->instance_used[id]) ++id;
          if (id >= a->max_instances) return fail(CT_E_CAPACITY, "update_scene: no free instance id");
          if (ti.primitive_count == 0) return fail(CT_E_INVALID, "update_scene: instance without primitives");
          if ((uint64_t)ti.first_primitive_index + ti.primitive_count > a->max_primitives)
            return fail(CT_E_INVALID, "update_scene: staged primitive range out of bounds");
          uint32_t off = 0;
          if (!a->prim_alloc.allocate(align32(ti.primitive_count), off))
            return fail(CT_E_CAPACITY, "update_scene: primitive free list exhausted");
          if (id >= a->instance_used.size()) {
            a->instance_used.resize(id + 1, 0);
            a->instances.resize(id + 1, ct_instance{});
            a->w2o.resize(16 * (size_t)(id + 1), 0.0f);
          }
          a->instance_used[id] = 1;
          uint32_t end = off + align32(ti.primitive_count);
          if (end > a->prim_high_water) {
            a->prim_high_water = end;
            a->aabbs.resize(end);
            a->prims.resize(end);
          }
          std::memcpy(&a->aabbs[off], &a->aabb_host[ti.first_primitive_index], sizeof(ct_aabb) * ti.primitive_count);
          std::memcpy(&a->prims[off], &a->primitives_host[ti.first_primitive_index], sizeof(ct_primitive) * ti.primitive_count);
          ct_instance &in = a->instances[id];
          std::memcpy(in.transform, t.u.blas_build_from_cpu.transform, sizeof(in.transform));
          in.first_primitive_index = off;
          in.primitive_count = ti.primitive_count;
          host_inverse(in.transform, &a->w2o[16 * (size_t)id]);
        }
        break;
      }
      case CT_TASK_DELETE_BLAS_FROM_CPU: {
        uint32_t id = t.u.blas_delete_from_cpu.instance_index;
        if (id >= a->instance_used.size() || !a->instance_used[id])
          return fail(CT_E_INVALID, "update_scene: DELETE_BLAS of an unused instance");
        a->prim_alloc.deallocate(a->instances[id].first_primitive_index, align32(a->instances[id].primitive_count));
        a->instance_used[id] = 0;
        a->instances[id] = ct_instance{};
        break;
      }
      case CT_TASK_UPDATE_BLAS_FROM_CPU: {
        uint32_t id = t.u.blas_update.instance_index;
        if (id >= a->instance_used.size() || !a->instance_used[id])
          return fail(CT_E_INVALID, "update_scene: UPDATE_BLAS of an unused instance");
        if (t.u.blas_update.primitive_count != 0)
          return fail(CT_E_UNSUPPORTED, "update_scene: UPDATE_BLAS with primitive changes is not supported");
        std::memcpy(a->instances[id].transform, t.u.blas_update.transform, sizeof(float) * 16);
        host_inverse(a->instances[id].transform, &a->w2o[16 * (size_t)id]);
        break;
      }
      default:
        return fail(CT_E_UNSUPPORTED, "update_scene: task type " + std::to_string(t.type) + " is not supported");
    }
  }
  a->temp_instance_count = 0;
  a->temp_primitive_count = 0;
  a->status = CT_AS_UPDATING;
  int rc = rebuild_all(*a);
  a->status = CT_AS_IDLE;
  if (rc) return rc;
  if (synchronize) CT_HIP(hipStreamSynchronize(a->stream));
  if (did_work) *did_work = 1;
  return CT_OK;
}

int ct_accel_get_status(ct_accel *a, uint32_t *status) {
  if (!a || !status) return fail(CT_E_INVALID, "get_status: null argument");
  *status = a->status;
  return CT_OK;
}

int ct_accel_is_remapping_active(ct_accel *a, int *active) {
  if (!a || !active) return fail(CT_E_INVALID, "is_remapping_active: null argument");
  *active = a->status == CT_AS_SETTLING ? 1 : 0;
  return CT_OK;
}

int ct_accel_get_tlas(ct_accel *a, int previous, ct_tlas *o) {
  if (!a || !o) return fail(CT_E_INVALID, "get_tlas: null argument");
  (void)previous;  // previous == current outside the SETTLING phase
  o->nodes = a->set.tlas_nodes.p;
  o->leaves = a->set.tlas_inst.p;
  o->root_ref = a->set.tlas_root;
  o->instance_count = a->set.tlas_count;
  o->depth = a->set.tlas_depth;
  o->generation = a->set.generation;
  return CT_OK;
}

int ct_accel_get_world(ct_accel *a, int previous, ct_world *o) {
  if (!a || !o) return fail(CT_E_INVALID, "get_world: null argument");
  (void)previous;
  o->instances = a->set.inst_aos.p;
  o->primitives = a->set.prims.p;
  o->aabbs = a->set.aabbs.p;
  o->cube_lights = a->set.cube_lights.p;
  o->instance_count = (uint32_t)a->instances.size();
  o->primitive_capacity = a->set.prim_capacity;
  o->cube_light_count = a->set.cube_light_count;
  o->pad = 0;
  return CT_OK;
}

int ct_accel_get_world_to_object(ct_accel *a, uint32_t id, float out[16]) {
  if (!a || !out) return fail(CT_E_INVALID, "get_world_to_object: null argument");
  if (id >= a->instance_used.size() || !a->instance_used[id]) return fail(CT_E_INVALID, "get_world_to_object: unused instance");
  std::memcpy(out, &a->w2o[16 * (size_t)id], 16 * sizeof(float));
  return CT_OK;
}

int ct_accel_read_instances(ct_accel *a, ct_instance *out, uint32_t capacity, uint32_t *count) {
  if (!a || !count || (capacity && !out)) return fail(CT_E_INVALID, "read_instances: null argument");
  const uint32_t n = (uint32_t)a->instances.size();
  *count = n;
  for (uint32_t i = 0; i < n && i < capacity; ++i) {
    out[i] = a->instances[i];
    if (i >= a->instance_used.size() || !a->instance_used[i]) out[i].primitive_count = 0;
  }
  return CT_OK;
}

int ct_accel_read_bvh(ct_accel *a, uint32_t id, ct_bvh_arrays *arr) {
  if (!a || !arr) return fail(CT_E_INVALID, "read_bvh: null argument");
  CT_HIP(hipSetDevice(a->device));
  CT_HIP(hipStreamSynchronize(a->stream));
  const BvhSet &S = a->set;
  uint32_t n, base, depth, root;
  const DNode *nodes;
  const uint32_t *keys, *ids, *pint, *pleaf;
  if (id == kInvalid) {
    n = S.tlas_count;
    if (n == 0) return fail(CT_E_STATE, "read_bvh: no TLAS");
    base = 0;
    nodes = S.tlas_nodes.as<DNode>();
    keys = S.tlas_keys.as<uint32_t>();
    ids = keys + n;
    pint = S.tlas_parent.as<uint32_t>();
    pleaf = pint + n;
    depth = S.tlas_depth;
    root = S.tlas_root;
  } else {
    if (id >= a->instance_used.size() || !a->instance_used[id]) return fail(CT_E_INVALID, "read_bvh: unused instance");
    n = a->instances[id].primitive_count;
    base = a->instances[id].first_primitive_index;
    nodes = S.nodes.as<DNode>();
    keys = S.blas_keys.as<uint32_t>();
    ids = S.blas_ids.as<uint32_t>();
    pint = S.blas_parent.as<uint32_t>();
    pleaf = pint + S.prim_capacity;
    depth = a->blas_depth[id];
    root = n == 1 ? kLeaf : 0u;
  }
  if (arr->keys) CT_HIP(hipMemcpy(arr->keys, keys + base, 4 * (size_t)n, hipMemcpyDeviceToHost));
  if (arr->ids) CT_HIP(hipMemcpy(arr->ids, ids + base, 4 * (size_t)n, hipMemcpyDeviceToHost));
  if (n > 1 && (arr->child || arr->boxes)) {
    std::vector<DNode> hn(n - 1);
    CT_HIP(hipMemcpy(hn.data(), nodes + base, sizeof(DNode) * (n - 1), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i + 1 < n; ++i) {
      if (arr->child) {
        arr->child[2 * i] = hn[i].d.x;
        arr->child[2 * i + 1] = hn[i].d.y;
      }
      if (arr->boxes) std::memcpy(arr->boxes + 12 * (size_t)i, &hn[i], 48);
    }
  }
  if (arr->parent) {
    if (n > 1) CT_HIP(hipMemcpy(arr->parent, pint + base, 4 * (size_t)(n - 1), hipMemcpyDeviceToHost));
    CT_HIP(hipMemcpy(arr->parent + (n - 1), pleaf + base, 4 * (size_t)n, hipMemcpyDeviceToHost));
  }
  arr->n = n;
  arr->depth = depth;
  arr->root_ref = root;
  return CT_OK;
}

}  // extern "C"
