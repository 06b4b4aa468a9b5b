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
#include <map>

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

// Per-segment centre bounds over a tile of kBoundItems rows of kBlock items:
// one set of 6 atomics per block when the tile lies in one segment (the
// common case); a tile across segments reduces row by row, and only a row
// that itself spans segments falls back to 6 atomics per item. Few, large
// tiles keep the atomics on a segment's six words from serialising.
constexpr int kBoundItems = 16;
constexpr uint32_t kBoundTile = kBlock * kBoundItems;
__device__ void block_bounds_to(uint32_t v[6], uint32_t *bb) {
  __shared__ uint32_t red[6][kBlock / 64];
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
    for (int a = 0; a < 6; ++a) {
      uint32_t r = red[a][0];
      for (int q = 1; q < kBlock / 64; ++q) r = a < 3 ? min(r, red[a][q]) : max(r, red[a][q]);
      if (a < 3) atomicMin(bb + a, r); else atomicMax(bb + a, r);
    }
  }
  __syncthreads();
}
template <class Centre>
__device__ void tile_bounds(SegView sv, Centre centre, uint32_t *bounds) {
  const uint32_t first = blockIdx.x * kBoundTile;
  if (first >= sv.n) return;  // the whole block
  const uint32_t last = min(first + kBoundTile, sv.n) - 1u;
  const uint32_t s0 = find_seg(sv.off, sv.nseg, first);
  const bool uniform = s0 == find_seg(sv.off, sv.nseg, last);
  uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  for (int i = 0; i < kBoundItems; ++i) {
    const uint32_t row = first + (uint32_t)i * kBlock;
    if (row >= sv.n) break;  // block-uniform
    const uint32_t k = row + threadIdx.x;
    const bool valid = k < sv.n;
    const uint32_t rs0 = uniform ? s0 : find_seg(sv.off, sv.nseg, row);
    const bool row_uniform = uniform || rs0 == find_seg(sv.off, sv.nseg, min(row + kBlock, sv.n) - 1u);
    const uint32_t s = row_uniform ? rs0 : (valid ? find_seg(sv.off, sv.nseg, k) : 0u);
    uint32_t e[3] = {0u, 0u, 0u};
    if (valid) {
      const float4 c = centre(k, s);
      e[0] = fenc(c.x);
      e[1] = fenc(c.y);
      e[2] = fenc(c.z);
    }
    if (uniform) {
      if (valid)
        for (int a = 0; a < 3; ++a) {
          v[a] = min(v[a], e[a]);
          v[3 + a] = max(v[3 + a], e[a]);
        }
    } else if (row_uniform) {
      uint32_t r[6] = {valid ? e[0] : 0xFFFFFFFFu, valid ? e[1] : 0xFFFFFFFFu, valid ? e[2] : 0xFFFFFFFFu,
                       valid ? e[0] : 0u, valid ? e[1] : 0u, valid ? e[2] : 0u};
      block_bounds_to(r, bounds + 6 * rs0);
    } else if (valid) {
      uint32_t *bb = bounds + 6 * s;
      for (int a = 0; a < 3; ++a) {
        atomicMin(bb + a, e[a]);
        atomicMax(bb + 3 + a, e[a]);
      }
    }
  }
  if (uniform) block_bounds_to(v, bounds + 6 * s0);
}

// Item centres of voxel AABBs and per-segment centre bounds.
__global__ void __launch_bounds__(kBlock) k_blas_items(SegView sv, const ct_aabb *aabbs, float4 *centers,
                                                        uint32_t *bounds) {
  tile_bounds(
      sv,
      [&](uint32_t k, uint32_t s) {
        const ct_aabb &b = aabbs[sv.src[s] + (k - sv.off[s])];
        const float4 c = make_float4((b.minimum[0] + b.maximum[0]) * 0.5f, (b.minimum[1] + b.maximum[1]) * 0.5f,
                                     (b.minimum[2] + b.maximum[2]) * 0.5f, 0.0f);
        centers[k] = c;
        return c;
      },
      bounds);
}

__global__ void __launch_bounds__(kBlock) k_item_bounds(SegView sv, const float4 *centers, uint32_t *bounds) {
  tile_bounds(sv, [&](uint32_t k, uint32_t) { return centers[k]; }, bounds);
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

// Exclusive scan of a long array in three passes: tile sums, a scan of the
// sums (k_rs_scan, one block), then each tile's own scan plus its offset.
// A tile is kBlock threads x kScanItems consecutive words per thread.
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kBlock * kScanItems;
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(incl, off);
    if ((int)lane >= off) incl += o;
  }
  if (lane == 63u) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0u;
  total = 0u;
  for (uint32_t q = 0; q < kBlock / 64; ++q) {
    if (q < w) before += wsum[q];
    total += wsum[q];
  }
  __syncthreads();
  return before + incl - v;
}
__global__ void __launch_bounds__(kBlock) k_scan_sums(const uint32_t *a, uint32_t n, uint32_t *sums) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t t = 0u;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) t += a[base + i];
  uint32_t total;
  (void)block_exclusive_scan(t, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
__global__ void __launch_bounds__(kBlock) k_scan_tiles(uint32_t *a, uint32_t n, const uint32_t *offsets) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t x[kScanItems];
  uint32_t t = 0u;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    x[i] = base + i < n ? a[base + i] : 0u;
    t += x[i];
  }
  uint32_t total;
  uint32_t run = offsets[blockIdx.x] + block_exclusive_scan(t, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) a[base + i] = run;
    run += x[i];
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
__device__ __forceinline__ void st_relaxed(float *p, float v) {
  __hip_atomic_store(reinterpret_cast<uint32_t *>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until this thread's earlier memory operations are done. Stores made
// as agent-scope atomics are then visible to every XCD, so a flag set after
// it publishes them without writing back and invalidating the whole L2
// (which is what an agent-scope __threadfence costs on a multi-XCD part).
__device__ __forceinline__ void wait_stores_done() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
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
    // Child boxes and heights go through agent-scope atomic stores and
    // loads, so the flag needs no full fence: the writer waits for its own
    // stores before the increment, and the second arrival reads after it.
    for (int a = 0; a < 6; ++a) st_relaxed(f + slot + a, b[a]);
    wait_stores_done();
    uint32_t old = atomicAdd(&flags[nb + node], 1u);
    if (old != 1u) return;  // first arrival (0) waits for its sibling; >1 cannot happen in a tree
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
    __hip_atomic_store(heights + nb + node, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Parent links of the segments' slot ranges, reset before their build (other
// instances keep theirs for later refits).
__global__ void k_reset_parents(SegView sv, uint32_t *parent_int, uint32_t *parent_leaf) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  uint32_t j = sv.base[s] + (k - sv.off[s]);
  parent_int[j] = kInvalid;
  parent_leaf[j] = kInvalid;
}

// Refit inputs of a built BLAS: its leaf order (local ids by Morton rank) in
// item order, and its leaf records rewritten with the voxels' current centres.
__global__ void k_refit_prep(SegView sv, const uint32_t *blas_ids, const float4 *centers, uint32_t *svals,
                             float4 *leaves) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sv.n) return;
  uint32_t s = find_seg(sv.off, sv.nseg, k);
  uint32_t j = k - sv.off[s];
  uint32_t p = blas_ids[sv.base[s] + j];
  svals[k] = p;
  float4 c = centers[sv.off[s] + p];
  leaves[sv.base[s] + j] = make_float4(c.x, c.y, c.z, __uint_as_float(p));
}

// dst[w.x] = w.y for each (index, value) pair; indices are checked on the host.
__global__ void k_scatter_u32(const uint2 *w, uint32_t n, uint32_t *dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[w[i].x] = w[i].y;
}

// Single slots uploaded as packed records: slot[i] takes aabb[i], prim[i].
__global__ void k_scatter_slots(const uint32_t *slot, const ct_aabb *aabb, const ct_primitive *prim, uint32_t n,
                                ct_aabb *dst_aabb, ct_primitive *dst_prim) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dst_aabb[slot[i]] = aabb[i];
    dst_prim[slot[i]] = prim[i];
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
      (rc = S.sorted_keys32.reserve(4 * (size_t)n)) || (rc = S.hist.reserve(4 * (256 * (size_t)(blocks(n, kRsTile) + 1) + blocks(256u * blocks(n, kRsTile), kScanTile) + 1))))
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
    hipLaunchKernelGGL(k_blas_items, dim3(blocks(n, kBoundTile)), dim3(kBlock), 0, st, sv, d_aabbs, d_centers, d_bounds);
  } else {
    CT_HIP(hipMemcpyAsync(d_centers, d_centers_in, 16 * (size_t)n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_item_bounds, dim3(blocks(n, kBoundTile)), dim3(kBlock), 0, st, sv, d_centers, d_bounds);
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
    const uint32_t nh = 256u * nb, nt = blocks(nh, kScanTile);
    hipLaunchKernelGGL(k_scan_sums, dim3(nt), dim3(kBlock), 0, st, d_hist, nh, d_hist + nh);
    hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, st, d_hist + nh, nt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, st, d_hist, nh, d_hist + nh);
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kBlock), 0, st, k0, v0, k1, v1, n, shift, d_hist);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  uint32_t *d_k32 = S.sorted_keys32.as<uint32_t>();
  hipLaunchKernelGGL(k_split_keys, dim3(blocks(n)), dim3(kBlock), 0, st, k0, d_k32, n);
  hipLaunchKernelGGL(k_reset_parents, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_parent_int, d_parent_leaf);
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

DScene previous_scene_view(const Accel &a) {
  if (!previous_distinct(a)) return scene_view(a);
  DScene s;
  s.blas_nodes = a.prev.nodes.as<DNode>();
  s.blas_leaves = a.prev.leaves.as<float4>();
  s.inst = a.prev.instances_d.as<DInst>();
  s.tlas_nodes = a.prev.tlas_nodes.as<DNode>();
  s.tlas_inst = a.prev.tlas_inst.as<uint32_t>();
  s.tlas_root = a.prev.tlas_root;
  s.n_inst = a.prev.n_inst;
  return s;
}

bool previous_distinct(const Accel &a) { return a.status == CT_AS_SETTLING && a.prev.valid; }

// What an update changed: instances whose BLAS is rebuilt from scratch,
// instances refitted in place, slot ranges whose AABBs/primitives are
// uploaded, and whether buffers must grow (everything is then rebuilt).
struct Dirty {
  bool full = false;
  std::vector<uint32_t> rebuild, refit;
  std::vector<std::pair<uint32_t, uint32_t>> ranges;  // (first slot, count)
};

namespace {

std::vector<uint32_t> used_ids(const Accel &a) {
  std::vector<uint32_t> ids;
  for (uint32_t i = 0; i < a.instance_used.size(); ++i)
    if (a.instance_used[i]) ids.push_back(i);
  return ids;
}

// Instance tables (DInst for traversal, the ct_instance AoS of the WORLD
// buffer, geometry_T = previous transform * world-to-object for motion
// vectors, motion_vectors.glsl:72-81, and the previous first indices of the
// remapping lookup) and the cube light list, from the host mirror.
int upload_tables(Accel &a, hipStream_t st) {
  BvhSet &S = a.set;
  int rc;
  a.light_high_water = std::max(a.light_high_water, a.cube_light_count);
  const size_t ni = std::max<size_t>(a.instances.size(), 1);
  if ((rc = S.inst_aos.reserve(sizeof(ct_instance) * ni)) || (rc = S.instances_d.reserve(sizeof(DInst) * ni)) ||
      (rc = S.geo_t.reserve(sizeof(float) * 16 * ni)) || (rc = S.prev_first.reserve(sizeof(uint32_t) * ni)) ||
      (rc = S.cube_lights.reserve(sizeof(ct_light) * (size_t)std::max<uint32_t>(a.light_high_water, 1))))
    return rc;
  std::vector<DInst> dinst(ni);
  std::memset(dinst.data(), 0, sizeof(DInst) * dinst.size());
  std::vector<float> geo(16 * ni, 0.0f);
  std::vector<uint32_t> pf(ni, 0u);
  for (size_t id = 0; id < a.instances.size(); ++id) {
    pf[id] = a.prev_first[id];  // a deleted instance keeps its previous range until the settle
    if (!a.instance_used[id]) continue;
    DInst &d = dinst[id];
    const float *w = &a.w2o[16 * id];
    const float *t = a.instances[id].transform;
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 3; ++r) d.w2o[c * 3 + r] = w[c * 4 + r];
    // rchit.glsl:29-33 puts gl_ObjectToWorldEXT[0][1] where [1][0] belongs.
    const float m[12] = {t[0], t[1], t[2], t[1], t[5], t[6], t[8], t[9], t[10], t[12], t[13], t[14]};
    std::memcpy(d.model, m, sizeof(m));
    d.first = a.instances[id].first_primitive_index;
    d.count = a.instances[id].primitive_count;
    d.root = d.count == 1 ? kLeaf : 0u;
    const float *pt = &a.prev_transform[16 * id];
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r)
        geo[16 * id + c * 4 + r] = pt[0 * 4 + r] * w[c * 4 + 0] + pt[1 * 4 + r] * w[c * 4 + 1] +
                                   pt[2 * 4 + r] * w[c * 4 + 2] + pt[3 * 4 + r] * w[c * 4 + 3];
  }
  std::vector<ct_instance> aos(a.instances);
  for (size_t id = 0; id < aos.size(); ++id)
    if (!a.instance_used[id]) aos[id] = ct_instance{};
  if (!a.instances.empty()) {
    CT_HIP(hipMemcpyAsync(S.instances_d.p, dinst.data(), sizeof(DInst) * a.instances.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipMemcpyAsync(S.inst_aos.p, aos.data(), sizeof(ct_instance) * aos.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipMemcpyAsync(S.geo_t.p, geo.data(), sizeof(float) * 16 * a.instances.size(), hipMemcpyHostToDevice, st));
    CT_HIP(hipMemcpyAsync(S.prev_first.p, pf.data(), sizeof(uint32_t) * a.instances.size(), hipMemcpyHostToDevice, st));
  }
  // Records past the count stay, as the reference's light buffer keeps them.
  if (a.light_high_water)
    CT_HIP(hipMemcpyAsync(S.cube_lights.p, a.cube_lights_host.data(), sizeof(ct_light) * a.light_high_water,
                          hipMemcpyHostToDevice, st));
  S.cube_light_count = a.cube_light_count;
  S.cube_light_hw = a.light_high_water;
  CT_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
  return CT_OK;
}

// Fresh LBVH of each listed BLAS (one segment per instance).
int build_blas_set(Accel &a, hipStream_t st, const std::vector<uint32_t> &ids) {
  if (ids.empty()) return CT_OK;
  BvhSet &S = a.set;
  const uint32_t cap = S.prim_capacity;
  std::vector<uint32_t> off(1, 0), base, src;
  for (uint32_t id : ids) {
    base.push_back(a.instances[id].first_primitive_index);
    src.push_back(a.instances[id].first_primitive_index);
    off.push_back(off.back() + a.instances[id].primitive_count);
  }
  BuildOut bo;
  uint32_t *par = S.blas_parent.as<uint32_t>();
  int rc = build_segments(a, st, off, base, src, true, S.aabbs.as<ct_aabb>(), nullptr, nullptr, nullptr, nullptr,
                          kHalfVoxel + kLeafPad, cap, S.nodes.as<DNode>(), S.leaves.as<float4>(),
                          S.blas_keys.as<uint32_t>(), S.blas_ids.as<uint32_t>(), par, par + cap, bo);
  if (rc) return rc;
  for (size_t q = 0; q < ids.size(); ++q) {
    for (int k = 0; k < 6; ++k) a.blas_root_box[6 * ids[q] + k] = bo.root_box[6 * q + k];
    a.blas_depth[ids[q]] = bo.depth[q];
  }
  return CT_OK;
}

// In-place refit of each listed BLAS (update_blases with src_blas = the
// BLAS itself, accel_struct_mngr.cpp:1672-1852): the leaf order and node
// topology stay; leaf centres are re-read from the AABBs and every node box
// is recomputed bottom up.
int refit_blas_set(Accel &a, hipStream_t st, const std::vector<uint32_t> &ids) {
  if (ids.empty()) return CT_OK;
  BvhSet &S = a.set;
  BuildScratch &B = a.scratch;
  const uint32_t cap = S.prim_capacity;
  const uint32_t nseg = (uint32_t)ids.size();
  std::vector<uint32_t> off(1, 0), base;
  for (uint32_t id : ids) {
    base.push_back(a.instances[id].first_primitive_index);
    off.push_back(off.back() + a.instances[id].primitive_count);
  }
  const uint32_t n = off.back();
  int rc;
  if ((rc = B.seg.reserve(sizeof(uint32_t) * (3 * (size_t)nseg + 1))) || (rc = B.centers.reserve(16 * (size_t)n)) ||
      (rc = B.vals[0].reserve(4 * (size_t)n)) || (rc = B.flags.reserve(4 * (size_t)cap)) ||
      (rc = B.heights.reserve(4 * (size_t)cap)) || (rc = B.root_box.reserve(24 * (size_t)nseg)) ||
      (rc = B.depth.reserve(4 * (size_t)nseg)) || (rc = B.lo.reserve(24 * (size_t)nseg)))
    return rc;
  uint32_t *d_off = B.seg.as<uint32_t>(), *d_base = d_off + nseg + 1, *d_src = d_base + nseg;
  CT_HIP(hipMemcpyAsync(d_off, off.data(), sizeof(uint32_t) * (nseg + 1), hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(d_base, base.data(), sizeof(uint32_t) * nseg, hipMemcpyHostToDevice, st));
  CT_HIP(hipMemcpyAsync(d_src, base.data(), sizeof(uint32_t) * nseg, hipMemcpyHostToDevice, st));
  SegView sv{d_off, d_base, d_src, nseg, n};
  std::vector<uint32_t> binit(6 * (size_t)nseg);
  for (uint32_t s = 0; s < nseg; ++s)
    for (int q = 0; q < 6; ++q) binit[6 * s + q] = q < 3 ? 0xFFFFFFFFu : 0u;
  CT_HIP(hipMemcpyAsync(B.lo.p, binit.data(), 24 * (size_t)nseg, hipMemcpyHostToDevice, st));
  float4 *d_centers = B.centers.as<float4>();
  uint32_t *d_vals = B.vals[0].as<uint32_t>();
  uint32_t *par = S.blas_parent.as<uint32_t>();
  hipLaunchKernelGGL(k_blas_items, dim3(blocks(n, kBoundTile)), dim3(kBlock), 0, st, sv, S.aabbs.as<ct_aabb>(), d_centers,
                     B.lo.as<uint32_t>());
  hipLaunchKernelGGL(k_refit_prep, dim3(blocks(n)), dim3(kBlock), 0, st, sv, S.blas_ids.as<uint32_t>(), d_centers,
                     d_vals, S.leaves.as<float4>());
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(cap)), dim3(kBlock), 0, st, B.flags.as<uint32_t>(), 0u, cap);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(nseg)), dim3(kBlock), 0, st, B.depth.as<uint32_t>(), 0u, nseg);
  hipLaunchKernelGGL(k_refit, dim3(blocks(n)), dim3(kBlock), 0, st, sv, d_vals, d_centers, nullptr, nullptr,
                     kHalfVoxel + kLeafPad, S.nodes.as<DNode>(), par, par + cap, B.flags.as<uint32_t>(),
                     B.heights.as<uint32_t>(), B.root_box.as<float>(), B.depth.as<uint32_t>());
  CT_HIP(hipGetLastError());
  std::vector<float> rb(6 * (size_t)nseg);
  std::vector<uint32_t> dp(nseg);
  CT_HIP(hipMemcpyAsync(rb.data(), B.root_box.p, 24 * (size_t)nseg, hipMemcpyDeviceToHost, st));
  CT_HIP(hipMemcpyAsync(dp.data(), B.depth.p, 4 * (size_t)nseg, hipMemcpyDeviceToHost, st));
  CT_HIP(hipStreamSynchronize(st));
  for (uint32_t q = 0; q < nseg; ++q) {
    for (int k = 0; k < 6; ++k) a.blas_root_box[6 * ids[q] + k] = rb[6 * q + k];
    a.blas_depth[ids[q]] = dp[q];  // unchanged topology: same height
  }
  return CT_OK;
}

// TLAS over the instances' padded world boxes, in slot order (build_tlas,
// accel_struct_mngr.cpp:1854-1973).
int build_tlas(Accel &a, hipStream_t st) {
  BvhSet &S = a.set;
  int rc;
  const std::vector<uint32_t> ids = used_ids(a);
  S.tlas_count = (uint32_t)ids.size();
  uint32_t dmax = 0;
  for (uint32_t id : ids) dmax = std::max(dmax, a.blas_depth[id]);
  S.blas_depth_max = dmax;
  if (ids.empty()) {
    S.tlas_root = kInvalid;
    S.tlas_depth = 0;
    S.generation = ++a.generation;
    return CT_OK;
  }
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
  rc = build_segments(a, st, toff, tbase, tsrc, false, nullptr, dc, dl, dh, dpay, 0.0f, ni, S.tlas_nodes.as<DNode>(),
                      tleaves.as<float4>(), tk, tk + ni, tpar, tpar + ni, to);
  if (rc) return rc;
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

}  // namespace

// Brings the device scene in line with the host mirror after an update:
// everything when buffers grow (or on the first build), else the touched
// slot ranges, the listed rebuilds and refits, the tables and the TLAS.
int commit(Accel &a, const Dirty &d) {
  hipStream_t st = a.stream;
  BvhSet &S = a.set;
  int rc;
  a.blas_root_box.resize(6 * a.instances.size(), 0.0f);
  a.blas_depth.resize(a.instances.size(), 0);
  const uint32_t cap = std::max<uint32_t>(a.prim_high_water, 1);
  const bool full = d.full || S.prim_capacity != cap || !S.nodes.p;
  if (full) {
    S.prim_capacity = cap;
    if ((rc = S.aabbs.reserve(sizeof(ct_aabb) * (size_t)cap)) || (rc = S.prims.reserve(sizeof(ct_primitive) * (size_t)cap)) ||
        (rc = S.nodes.reserve(sizeof(DNode) * (size_t)cap)) || (rc = S.leaves.reserve(16 * (size_t)cap)) ||
        (rc = S.blas_keys.reserve(4 * (size_t)cap)) || (rc = S.blas_ids.reserve(4 * (size_t)cap)) ||
        (rc = S.blas_parent.reserve(8 * (size_t)cap)) || (rc = S.remap_prim.reserve(4 * (size_t)cap)))
      return rc;
    if (a.prim_high_water)
      CT_HIP(hipMemcpyAsync(S.aabbs.p, a.aabbs.data(), sizeof(ct_aabb) * (size_t)a.prim_high_water,
                            hipMemcpyHostToDevice, st));
    if (a.prim_high_water)
      CT_HIP(hipMemcpyAsync(S.prims.p, a.prims.data(), sizeof(ct_primitive) * (size_t)a.prim_high_water,
                            hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(cap)), dim3(kBlock), 0, st, S.remap_prim.as<uint32_t>(), 0u, cap);
    S.remap_prim_cap = cap;
  } else {
    // Touched slot ranges, sorted and merged; runs of a few slots (a
    // relinked light owner, one per moved light) go up together as packed
    // records with one copy and one scatter kernel, longer runs by copies.
    std::vector<std::pair<uint32_t, uint32_t>> runs;
    for (const auto &r : d.ranges)
      if (r.second) runs.push_back(r);
    std::sort(runs.begin(), runs.end());
    size_t m = 0;
    for (size_t k = 1; k < runs.size(); ++k) {
      const uint64_t end = (uint64_t)runs[m].first + runs[m].second;
      if ((uint64_t)runs[k].first <= end)
        runs[m].second = (uint32_t)std::max<uint64_t>(end, (uint64_t)runs[k].first + runs[k].second) - runs[m].first;
      else
        runs[++m] = runs[k];
    }
    if (!runs.empty()) runs.resize(m + 1);
    constexpr uint32_t kPackedRun = 8;
    std::vector<uint32_t> one;
    for (const auto &r : runs) {
      if (r.second <= kPackedRun) {
        for (uint32_t k = 0; k < r.second; ++k) one.push_back(r.first + k);
        continue;
      }
      CT_HIP(hipMemcpyAsync(S.aabbs.as<ct_aabb>() + r.first, &a.aabbs[r.first], sizeof(ct_aabb) * (size_t)r.second,
                            hipMemcpyHostToDevice, st));
      CT_HIP(hipMemcpyAsync(S.prims.as<ct_primitive>() + r.first, &a.prims[r.first],
                            sizeof(ct_primitive) * (size_t)r.second, hipMemcpyHostToDevice, st));
    }
    if (!one.empty()) {
      const size_t n = one.size();
      std::vector<unsigned char> pk((4 + sizeof(ct_aabb) + sizeof(ct_primitive)) * n);
      uint32_t *ps = reinterpret_cast<uint32_t *>(pk.data());
      ct_aabb *pa = reinterpret_cast<ct_aabb *>(ps + n);
      ct_primitive *pp = reinterpret_cast<ct_primitive *>(pa + n);
      for (size_t k = 0; k < n; ++k) {
        ps[k] = one[k];
        pa[k] = a.aabbs[one[k]];
        pp[k] = a.prims[one[k]];
      }
      if ((rc = a.upload.reserve(pk.size()))) return rc;
      CT_HIP(hipMemcpyAsync(a.upload.p, pk.data(), pk.size(), hipMemcpyHostToDevice, st));
      const uint32_t *ds = a.upload.as<uint32_t>();
      const ct_aabb *da = reinterpret_cast<const ct_aabb *>(ds + n);
      const ct_primitive *dp = reinterpret_cast<const ct_primitive *>(da + n);
      hipLaunchKernelGGL(k_scatter_slots, dim3(blocks(n)), dim3(kBlock), 0, st, ds, da, dp, (uint32_t)n,
                         S.aabbs.as<ct_aabb>(), S.prims.as<ct_primitive>());
      CT_HIP(hipGetLastError());
      CT_HIP(hipStreamSynchronize(st));  // `pk` is released on return
    }
  }
  const uint32_t lcap = std::max<uint32_t>(a.max_cube_lights, 1);
  if (S.remap_light_cap != lcap) {
    if ((rc = S.remap_light.reserve(4 * (size_t)lcap))) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(lcap)), dim3(kBlock), 0, st, S.remap_light.as<uint32_t>(), 0u, lcap);
    S.remap_light_cap = lcap;
  }
  if (full) {
    if ((rc = build_blas_set(a, st, used_ids(a)))) return rc;
  } else {
    if ((rc = build_blas_set(a, st, d.rebuild))) return rc;
    if ((rc = refit_blas_set(a, st, d.refit))) return rc;
  }
  if ((rc = upload_tables(a, st))) return rc;
  return build_tlas(a, st);
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
  if (count != a->cube_light_count && ct::deletions_pending(*a))
    return fail(CT_E_STATE, "ct_accel_set_cube_light_count: primitive deletions are queued; run update_scene first");
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

int ct_accel_request_primitive_index_host_buffer_count(ct_accel *a, uint32_t count, uint32_t **base, uint32_t *offset) {
  if (!a || !base || !offset) return fail(CT_E_INVALID, "request_primitive_index_host_buffer_count: null argument");
  if ((uint64_t)a->temp_prim_index_count + count > a->max_primitives)
    return fail(CT_E_CAPACITY, "request_primitive_index_host_buffer_count: exceeds max_primitive_count");
  if (a->prim_index_host.size() < a->max_primitives) a->prim_index_host.resize(a->max_primitives, 0u);
  *offset = a->temp_prim_index_count;
  *base = a->prim_index_host.data();
  a->temp_prim_index_count += count;
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
  ++a->task_serial;
  return CT_OK;
}

}  // extern "C"

namespace ct {
namespace {

constexpr size_t kMaxDone = 1u << 16;  // undo depth kept

bool used(const Accel &a, uint32_t id) { return id < a.instance_used.size() && a.instance_used[id]; }

void mark(std::vector<uint32_t> &v, uint32_t id) {
  if (std::find(v.begin(), v.end(), id) == v.end()) v.push_back(id);
}

// Column-major product m * t with the operation order of
// daxa_f32mat4x4_mult (math.inl:92-117).
void mat4_mul_host(const float *m, const float *t, float *out) {
  float r[16];
  for (int c = 0; c < 4; ++c)
    for (int row = 0; row < 4; ++row)
      r[c * 4 + row] = m[0 * 4 + row] * t[c * 4 + 0] + m[1 * 4 + row] * t[c * 4 + 1] + m[2 * 4 + row] * t[c * 4 + 2] +
                       m[3 * 4 + row] * t[c * 4 + 3];
  std::memcpy(out, r, sizeof(r));
}

// Per-update state: remapping writes and what the commit must redo.
// Cube-light moves are tracked in the index space of the list as the update
// found it, which is the space the previous frame's reservoirs use:
// origin[l] is the original index of the light now at l (kNoOrigin for a
// light an undo brought back), final_pos[o] where original light o ended up
// (kInvalid once removed, kUnmoved if it never moved).  GPU deletions, CPU
// deletions and undos of one update then compose, and each original index
// gets one remapping entry.
// Voxel moves are tracked the same way per instance, in local indices.
constexpr uint32_t kNoOrigin = 0xFFFFFFFEu, kUnmoved = 0xFFFFFFFEu;
struct PrimMoves {
  uint32_t first = 0;
  std::vector<uint32_t> origin;
  std::map<uint32_t, uint32_t> final_pos;
};
struct Update {
  Dirty dirty;
  std::vector<uint2> remap_prims, remap_lights;  // (index, value)
  std::vector<uint32_t> light_origin, light_final;
  std::map<uint32_t, PrimMoves> prim_moves;
  std::vector<DoneTask> completed;
  int first_error = CT_OK;
  std::string message;
  void error(int code, const std::string &m) {
    if (first_error == CT_OK) {
      first_error = code;
      message = m;
    }
  }
  void lights_begin(uint32_t n) {
    light_origin.resize(n);
    for (uint32_t l = 0; l < n; ++l) light_origin[l] = l;
    light_final.assign(n, kUnmoved);
  }
  uint32_t origin_of(uint32_t l) const { return l < light_origin.size() ? light_origin[l] : kNoOrigin; }
  void light_removed(uint32_t l) {
    const uint32_t o = origin_of(l);
    if (o < light_final.size()) light_final[o] = kInvalid;
  }
  void light_moved(uint32_t from, uint32_t to) {
    const uint32_t o = origin_of(from);
    if (to >= light_origin.size()) light_origin.resize(to + 1, kNoOrigin);
    light_origin[to] = o;
    if (o < light_final.size()) light_final[o] = to;
  }
  void light_restored(uint32_t l) {
    if (l >= light_origin.size()) light_origin.resize(l + 1, kNoOrigin);
    light_origin[l] = kNoOrigin;
  }
  void lights_end() {
    for (uint32_t o = 0; o < light_final.size(); ++o)
      if (light_final[o] != kUnmoved) remap_lights.push_back(make_uint2(o, light_final[o]));
  }
  PrimMoves &moves(uint32_t id, uint32_t first, uint32_t count) {
    auto it = prim_moves.find(id);
    if (it != prim_moves.end()) return it->second;
    PrimMoves &m = prim_moves[id];
    m.first = first;
    m.origin.resize(count);
    for (uint32_t k = 0; k < count; ++k) m.origin[k] = k;
    return m;
  }
  static uint32_t origin_in(const PrimMoves &m, uint32_t l) { return l < m.origin.size() ? m.origin[l] : kNoOrigin; }
  void prim_removed(PrimMoves &m, uint32_t l) {
    const uint32_t o = origin_in(m, l);
    if (o != kNoOrigin) m.final_pos[o] = kInvalid;
  }
  void prim_moved(PrimMoves &m, uint32_t from, uint32_t to) {
    const uint32_t o = origin_in(m, from);
    if (to >= m.origin.size()) m.origin.resize(to + 1, kNoOrigin);
    m.origin[to] = o;
    if (o != kNoOrigin) m.final_pos[o] = to;
  }
  void prim_restored(PrimMoves &m, uint32_t l) {
    if (l >= m.origin.size()) m.origin.resize(l + 1, kNoOrigin);
    m.origin[l] = kNoOrigin;
  }
  void prims_end() {
    for (const auto &im : prim_moves)
      for (const auto &of : im.second.final_pos) remap_prims.push_back(make_uint2(im.second.first + of.first, of.second));
  }
};

void touch(Accel &a, Update &u, uint32_t id, uint32_t count) {
  u.dirty.ranges.emplace_back(a.instances[id].first_primitive_index, count);
}

void free_instance(Accel &a, Update &u, uint32_t id) {
  a.prim_alloc.deallocate(a.instances[id].first_primitive_index, align32(a.instances[id].primitive_count));
  a.instance_used[id] = 0;
  a.instances[id] = ct_instance{};
  auto drop = [&](std::vector<uint32_t> &v) { v.erase(std::remove(v.begin(), v.end(), id), v.end()); };
  drop(u.dirty.rebuild);
  drop(u.dirty.refit);
}

// The light at index `l` records its owner (instance, primitive); make the
// owner's PRIMITIVE.light_index point back at `l` (the owner may sit in
// another instance: its slot is uploaded too).
void relink_light_owner(Accel &a, Update &u, uint32_t l) {
  const ct_object_info o = a.cube_lights_host[l].instance_info;
  if (used(a, o.instance_id) && o.primitive_id < a.instances[o.instance_id].primitive_count) {
    const uint32_t slot = a.instances[o.instance_id].first_primitive_index + o.primitive_id;
    a.prims[slot].light_index = l;
    u.dirty.ranges.emplace_back(slot, 1u);
  }
}

void task_build(Accel &a, Update &u, const ct_task &t, uint32_t &cursor) {
  for (uint32_t q = 0; q < t.u.blas_build_from_cpu.instance_count; ++q, ++cursor) {
    if (cursor >= a.max_instances) return u.error(CT_E_CAPACITY, "update_scene: staging cursor out of range");
    const ct_instance &ti = a.temp_instances[cursor];
    uint32_t id = 0;
    while (id < a.instance_used.size() && a.instance_used[id]) ++id;
    if (id >= a.max_instances) return u.error(CT_E_CAPACITY, "update_scene: no free instance id");
    if (ti.primitive_count == 0) {
      u.error(CT_E_INVALID, "update_scene: instance without primitives");
      continue;
    }
    if ((uint64_t)ti.first_primitive_index + ti.primitive_count > a.max_primitives) {
      u.error(CT_E_INVALID, "update_scene: staged primitive range out of bounds");
      continue;
    }
    uint32_t off = 0;
    if (!a.prim_alloc.allocate(align32(ti.primitive_count), off))
      return u.error(CT_E_CAPACITY, "update_scene: primitive free list exhausted");
    if (id >= a.instance_used.size()) {
      a.instance_used.resize(id + 1, 0);
      a.instances.resize(id + 1, ct_instance{});
      a.w2o.resize(16 * (size_t)(id + 1), 0.0f);
      a.prev_transform.resize(16 * (size_t)(id + 1), 0.0f);
      a.prev_first.resize(id + 1, 0u);
    }
    a.instance_used[id] = 1;
    uint32_t end = off + align32(ti.primitive_count);
    if (end > a.prim_high_water) {
      a.prim_high_water = end;
      a.aabbs.resize(end);
      a.prims.resize(end);
      u.dirty.full = true;
    }
    std::memcpy(&a.aabbs[off], &a.aabb_host[ti.first_primitive_index], sizeof(ct_aabb) * ti.primitive_count);
    std::memcpy(&a.prims[off], &a.primitives_host[ti.first_primitive_index], sizeof(ct_primitive) * ti.primitive_count);
    ct_instance &in = a.instances[id];
    std::memcpy(in.transform, t.u.blas_build_from_cpu.transform, sizeof(in.transform));
    in.first_primitive_index = off;
    in.primitive_count = ti.primitive_count;
    host_inverse(in.transform, &a.w2o[16 * (size_t)id]);
    // A new instance has no earlier frame: its previous state is its own.
    std::memcpy(&a.prev_transform[16 * (size_t)id], in.transform, sizeof(in.transform));
    a.prev_first[id] = off;
    mark(u.dirty.rebuild, id);
    touch(a, u, id, in.primitive_count);
  }
}

// delete_aabb_device_buffer + update_remapping_buffer +
// update_light_remapping_buffer + delete_light_device_buffer
// (accel_struct_mngr.cpp:607-817, 1155-1196).
bool task_delete_primitive(Accel &a, Update &u, ct_task &t, DoneTask &d) {
  const uint32_t id = t.u.blas_delete_primitive_from_cpu.instance_index;
  const uint32_t del = t.u.blas_delete_primitive_from_cpu.del_primitive_index;
  if (!used(a, id)) return u.error(CT_E_INVALID, "update_scene: DELETE_PRIMITIVE of an unused instance"), false;
  ct_instance &in = a.instances[id];
  if (del >= in.primitive_count)
    return u.error(CT_E_INVALID, "update_scene: DELETE_PRIMITIVE index out of range"), false;
  if (in.primitive_count == 1)
    return u.error(CT_E_INVALID, "update_scene: DELETE_PRIMITIVE would empty the instance (use DELETE_BLAS)"), false;
  const uint32_t first = in.first_primitive_index, last = in.primitive_count - 1;
  touch(a, u, id, in.primitive_count);
  d.task = t;
  d.last = last;
  d.aabb = a.aabbs[first + del];
  d.prim = a.prims[first + del];
  uint32_t L = a.prims[first + del].light_index, Lx = kInvalid;
  if (L != kInvalid && L >= a.cube_light_count) L = kInvalid;  // a stale link names no light
  if (L != kInvalid) {
    d.light_rec = a.cube_lights_host[L];
    Lx = a.cube_light_count - 1;
    if (Lx == L) Lx = kInvalid;
    u.light_removed(L);
    if (Lx != kInvalid) {
      a.cube_lights_host[L] = a.cube_lights_host[Lx];
      u.light_moved(Lx, L);
      relink_light_owner(a, u, L);
    }
    --a.cube_light_count;
  }
  uint32_t Lm = kInvalid;
  if (last != del) {
    a.aabbs[first + del] = a.aabbs[first + last];
    a.prims[first + del] = a.prims[first + last];
    Lm = a.prims[first + del].light_index;
    if (Lm != kInvalid && Lm < a.cube_light_count) a.cube_lights_host[Lm].instance_info.primitive_id = del;
  }
  in.primitive_count = last;
  d.light = L;
  d.light_moved = Lx;
  d.light_count_after = a.cube_light_count;
  if (Lx != kInvalid) d.moved_rec = a.cube_lights_host[L];
  t.u.blas_delete_primitive_from_cpu.remap_primitive_index = last;
  t.u.blas_delete_primitive_from_cpu.del_light_index = L;
  t.u.blas_delete_primitive_from_cpu.remap_light_index = Lx;
  t.u.blas_delete_primitive_from_cpu.remap_primitive_light_index = Lm;
  d.task = t;
  PrimMoves &pm = u.moves(id, first, last + 1);
  u.prim_removed(pm, del);
  if (last != del) u.prim_moved(pm, last, del);
  mark(u.dirty.rebuild, id);
  return true;
}

// Whether the light list is still the one a primitive deletion left: same
// length, and the record it moved into the hole still there.  Any other
// change to the list (a GPU deletion that removed lights, a load that
// appended some) invalidates the deletion's saved light indices.
bool lights_as_left(const Accel &a, const DoneTask &d) {
  if (d.light == kInvalid) return true;
  if (a.cube_light_count != d.light_count_after) return false;
  if (d.light_moved == kInvalid) return true;
  return std::memcmp(&a.cube_lights_host[d.light], &d.moved_rec, sizeof(ct_light)) == 0;
}

// Exact inverse of task_delete_primitive (the reference's own undo,
// restore_aabb_device_buffer / restore_remapping_buffer /
// restore_cube_light_remapping_buffer, accel_struct_mngr.cpp:915-1083).
void undo_delete_primitive(Accel &a, Update &u, const DoneTask &d) {
  const uint32_t id = d.task.u.blas_delete_primitive_from_cpu.instance_index;
  const uint32_t del = d.task.u.blas_delete_primitive_from_cpu.del_primitive_index;
  if (!used(a, id) || a.instances[id].primitive_count != d.last)
    return u.error(CT_E_STATE, "update_scene: UNDO of a primitive deletion whose instance has changed since");
  if (!lights_as_left(a, d))
    return u.error(CT_E_STATE, "update_scene: UNDO of a primitive deletion whose cube light list has changed since");
  if (d.light != kInvalid && a.cube_light_count >= a.max_cube_lights)
    return u.error(CT_E_CAPACITY, "update_scene: UNDO needs a cube light slot");
  ct_instance &in = a.instances[id];
  const uint32_t first = in.first_primitive_index, last = d.last;
  in.primitive_count = last + 1;
  touch(a, u, id, in.primitive_count);
  PrimMoves &pm = u.moves(id, first, last);
  if (last != del) {
    a.aabbs[first + last] = a.aabbs[first + del];
    a.prims[first + last] = a.prims[first + del];
    const uint32_t Lm = a.prims[first + last].light_index;
    if (Lm != kInvalid && Lm < a.cube_light_count) a.cube_lights_host[Lm].instance_info.primitive_id = last;
    u.prim_moved(pm, del, last);
  }
  u.prim_restored(pm, del);
  a.aabbs[first + del] = d.aabb;
  a.prims[first + del] = d.prim;
  if (d.light != kInvalid) {
    ++a.cube_light_count;
    if (d.light_moved != kInvalid) {
      a.cube_lights_host[d.light_moved] = a.cube_lights_host[d.light];
      u.light_moved(d.light, d.light_moved);
      relink_light_owner(a, u, d.light_moved);
    }
    a.cube_lights_host[d.light] = d.light_rec;
    u.light_restored(d.light);
  }
  mark(u.dirty.rebuild, id);
}

void task_update(Accel &a, Update &u, const ct_task &t, bool &touched) {
  const uint32_t id = t.u.blas_update.instance_index;
  if (!used(a, id)) return u.error(CT_E_INVALID, "update_scene: UPDATE_BLAS of an unused instance");
  const uint32_t n = t.u.blas_update.primitive_count;
  const uint32_t pio = t.u.blas_update.primitive_index_buf_offset, ao = t.u.blas_update.aabb_buf_offset;
  if (n && ((uint64_t)pio + n > a.prim_index_host.size() || (uint64_t)ao + n > a.aabb_host.size()))
    return u.error(CT_E_INVALID, "update_scene: UPDATE_BLAS staging range out of bounds");
  ct_instance &in = a.instances[id];
  mat4_mul_host(in.transform, t.u.blas_update.transform, in.transform);
  host_inverse(in.transform, &a.w2o[16 * (size_t)id]);
  touched = true;
  bool moved = false;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t p = a.prim_index_host[pio + i];
    if (p >= in.primitive_count) continue;  // update_aabb_device_buffer skips these (:532-538)
    a.aabbs[in.first_primitive_index + p] = a.aabb_host[ao + i];
    moved = true;
  }
  if (moved) {
    touch(a, u, id, in.primitive_count);
    if (std::find(u.dirty.rebuild.begin(), u.dirty.rebuild.end(), id) == u.dirty.rebuild.end()) mark(u.dirty.refit, id);
  }
}

// DELETE_PRIMITIVE_BLAS_FROM_GPU (:2258-2290).  The device slots already
// hold the compacted primitives; the host mirror was refreshed from them.
// Cube lights of the GPU deletions of one update (the rearrangement leaves
// them alone), on indices taken before any of them moves: every light whose
// voxel survived gets the voxel's new local index, then every light whose
// voxel was deleted is removed as the CPU deletion removes one (the last
// light moves into its place and its owner is relinked), highest index
// first so that a light about to be removed is never the one moved.  The
// light remapping entries map each original index to its final one.
// A GPU-deleted instance's primitive records before its compaction: all n
// of them, or those of the deleted slots (the lights the deletion can
// remove: a kept voxel's light stays alive).
struct GpuOld {
  uint32_t id = 0, n = 0;
  std::vector<ct_primitive> recs;
};

void gpu_delete_lights(Accel &a, Update &u, const std::vector<ct_task> &queue, const std::vector<GpuOld> &old) {
  std::vector<uint8_t> alive(a.cube_light_count, 0);
  // one entry per instance (its first task); the instances own disjoint
  // lights, so they are scanned in parallel
  std::vector<std::pair<const ct_task *, const GpuOld *>> work;
  for (const ct_task &t : queue) {
    if (t.type != CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU) continue;
    const uint32_t id = t.u.blas_del_prim_gpu.instance_index;
    const GpuOld *o = nullptr;
    for (const GpuOld &g : old)
      if (g.id == id) o = &g;
    if (!o || !used(a, id)) continue;
    if (std::any_of(work.begin(), work.end(),
                    [&](const auto &w) { return w.first->u.blas_del_prim_gpu.instance_index == id; }))
      continue;
    work.emplace_back(&t, o);
  }
  std::vector<std::vector<uint32_t>> gone_of(work.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < (int64_t)work.size(); ++q) {
    const ct_task &t = *work[q].first;
    const GpuOld &o = *work[q].second;
    const uint32_t id = t.u.blas_del_prim_gpu.instance_index;
    const uint32_t f = a.instances[id].first_primitive_index, n = o.n;
    const uint32_t keep = n - std::min(t.u.blas_del_prim_gpu.del_prim_count, n);
    for (uint32_t k = 0; k < keep; ++k) {
      const uint32_t L = a.prims[f + k].light_index;
      if (L != kInvalid && L < a.cube_light_count) {
        a.cube_lights_host[L].instance_info.primitive_id = k;
        alive[L] = 1;
      }
    }
    for (const ct_primitive &p : o.recs)
      if (p.light_index != kInvalid && p.light_index < a.cube_light_count) gone_of[q].push_back(p.light_index);
  }
  std::vector<uint32_t> gone;
  for (const auto &g : gone_of) gone.insert(gone.end(), g.begin(), g.end());
  gone.erase(std::remove_if(gone.begin(), gone.end(), [&](uint32_t L) { return alive[L] != 0; }), gone.end());
  std::sort(gone.begin(), gone.end());
  gone.erase(std::unique(gone.begin(), gone.end()), gone.end());
  for (auto it = gone.rbegin(); it != gone.rend(); ++it) {
    const uint32_t L = *it, Lx = a.cube_light_count - 1;
    u.light_removed(L);
    if (Lx != L) {
      a.cube_lights_host[L] = a.cube_lights_host[Lx];
      u.light_moved(Lx, L);
      relink_light_owner(a, u, L);
    }
    --a.cube_light_count;
  }
}

void task_delete_from_gpu(Accel &a, Update &u, const ct_task &t) {
  const uint32_t id = t.u.blas_del_prim_gpu.instance_index;
  if (!used(a, id)) return u.error(CT_E_INVALID, "update_scene: DELETE_PRIMITIVE_FROM_GPU of an unused instance");
  ct_instance &in = a.instances[id];
  if (in.primitive_count <= t.u.blas_del_prim_gpu.del_prim_count) {
    free_instance(a, u, id);
    return;
  }
  in.primitive_count -= t.u.blas_del_prim_gpu.del_prim_count;
  // No upload: the device slots are the compaction's, which the host
  // mirror was refreshed from.
  mark(u.dirty.rebuild, id);
}

// process_settling_task_queue (:2586-2680): the remapping tables are
// cleared and the previous instance state becomes the current one.
int scatter(hipStream_t st, const std::vector<uint2> &w, uint32_t *dst, DevBuf &buf);

int settle(Accel &a, hipStream_t st) {
  BvhSet &S = a.set;
  // Single entries (most of them) are cleared by one scatter per table,
  // longer ranges by a fill each.
  std::vector<uint2> zp, zl;
  for (const auto &r : a.settle_prims) {
    if (!S.remap_prim.p || r.first >= S.remap_prim_cap) continue;
    const uint32_t n = std::min(r.second, S.remap_prim_cap - r.first);
    if (n == 1)
      zp.push_back(make_uint2(r.first, 0u));
    else if (n)
      hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n)), dim3(kBlock), 0, st, S.remap_prim.as<uint32_t>() + r.first, 0u, n);
  }
  for (uint32_t l : a.settle_lights)
    if (S.remap_light.p && l < S.remap_light_cap) zl.push_back(make_uint2(l, 0u));
  CT_HIP(hipGetLastError());
  int rc;
  if ((rc = scatter(st, zp, S.remap_prim.as<uint32_t>(), a.upload)) ||
      (rc = scatter(st, zl, S.remap_light.as<uint32_t>(), a.upload)))
    return rc;
  a.settle_prims.clear();
  a.settle_lights.clear();
  for (size_t id = 0; id < a.instances.size(); ++id) {
    std::memcpy(&a.prev_transform[16 * id], a.instances[id].transform, 16 * sizeof(float));
    a.prev_first[id] = a.instances[id].first_primitive_index;
  }
  return upload_tables(a, st);
}

// One write per index, the last one queued: the scatter kernel writes its
// pairs in parallel, so duplicate indices would race.
void last_write_wins(std::vector<uint2> &w) {
  std::stable_sort(w.begin(), w.end(), [](const uint2 &p, const uint2 &q) { return p.x < q.x; });
  size_t n = 0;
  for (size_t i = 0; i < w.size(); ++i) {
    if (i + 1 < w.size() && w[i + 1].x == w[i].x) continue;
    w[n++] = w[i];
  }
  w.resize(n);
}

int copy_device(DevBuf &dst, const DevBuf &src, size_t bytes, hipStream_t st) {
  if (!src.p || bytes == 0) return CT_OK;
  int rc;
  if ((rc = dst.reserve(bytes))) return rc;
  CT_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, st));
  return CT_OK;
}

// Copies the committed scene into the previous set before an update
// changes it (stream-ordered device copies; about 116 B per primitive slot).
int snapshot_previous(Accel &a, hipStream_t st) {
  const BvhSet &S = a.set;
  PrevSet &P = a.prev;
  P.valid = false;
  if (!S.nodes.p || S.tlas_count == 0) return CT_OK;
  const size_t cap = S.prim_capacity, ni = a.instances.size();
  int rc;
  if ((rc = copy_device(P.nodes, S.nodes, sizeof(DNode) * cap, st)) ||
      (rc = copy_device(P.leaves, S.leaves, 16 * cap, st)) ||
      (rc = copy_device(P.aabbs, S.aabbs, sizeof(ct_aabb) * cap, st)) ||
      (rc = copy_device(P.prims, S.prims, sizeof(ct_primitive) * cap, st)) ||
      (rc = copy_device(P.instances_d, S.instances_d, sizeof(DInst) * ni, st)) ||
      (rc = copy_device(P.inst_aos, S.inst_aos, sizeof(ct_instance) * ni, st)) ||
      (rc = copy_device(P.tlas_nodes, S.tlas_nodes, S.tlas_nodes.bytes, st)) ||
      (rc = copy_device(P.tlas_inst, S.tlas_inst, S.tlas_inst.bytes, st)))
    return rc;
  P.tlas_root = S.tlas_root;
  P.tlas_count = S.tlas_count;
  P.tlas_depth = S.tlas_depth;
  P.generation = S.generation;
  P.n_inst = (uint32_t)ni;
  P.prim_capacity = (uint32_t)cap;
  P.valid = true;
  return CT_OK;
}

// Writes dst[w.x] = w.y for every pair through `buf` (device scratch).
int scatter(hipStream_t st, const std::vector<uint2> &w, uint32_t *dst, DevBuf &buf) {
  if (w.empty()) return CT_OK;
  int rc;
  if ((rc = buf.reserve(sizeof(uint2) * w.size()))) return rc;
  CT_HIP(hipMemcpyAsync(buf.p, w.data(), sizeof(uint2) * w.size(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_scatter_u32, dim3(blocks(w.size())), dim3(kBlock), 0, st, buf.as<uint2>(), (uint32_t)w.size(), dst);
  CT_HIP(hipGetLastError());
  CT_HIP(hipStreamSynchronize(st));  // `w` and the next use of `buf` wait for the kernel
  return CT_OK;
}

}  // namespace
}  // namespace ct

extern "C" {

int ct_accel_update_scene(ct_accel *a, int synchronize, void *stream, int *did_work) {
  if (!a) return fail(CT_E_INVALID, "update_scene: null handle");
  if (did_work) *did_work = 0;
  if (a->tasks.empty() && a->status != CT_AS_SETTLING) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  if (stream) CT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  hipStream_t st = a->stream;
  int rc;
  // A caller's compaction pass ahead of DELETE_PRIMITIVE_BLAS_FROM_GPU has
  // written device slots and remapping entries: take them before the
  // settle clears the tables.
  const uint64_t first_serial = a->task_serial - a->tasks.size();
  std::vector<ct_task> queue;
  while (!a->tasks.empty()) {
    queue.push_back(a->tasks.front());
    a->tasks.pop();
  }
  std::vector<Accel::Compaction> comps;  // the library's own compactions, consumed here
  comps.swap(a->compactions);
  std::vector<uint2> gpu_entries;  // the caller's compaction's remapping entries (slot, value), value != 0
  std::vector<GpuOld> gpu_old;     // records before the compaction
  // An instance the library's compaction pass rearranged, with its task the
  // only GPU deletion of it in the queue, takes the pass's moves on the host
  // mirror (the pass wrote nothing else: the tables were clear, as it runs
  // only on an idle accel).  Any other GPU-deleted instance's slots (AABB,
  // primitive, remapping entry) come back through one pinned staging
  // buffer, then into the host mirror.  Slot ranges separated only by
  // alignment padding are read as one run (the padding belongs to no
  // primitive).
  std::vector<std::pair<uint32_t, uint32_t>> rb;  // (first slot, count) per instance
  std::vector<const Accel::Compaction *> moves;   // per entry of gpu_old; null: read back
  for (size_t j = 0; j < queue.size(); ++j) {
    const ct_task &t = queue[j];
    if (t.type != CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU) continue;
    const uint32_t id = t.u.blas_del_prim_gpu.instance_index;
    if (!used(*a, id) || !a->set.aabbs.p) continue;
    const ct_instance &in = a->instances[id];
    const uint32_t f = in.first_primitive_index, n = in.primitive_count;
    if (std::any_of(gpu_old.begin(), gpu_old.end(), [&](const GpuOld &g) { return g.id == id; })) continue;
    const Accel::Compaction *c = nullptr;
    for (const Accel::Compaction &k : comps)
      if (k.serial == first_serial + j && k.id == id && k.first == f && k.count == n) c = &k;
    for (size_t k = j + 1; c && k < queue.size(); ++k)
      if (queue[k].type == CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU && queue[k].u.blas_del_prim_gpu.instance_index == id)
        c = nullptr;
    GpuOld g;
    g.id = id;
    g.n = n;
    if (c) {
      g.recs.reserve(c->del.size());
      for (uint32_t d : c->del) g.recs.push_back(a->prims[f + d]);
    } else {
      g.recs.assign(a->prims.begin() + f, a->prims.begin() + f + n);
      rb.emplace_back(f, n);
    }
    gpu_old.push_back(std::move(g));
    moves.push_back(c);
  }
  {
    std::vector<std::vector<uint2>> found(moves.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)moves.size(); ++q) {
      const Accel::Compaction *c = moves[q];
      if (!c) continue;
      const uint32_t f = c->first;
      for (size_t k = 0; k < c->sur.size(); ++k) {
        a->aabbs[f + c->del[k]] = a->aabbs[f + c->sur[k]];
        a->prims[f + c->del[k]] = a->prims[f + c->sur[k]];
      }
      std::vector<uint2> &e = found[q];
      e.reserve(c->del.size() + c->sur.size());
      for (uint32_t d : c->del) e.push_back(make_uint2(f + d, 0xFFFFFFFFu));
      for (size_t k = 0; k < c->sur.size(); ++k)
        if (c->del[k] != 0u) e.push_back(make_uint2(f + c->sur[k], c->del[k]));
    }
    for (const auto &e : found) gpu_entries.insert(gpu_entries.end(), e.begin(), e.end());
  }
  if (!rb.empty()) {
    std::vector<std::pair<uint32_t, uint32_t>> runs(rb);  // merged (first, count)
    std::sort(runs.begin(), runs.end());
    size_t m = 0;
    for (size_t k = 1; k < runs.size(); ++k) {
      const uint64_t end = (uint64_t)runs[m].first + runs[m].second;
      // a gap of fewer than CT_PRIMITIVE_ALIGNMENT slots is an instance's
      // alignment padding (instances start aligned): no other instance
      if ((uint64_t)runs[k].first < end + CT_PRIMITIVE_ALIGNMENT) {
        runs[m].second = (uint32_t)std::max<uint64_t>(end, (uint64_t)runs[k].first + runs[k].second) - runs[m].first;
      } else {
        runs[++m] = runs[k];
      }
    }
    runs.resize(m + 1);
    // Chunks of at most 64 Ki slots, copied in rounds through at most
    // kStageSlots slots of pinned staging (72 MiB), each round's chunks then
    // into the host mirror and the compaction's remapping entries collected,
    // over the host threads.
    constexpr size_t kSlotBytes = sizeof(ct_aabb) + sizeof(ct_primitive) + 4;
    constexpr uint32_t kChunk = 1u << 16;
    constexpr size_t kStageSlots = (size_t)1 << 21;
    std::vector<std::pair<uint32_t, uint32_t>> chunks;  // (first slot, count)
    size_t slots = 0;
    for (const auto &r : runs)
      for (uint64_t c = r.first; c < (uint64_t)r.first + r.second; c += kChunk) {
        chunks.emplace_back((uint32_t)c, (uint32_t)std::min<uint64_t>(kChunk, (uint64_t)r.first + r.second - c));
        slots += chunks.back().second;
      }
    const size_t stage = std::min(slots, kStageSlots);
    if ((rc = a->readback.reserve(kSlotBytes * stage))) return rc;
    ct_aabb *pa = a->readback.as<ct_aabb>();
    ct_primitive *pp = reinterpret_cast<ct_primitive *>(pa + stage);
    uint32_t *pr = reinterpret_cast<uint32_t *>(pp + stage);
    std::vector<std::vector<uint2>> found(chunks.size());
    std::vector<size_t> at(chunks.size());
    for (size_t q0 = 0; q0 < chunks.size();) {
      size_t q1 = q0, off = 0;
      for (; q1 < chunks.size() && off + chunks[q1].second <= stage; ++q1) {
        const uint32_t f = chunks[q1].first, n = chunks[q1].second;
        at[q1] = off;
        CT_HIP(hipMemcpyAsync(pa + off, a->set.aabbs.as<ct_aabb>() + f, sizeof(ct_aabb) * n, hipMemcpyDeviceToHost, st));
        CT_HIP(hipMemcpyAsync(pp + off, a->set.prims.as<ct_primitive>() + f, sizeof(ct_primitive) * n,
                              hipMemcpyDeviceToHost, st));
        CT_HIP(hipMemcpyAsync(pr + off, a->set.remap_prim.as<uint32_t>() + f, 4 * (size_t)n, hipMemcpyDeviceToHost,
                              st));
        off += n;
      }
      CT_HIP(hipStreamSynchronize(st));  // one wait per round
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t q = (int64_t)q0; q < (int64_t)q1; ++q) {
        const uint32_t c = chunks[q].first, n = chunks[q].second;
        const size_t src = at[q];
        std::memcpy(&a->aabbs[c], pa + src, sizeof(ct_aabb) * n);
        std::memcpy(&a->prims[c], pp + src, sizeof(ct_primitive) * n);
        for (uint32_t i = 0; i < n; ++i)
          if (pr[src + i] != 0u) found[q].push_back(make_uint2(c + i, pr[src + i]));
      }
      q0 = q1;
    }
    for (const auto &f : found) gpu_entries.insert(gpu_entries.end(), f.begin(), f.end());
  }
  std::sort(gpu_entries.begin(), gpu_entries.end(), [](uint2 x, uint2 y) { return x.x < y.x; });
  bool worked = false;
  if (a->status == CT_AS_SETTLING) {
    if ((rc = settle(*a, st))) return rc;
    a->status = CT_AS_IDLE;
    a->prev.valid = false;
    worked = true;
  }
  if (queue.empty()) {
    if (did_work) *did_work = worked ? 1 : 0;
    return CT_OK;
  }
  if ((rc = snapshot_previous(*a, st))) return rc;
  // process_task_queue, accel_struct_mngr.cpp:2115-2443
  a->status = CT_AS_UPDATING;
  Update u;
  u.lights_begin(a->cube_light_count);
  gpu_delete_lights(*a, u, queue, gpu_old);
  uint32_t cursor = 0;
  bool tables = false;
  for (ct_task &t : queue) {
    switch (t.type) {
      case CT_TASK_BUILD_BLAS_FROM_CPU:
        task_build(*a, u, t, cursor);
        tables = true;
        break;
      case CT_TASK_DELETE_BLAS_FROM_CPU: {
        const uint32_t id = t.u.blas_delete_from_cpu.instance_index;
        if (!used(*a, id)) {
          u.error(CT_E_INVALID, "update_scene: DELETE_BLAS of an unused instance");
          continue;
        }
        t.u.blas_delete_from_cpu.first_primitive_index = a->instances[id].first_primitive_index;
        t.u.blas_delete_from_cpu.deleted_primitive_count = a->instances[id].primitive_count;
        free_instance(*a, u, id);
        tables = true;
        DoneTask d;
        d.task = t;
        u.completed.push_back(d);
        break;
      }
      case CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU: {
        DoneTask d;
        if (task_delete_primitive(*a, u, t, d)) u.completed.push_back(d);
        tables = true;
        break;
      }
      case CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU: {
        task_delete_from_gpu(*a, u, t);
        tables = true;
        DoneTask d;
        d.task = t;
        u.completed.push_back(d);
        break;
      }
      case CT_TASK_UPDATE_BLAS_FROM_CPU: {
        bool touched = false;
        task_update(*a, u, t, touched);
        if (touched) {
          tables = true;
          DoneTask d;
          d.task = t;
          u.completed.push_back(d);
        }
        break;
      }
      case CT_TASK_UNDO_OP_CPU: {
        if (a->done.empty()) break;
        DoneTask d = a->done.back();
        a->done.pop_back();
        if (d.task.type == CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU) undo_delete_primitive(*a, u, d);
        tables = true;
        break;
      }
      default:
        break;
    }
  }
  u.lights_end();
  u.prims_end();
  a->temp_instance_count = 0;
  a->temp_primitive_count = 0;
  a->temp_prim_index_count = 0;
  if (tables) {
    if ((rc = commit(*a, u.dirty))) {
      a->status = CT_AS_IDLE;
      return rc;
    }
  }
  // Remapping writes of this update (and the caller's, re-applied), then
  // remembered for the settle.
  // The caller's compaction came first, so its entries go first and a later
  // task's entry for the same slot wins.
  u.remap_prims.insert(u.remap_prims.begin(), gpu_entries.begin(), gpu_entries.end());
  for (const uint2 &w : u.remap_prims)
    if (w.x < a->set.remap_prim_cap) a->settle_prims.emplace_back(w.x, 1u);
  for (const uint2 &w : u.remap_lights)
    if (w.x < a->set.remap_light_cap) a->settle_lights.push_back(w.x);
  std::vector<uint2> wp, wl;
  for (const uint2 &w : u.remap_prims)
    if (w.x < a->set.remap_prim_cap) wp.push_back(w);
  for (const uint2 &w : u.remap_lights)
    if (w.x < a->set.remap_light_cap) wl.push_back(w);
  last_write_wins(wp);
  last_write_wins(wl);
  if ((rc = scatter(st, wp, a->set.remap_prim.as<uint32_t>(), a->upload)) ||
      (rc = scatter(st, wl, a->set.remap_light.as<uint32_t>(), a->upload)))
    return rc;
  for (DoneTask &d : u.completed) {
    a->done.push_back(d);
    if (a->done.size() > kMaxDone) a->done.pop_front();
  }
  a->status = CT_AS_SETTLING;
  if (synchronize) CT_HIP(hipStreamSynchronize(st));
  if (did_work) *did_work = 1;
  if (u.first_error != CT_OK) return fail(u.first_error, u.message);
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
  if (previous && ct::previous_distinct(*a)) {  // get_previous_tlas, hpp:219-223
    o->nodes = a->prev.tlas_nodes.p;
    o->leaves = a->prev.tlas_inst.p;
    o->root_ref = a->prev.tlas_root;
    o->instance_count = a->prev.tlas_count;
    o->depth = a->prev.tlas_depth;
    o->generation = a->prev.generation;
    return CT_OK;
  }
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
  const bool prev = previous && ct::previous_distinct(*a);  // get_previous_*_buffer, hpp:225-250
  o->instances = prev ? a->prev.inst_aos.p : a->set.inst_aos.p;
  o->primitives = prev ? a->prev.prims.p : a->set.prims.p;
  o->aabbs = prev ? a->prev.aabbs.p : a->set.aabbs.p;
  o->cube_lights = a->set.cube_lights.p;
  o->remapped_primitives = a->set.remap_prim.p;
  o->remapped_cube_lights = a->set.remap_light.p;
  o->instance_count = prev ? a->prev.n_inst : (uint32_t)a->instances.size();
  o->primitive_capacity = prev ? a->prev.prim_capacity : a->set.prim_capacity;
  o->cube_light_count = a->set.cube_light_count;
  o->cube_light_capacity = a->set.remap_light_cap;
  return CT_OK;
}

int ct_accel_read_world(ct_accel *a, ct_aabb *aabbs, ct_primitive *prims, uint32_t prim_capacity, ct_light *lights,
                        uint32_t light_capacity, uint32_t *remap_prims, uint32_t *remap_lights) {
  if (!a) return fail(CT_E_INVALID, "read_world: null handle");
  CT_HIP(hipSetDevice(a->device));
  CT_HIP(hipStreamSynchronize(a->stream));
  const BvhSet &S = a->set;
  const size_t np = std::min<size_t>(prim_capacity, S.prim_capacity);
  const size_t nl = std::min<size_t>(light_capacity, S.cube_light_hw);
  if (aabbs && np && S.aabbs.p) CT_HIP(hipMemcpy(aabbs, S.aabbs.p, sizeof(ct_aabb) * np, hipMemcpyDeviceToHost));
  if (prims && np && S.prims.p) CT_HIP(hipMemcpy(prims, S.prims.p, sizeof(ct_primitive) * np, hipMemcpyDeviceToHost));
  if (lights && nl && S.cube_lights.p) CT_HIP(hipMemcpy(lights, S.cube_lights.p, sizeof(ct_light) * nl, hipMemcpyDeviceToHost));
  const size_t nr = std::min<size_t>(prim_capacity, S.remap_prim_cap);
  if (remap_prims && nr && S.remap_prim.p)
    CT_HIP(hipMemcpy(remap_prims, S.remap_prim.p, 4 * nr, hipMemcpyDeviceToHost));
  const size_t nrl = std::min<size_t>(light_capacity, S.remap_light_cap);
  if (remap_lights && nrl && S.remap_light.p)
    CT_HIP(hipMemcpy(remap_lights, S.remap_light.p, 4 * nrl, hipMemcpyDeviceToHost));
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
