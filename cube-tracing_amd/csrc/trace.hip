// trace.hip -- ct_trace_closest / ct_trace_any on caller-owned host arrays:
// the traversal kernels of the frame path, exposed for tests and tools.

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "ct_internal.h"

using namespace ct;

namespace {

struct Staged {
  DevBuf o, d, t, occ, hit, cnt, stats;
};

int upload_rays(Staged &s, uint32_t n, const float *origins, const float *directions, const float *extra) {
  std::vector<float4> ho(n), hd(n);
  for (uint32_t i = 0; i < n; ++i) {
    ho[i] = make_float4(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2], extra ? extra[i] : 0.0f);
    hd[i] = make_float4(directions[3 * i], directions[3 * i + 1], directions[3 * i + 2], 0.0f);
  }
  int rc;
  if ((rc = s.o.reserve(16 * (size_t)n)) || (rc = s.d.reserve(16 * (size_t)n))) return rc;
  CT_HIP(hipMemcpy(s.o.p, ho.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(s.d.p, hd.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
  return CT_OK;
}

// ------------------------------------------------ allocation-free any hit
// ct_trace_any_noalloc stages its rays and work counters in one file-scope
// device array instead of memory it allocates.  Layout: ray_o [max] float4,
// ray_d [max] float4, ray_t [max] u32, then the queue length and the work
// counter.  Every GPU holds its own copy, at its own address, which the host
// looks up once per device with hipGetSymbolAddress and uses from then on.
constexpr uint32_t kNoAllocMaxRays = 1u << 20;
constexpr size_t kNoAllocO = 0, kNoAllocD = kNoAllocO + 16 * (size_t)kNoAllocMaxRays,
                 kNoAllocT = kNoAllocD + 16 * (size_t)kNoAllocMaxRays,
                 kNoAllocCnt = kNoAllocT + 4 * (size_t)kNoAllocMaxRays, kNoAllocBytes = kNoAllocCnt + 256;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch;  // per device, null until looked up

int scratch_on(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)device >= g_scratch.size()) g_scratch.resize(device + 1, nullptr);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    CT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(CT_E_HIP, "ct_trace_any_noalloc: no device address for the scratch array");
    g_scratch[device] = static_cast<unsigned char *>(ptr);
  }
  *out = g_scratch[device];
  return CT_OK;
}

// Packs the caller's device arrays into the staged ray records and sets the
// queue length (block 0, thread 0).
__global__ void k_stage_any(uint32_t n, const float *o, const float *d, const float *dist, const uint32_t *targets,
                            float4 *ro, float4 *rd, uint32_t *rt, uint32_t *cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) cnt[0] = n;
  if (i >= n) return;
  const uint32_t ti = targets ? targets[2 * i] : kInvalid;
  ro[i] = make_float4(o[3 * i], o[3 * i + 1], o[3 * i + 2], dist[i]);
  rd[i] = make_float4(d[3 * i], d[3 * i + 1], d[3 * i + 2], __uint_as_float(ti));
  rt[i] = targets ? targets[2 * i + 1] : kInvalid;
}

}  // namespace

extern "C" {

int ct_trace_any_noalloc(ct_accel *a, uint32_t n, const float *origins, const float *directions,
                         const float *distances, const uint32_t *targets, uint8_t *occluded_out, void *stream) {
  if (!a || (n && (!origins || !directions || !distances || !occluded_out)))
    return fail(CT_E_INVALID, "ct_trace_any_noalloc: null argument");
  if (n > kNoAllocMaxRays) return fail(CT_E_CAPACITY, "ct_trace_any_noalloc: more than 2^20 rays in one call");
  if (a->set.tlas_count == 0) return fail(CT_E_STATE, "ct_trace_any_noalloc: empty scene");
  if (n == 0) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  unsigned char *base = nullptr;
  int rc;
  if ((rc = scratch_on(a->device, &base))) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : a->stream;
  float4 *ro = reinterpret_cast<float4 *>(base + kNoAllocO), *rd = reinterpret_cast<float4 *>(base + kNoAllocD);
  uint32_t *rt = reinterpret_cast<uint32_t *>(base + kNoAllocT), *cnt = reinterpret_cast<uint32_t *>(base + kNoAllocCnt);
  hipLaunchKernelGGL(k_stage_any, dim3((n + 255) / 256), dim3(256), 0, st, n, origins, directions, distances, targets,
                     ro, rd, rt, cnt);
  CT_HIP(hipGetLastError());
  TraceAnyArgs t;
  std::memset(&t, 0, sizeof(t));
  t.ray_o = ro;
  t.ray_d = rd;
  t.ray_t = rt;
  t.queue = nullptr;
  t.count = cnt;
  t.head = cnt + 1;  // zeroed by the launch
  t.mask = nullptr;
  t.occ = occluded_out;
  t.stats = nullptr;
  t.max_rays = n;
  return launch_trace_any(scene_view(*a), t, st);
}

int ct_trace_closest(ct_accel *a, uint32_t n, const float *origins, const float *directions, float tmax,
                     float *t_out, uint32_t *ids_out, ct_trace_stats *stats) {
  if (!a || (n && (!origins || !directions || !t_out || !ids_out)))
    return fail(CT_E_INVALID, "ct_trace_closest: null argument");
  if (a->set.tlas_count == 0) return fail(CT_E_STATE, "ct_trace_closest: empty scene");
  if (n == 0) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  Staged s;
  int rc;
  if ((rc = upload_rays(s, n, origins, directions, nullptr)) || (rc = s.hit.reserve(16 * (size_t)n)) ||
      (rc = s.stats.reserve(3 * sizeof(unsigned long long))))
    return rc;
  CT_HIP(hipMemset(s.stats.p, 0, 3 * sizeof(unsigned long long)));
  TraceClosestArgs t;
  t.ray_o = s.o.as<float4>();
  t.ray_d = s.d.as<float4>();
  t.tmax = tmax;
  t.n = n;
  t.hit = s.hit.as<float4>();
  t.stats = stats ? s.stats.as<unsigned long long>() : nullptr;
  if ((rc = launch_trace_closest(scene_view(*a), t, a->stream))) return rc;
  CT_HIP(hipStreamSynchronize(a->stream));
  std::vector<float4> h(n);
  CT_HIP(hipMemcpy(h.data(), s.hit.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    t_out[i] = h[i].x;
    std::memcpy(&ids_out[2 * i], &h[i].y, 4);
    std::memcpy(&ids_out[2 * i + 1], &h[i].z, 4);
  }
  if (stats) {
    unsigned long long v[3];
    CT_HIP(hipMemcpy(v, s.stats.p, sizeof(v), hipMemcpyDeviceToHost));
    stats->rays = n;
    stats->nodes = v[0];
    stats->leaves = v[1];
    stats->instances = v[2];
  }
  return CT_OK;
}

int ct_trace_any(ct_accel *a, uint32_t n, const float *origins, const float *directions, const float *distances,
                 const uint32_t *targets, uint8_t *occluded_out, ct_trace_stats *stats) {
  if (!a || (n && (!origins || !directions || !distances || !occluded_out)))
    return fail(CT_E_INVALID, "ct_trace_any: null argument");
  if (a->set.tlas_count == 0) return fail(CT_E_STATE, "ct_trace_any: empty scene");
  if (n == 0) return CT_OK;
  CT_HIP(hipSetDevice(a->device));
  Staged s;
  int rc;
  if ((rc = upload_rays(s, n, origins, directions, distances)) || (rc = s.t.reserve(4 * (size_t)n)) ||
      (rc = s.occ.reserve((size_t)n)) || (rc = s.cnt.reserve(2 * sizeof(uint32_t))) ||
      (rc = s.stats.reserve(3 * sizeof(unsigned long long))))
    return rc;
  // targets: instance id rides in the direction's w lane, the primitive id
  // in ray_t.
  std::vector<float4> hd(n);
  CT_HIP(hipMemcpy(hd.data(), s.d.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<uint32_t> tp(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t ti = targets ? targets[2 * i] : kInvalid;
    tp[i] = targets ? targets[2 * i + 1] : kInvalid;
    std::memcpy(&hd[i].w, &ti, 4);
  }
  CT_HIP(hipMemcpy(s.d.p, hd.data(), 16 * (size_t)n, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(s.t.p, tp.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
  uint32_t cnt[2] = {n, 0};
  CT_HIP(hipMemcpy(s.cnt.p, cnt, sizeof(cnt), hipMemcpyHostToDevice));
  CT_HIP(hipMemset(s.stats.p, 0, 3 * sizeof(unsigned long long)));
  TraceAnyArgs t;
  std::memset(&t, 0, sizeof(t));
  t.ray_o = s.o.as<float4>();
  t.ray_d = s.d.as<float4>();
  t.ray_t = s.t.as<uint32_t>();
  t.queue = nullptr;
  t.count = s.cnt.as<uint32_t>();
  t.head = s.cnt.as<uint32_t>() + 1;
  t.mask = nullptr;
  t.occ = s.occ.as<uint8_t>();
  t.stats = stats ? s.stats.as<unsigned long long>() : nullptr;
  t.max_rays = n;
  if ((rc = launch_trace_any(scene_view(*a), t, a->stream))) return rc;
  CT_HIP(hipStreamSynchronize(a->stream));
  CT_HIP(hipMemcpy(occluded_out, s.occ.p, (size_t)n, hipMemcpyDeviceToHost));
  if (stats) {
    unsigned long long v[3];
    CT_HIP(hipMemcpy(v, s.stats.p, sizeof(v), hipMemcpyDeviceToHost));
    stats->rays = n;
    stats->nodes = v[0];
    stats->leaves = v[1];
    stats->instances = v[2];
  }
  return CT_OK;
}

}  // extern "C"
