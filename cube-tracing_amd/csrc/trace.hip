// trace.hip -- ct_trace_closest / ct_trace_any on caller-owned host arrays:
// the traversal kernels of the frame path, exposed for tests and tools.

#include <hip/hip_runtime.h>

#include <cstring>
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

}  // namespace

extern "C" {

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
