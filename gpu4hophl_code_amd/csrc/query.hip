// query.hip -- batched (s,t,k) hop-bounded distance queries on gfx950.
//
// Semantics of hop_constrained_extract_distance
// (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:144-222):
//   k < 0 -> INT32_MAX; s == t -> 0; k == 0 -> INT32_MAX; otherwise
//   min{ d1 + d2 : (w,h1,d1) in L(s), (w,h2,d2) in L(t), h1 + h2 <= k }
//   or INT32_MAX when no pair qualifies.
// The reference's GPU query (Query_mdnd/src/gpu_query.cu:5-177) launches three
// kernels per query; here one launch answers a whole batch.
//
// One wavefront per query.  The hubs of the longer list are staged in LDS
// with coalesced loads (64 labels per instruction) -- all of them for lists
// up to 2048 labels, every stride-th one (a skip index) for longer lists; the
// lanes then take the shorter list 64 labels at a time, binary-search each
// hub in LDS (plus at most log2(stride) global probes inside the block),
// walk that hub's run (<= K+1 labels, hops ascending) while h1 + h2 <= k and
// keep a running minimum; a wave min-reduction gives the answer.  A query's
// critical path is a few coalesced loads plus LDS probes, which keeps the
// latency of small serving batches low.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "device.hpp"

namespace hsdl {

constexpr uint32_t kQueryWaves = 4;        // queries (wavefronts) per workgroup
constexpr uint32_t kQueryLdsHubs = 2048;   // hubs of the longer list per wave

template <class Getter>
__device__ inline uint32_t lower_bound_by(Getter get, uint32_t len, uint32_t w) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (get(mid) < w) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_query(uint32_t n, const unsigned long long *off,
                                               const unsigned long long *lab, uint64_t nq,
                                               const uint32_t *qs, const uint32_t *qt,
                                               const int8_t *qk, int32_t *out, uint32_t lds_hubs) {
  __shared__ uint32_t hubs_all[kQueryWaves * kQueryLdsHubs];
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t *hubs = hubs_all + wv * kQueryLdsHubs;
  const uint64_t q = (uint64_t)blockIdx.x * kQueryWaves + wv;
  if (q >= nq) return;  // whole wave leaves together
  const uint32_t s = qs[q], t = qt[q];
  const int k = qk[q];
  if (k < 0 || s >= n || t >= n || s == t || k == 0) {
    if (lane == 0) out[q] = (k >= 0 && s == t && s < n) ? 0 : INT32_MAX;
    return;
  }
  const unsigned long long sb = off[s], tb = off[t];
  uint32_t ls = (uint32_t)(off[s + 1] - sb), lt = (uint32_t)(off[t + 1] - tb);
  const unsigned long long *A = lab + sb, *B = lab + tb;
  if (ls > lt) {  // A = shorter list, B = longer list
    const unsigned long long *T = A; A = B; B = T;
    const uint32_t tl = ls; ls = lt; lt = tl;
  }
  if (ls == 0) {  // an empty list shares no hub
    if (lane == 0) out[q] = INT32_MAX;
    return;
  }
  const uint32_t kk = (uint32_t)k;
  // LDS skip index of B: every stride-th hub (stride 1 = all of B for lists
  // up to kQueryLdsHubs).  A lookup searches LDS for the block, then at most
  // log2(stride) labels of B in global memory.
  const uint32_t stride = (lt + lds_hubs - 1) / lds_hubs;
  const uint32_t ns = (lt + stride - 1) / stride;
  for (uint32_t i = lane; i < ns; i += kWave) hubs[i] = label_hub(B[(uint64_t)i * stride]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t best = kInf32;
  for (uint32_t j = lane; j < ls; j += kWave) {
    const unsigned long long a = A[j];
    const uint32_t w = label_hub(a), h1 = label_hop(a), d1 = label_dist(a);
    if (h1 > kk) continue;
    // first sample >= w, then the first position >= w inside its block
    const uint32_t si = lower_bound_by([&](uint32_t i) { return hubs[i]; }, ns, w);
    const uint32_t lo = si == 0 ? 0 : (si - 1) * stride + 1;
    const uint32_t hi = si < ns ? si * stride : lt;
    uint32_t p = lo + lower_bound_by([&](uint32_t i) { return label_hub(B[lo + i]); }, hi - lo, w);
    for (; p < lt; ++p) {
      const unsigned long long b = B[p];
      if (label_hub(b) != w || h1 + label_hop(b) > kk) break;  // hops ascend within a hub
      const uint32_t dsum = d1 + label_dist(b);
      best = dsum < best ? dsum : best;
    }
  }
  best = wave_min_u32(best);
  if (lane == 0) out[q] = (best == kInf32) ? INT32_MAX : (int32_t)best;
}

int launch_query(const hsdl_index *idx, size_t nq, const uint32_t *d_s, const uint32_t *d_t,
                 const int8_t *d_k, int32_t *d_out, hipStream_t st) {
  if (nq == 0) return HSDL_OK;
  const uint64_t blocks = (nq + kQueryWaves - 1) / kQueryWaves;
  if (blocks > 0x7FFFFFFFull) return fail(HSDL_EINVAL, "query batch too large for one launch");
  // LDS skip-index capacity per query (tuning knob HSDL_QUERY_LDS_HUBS, <= 2048)
  const uint32_t lds_hubs = env_u32("HSDL_QUERY_LDS_HUBS", kQueryLdsHubs, 1, kQueryLdsHubs);
  hipLaunchKernelGGL(k_query, dim3((uint32_t)blocks), dim3(kQueryWaves * kWave), 0, st, idx->n,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     (uint64_t)nq, d_s, d_t, d_k, d_out, lds_hubs);
  HSDL_HIP(hipGetLastError());
  return HSDL_OK;
}

}  // namespace hsdl

using namespace hsdl;

extern "C" {

int hsdl_query_device(const hsdl_index *idx, size_t n, const uint32_t *d_s, const uint32_t *d_t,
                      const int8_t *d_k, int32_t *d_out, void *hip_stream) {
  if (!idx) return fail(HSDL_EINVAL, "null index");
  if (n && (!d_s || !d_t || !d_k || !d_out)) return fail(HSDL_EINVAL, "null query buffers");
  return launch_query(idx, n, d_s, d_t, d_k, d_out, (hipStream_t)hip_stream);
}

int hsdl_query(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
               const int8_t *k, int32_t *out) {
  if (!idx) return fail(HSDL_EINVAL, "null index");
  if (n == 0) return HSDL_OK;
  if (!s || !t || !k || !out) return fail(HSDL_EINVAL, "null query arrays");
  for (size_t i = 0; i < n; ++i)
    if (s[i] >= idx->n || t[i] >= idx->n) return fail(HSDL_EINVAL, "query vertex out of range");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  DevBuf ds, dt, dk, dout;
  if ((rc = ds.alloc(n * 4)) || (rc = dt.alloc(n * 4)) || (rc = dk.alloc(n)) ||
      (rc = dout.alloc(n * 4)))
    return rc;
  HSDL_HIP(hipMemcpy(ds.p, s, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dt.p, t, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dk.p, k, n, hipMemcpyHostToDevice));
  if ((rc = launch_query(idx, n, ds.as<uint32_t>(), dt.as<uint32_t>(), dk.as<int8_t>(),
                         dout.as<int32_t>(), nullptr)))
    return rc;
  HSDL_HIP(hipMemcpy(out, dout.p, n * 4, hipMemcpyDeviceToHost));
  return HSDL_OK;
}

}  // extern "C"
