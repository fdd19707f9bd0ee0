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
// A group of G lanes answers one query (G = 16: four queries per wavefront,
// which suits the short cleaned lists of degree-ranked graphs).  The group
// streams the shorter list in chunks of G labels; every lane binary-searches
// its label's hub in the longer list, walks that hub's (<= K+1) labels in hop
// order while h1 + h2 <= k, and keeps a running minimum.  A shuffle
// min-reduction inside the group gives the answer.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "device.hpp"

namespace hsdl {

constexpr int kQueryGroup = 16;

__device__ inline uint32_t lower_bound_hub(const unsigned long long *L, uint32_t len, uint32_t w) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (label_hub(L[mid]) < w) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_query(uint32_t n, const unsigned long long *off,
                                               const unsigned long long *lab, uint64_t nq,
                                               const uint32_t *qs, const uint32_t *qt,
                                               const int8_t *qk, int32_t *out) {
  const uint32_t lane = threadIdx.x & (kQueryGroup - 1);
  const uint64_t q = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kQueryGroup;
  const bool valid = q < nq;
  uint32_t s = 0, t = 0;
  int k = -1;
  if (valid) {
    s = qs[q];
    t = qt[q];
    k = qk[q];
  }
  // Cases decided without labels; they are uniform within a group.
  bool trivial = !valid || k < 0 || s == t || k == 0 || s >= n || t >= n;
  uint32_t best = kInf32;
  if (!trivial) {
    const unsigned long long sb = off[s], tb = off[t];
    uint32_t ls = (uint32_t)(off[s + 1] - sb), lt = (uint32_t)(off[t + 1] - tb);
    const unsigned long long *A = lab + sb, *B = lab + tb;
    if (ls > lt) {  // stream the shorter list
      const unsigned long long *T = A; A = B; B = T;
      uint32_t tl = ls; ls = lt; lt = tl;
    }
    const uint32_t kk = (uint32_t)k;
    for (uint32_t j = lane; j < ls; j += kQueryGroup) {
      const unsigned long long a = A[j];
      const uint32_t w = label_hub(a), h1 = label_hop(a), d1 = label_dist(a);
      if (h1 > kk) continue;
      for (uint32_t p = lower_bound_hub(B, lt, w); p < lt; ++p) {
        const unsigned long long b = B[p];
        if (label_hub(b) != w || h1 + label_hop(b) > kk) break;  // hops ascend within a hub
        const uint32_t dsum = d1 + label_dist(b);
        best = dsum < best ? dsum : best;
      }
    }
  }
#pragma unroll
  for (int o = kQueryGroup / 2; o > 0; o >>= 1) {
    const uint32_t other = __shfl_xor(best, o, kQueryGroup);
    best = other < best ? other : best;
  }
  if (valid && lane == 0) {
    int32_t r;
    if (k < 0 || s >= n || t >= n) r = INT32_MAX;
    else if (s == t) r = 0;
    else if (k == 0) r = INT32_MAX;
    else r = (best == kInf32) ? INT32_MAX : (int32_t)best;
    out[q] = r;
  }
}

int launch_query(const hsdl_index *idx, size_t nq, const uint32_t *d_s, const uint32_t *d_t,
                 const int8_t *d_k, int32_t *d_out, hipStream_t st) {
  if (nq == 0) return HSDL_OK;
  const uint64_t threads = (uint64_t)nq * kQueryGroup;
  const uint64_t blocks = (threads + 255) / 256;
  if (blocks > 0x7FFFFFFFull) return fail(HSDL_EINVAL, "query batch too large for one launch");
  hipLaunchKernelGGL(k_query, dim3((uint32_t)blocks), dim3(256), 0, st, idx->n,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     (uint64_t)nq, d_s, d_t, d_k, d_out);
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
