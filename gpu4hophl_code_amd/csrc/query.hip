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
// One workgroup of four wavefronts per query, in a persistent loop over the
// batch (workgroup g takes queries g, g+G, g+2G, ...).  Two paths:
//  * merge path (|L(s)| + |L(t)| <= 4094): both lists are staged in LDS as
//    packed labels, then the 256 lanes split the merged hub order evenly
//    (merge path: each lane binary-searches its start on the diagonal and
//    walks ceil(total/256) merged positions).  The shorter list A is the
//    probing side: an element of A, met at B position j (= first B label
//    with hub >= its hub, since A goes first on equal hubs), walks B's run of
//    that hub (<= K+1 labels, hops ascending) while h1 + h2 <= k.
//  * long path (longer pairs): the hubs of the longer list B are staged in
//    LDS -- all of them up to 8188, else every stride-th one (a skip index);
//    lanes take A 256 labels at a time, binary-search each hub in LDS (plus
//    at most log2(stride) global probes) and walk B's run in global memory.
// Memory latency is hidden by software pipelining across the workgroup's
// queries: while query i is joined out of LDS, the labels of query i+1
// (16 per lane in registers), the label offsets of query i+2 and the
// (s,t,k) of query i+3 are in flight.  The prefetches are vector loads by
// lane 0 of each wave (broadcast with readfirstlane), so they never hold up
// the LDS waits of the join.  A min-reduction over the four waves gives the
// answer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "device.hpp"

namespace hsdl {

constexpr uint32_t kQThreads = 256;                 // four wavefronts per query
constexpr uint32_t kQWaves = kQThreads / kWave;
constexpr uint32_t kStageWords = 4096;              // 32 KB of LDS per workgroup
constexpr uint32_t kMergeCap = kStageWords - 2;     // last 16 B hold the wave minima
constexpr uint32_t kLongHubs = 2 * kMergeCap;       // hubs of B staged on the long path
constexpr uint32_t kLoadPerLane = kStageWords / kQThreads;  // 16 labels in flight per lane

template <class Getter>
__device__ inline uint32_t lower_bound_by(Getter get, uint32_t len, uint32_t w) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (get(mid) < w) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t lane_u32(uint32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline uint64_t lane_u64(uint64_t v, int l) {
  return ((uint64_t)lane_u32((uint32_t)(v >> 32), l) << 32) | lane_u32((uint32_t)v, l);
}

// Prefetch loads.  Lanes 0-2 load (s, t, k) of a query and lanes 0-3 its two
// label ranges; the per-lane addresses keep them vector loads (a scalar load
// would share the wait counter of the LDS traffic and stall the join).
__device__ inline uint32_t load_params(uint32_t lane, uint64_t x, const uint32_t *qs,
                                       const uint32_t *qt, const int8_t *qk) {
  uint32_t v = 0;
  if (lane < 2) v = (lane == 0 ? qs : qt)[x];
  else if (lane == 2) v = (uint32_t)(int32_t)qk[x];
  return v;
}
__device__ inline uint64_t load_ranges(uint32_t lane, const unsigned long long *off, uint32_t s,
                                       uint32_t t) {
  uint64_t v = 0;
  if (lane < 4) v = off[((lane & 2) ? t : s) + (lane & 1)];
  return v;
}

// Label formats the kernel reads: the index's 8-byte labels, or its 32-bit
// query copy (hub << (hop_bits + dist_bits) | hop << dist_bits | dist, same
// per-vertex order), which halves the bytes, the registers of the label
// prefetch and the LDS of the staging.
struct Lab64 {
  using T = unsigned long long;
  __device__ uint32_t hub(T l) const { return label_hub(l); }
  __device__ uint32_t hop(T l) const { return label_hop(l); }
  __device__ uint32_t dist(T l) const { return label_dist(l); }
};
struct Lab32 {
  using T = uint32_t;
  uint32_t db, hsh, hmask, dmask;  // dist bits, hop + dist bits, field masks
  __device__ uint32_t hub(T l) const { return l >> hsh; }
  __device__ uint32_t hop(T l) const { return (l >> db) & hmask; }
  __device__ uint32_t dist(T l) const { return l & dmask; }
};

// A query after its (s,t,k) and label offsets are known; uniform over the
// workgroup.  fixed >= 0 is an answer that needs no join.
template <class T>
struct QJob {
  const T *A, *B;
  uint32_t ls, lt, kk;
  int32_t fixed;
};

__device__ inline bool needs_join(uint32_t n, uint32_t s, uint32_t t, int k) {
  return k > 0 && s < n && t < n && s != t;
}

template <class T>
__device__ inline QJob<T> make_job(uint32_t n, const T *lab, uint32_t s, uint32_t t, int k,
                                   uint64_t sb, uint64_t se, uint64_t tb, uint64_t te) {
  QJob<T> j;
  j.A = j.B = lab;
  j.ls = j.lt = 0;
  j.kk = 0;
  if (!needs_join(n, s, t, k)) {
    j.fixed = (k >= 0 && s == t && s < n) ? 0 : INT32_MAX;
    return j;
  }
  uint32_t ls = (uint32_t)(se - sb), lt = (uint32_t)(te - tb);
  const T *A = lab + sb, *B = lab + tb;
  if (ls > lt) {  // A = shorter list, B = longer list
    const T *X = A; A = B; B = X;
    const uint32_t tl = ls; ls = lt; lt = tl;
  }
  j.A = A; j.B = B; j.ls = ls; j.lt = lt; j.kk = (uint32_t)k;
  j.fixed = ls == 0 ? INT32_MAX : -1;  // an empty list shares no hub
  return j;
}

template <class T>
__device__ inline bool merge_job(const QJob<T> &j, uint32_t merge_cap) {
  return j.fixed < 0 && j.ls + j.lt <= merge_cap;
}

// LDS of a workgroup: kStageWords labels, then (32-bit labels) the four
// wave minima, which the 8-byte form keeps in its last 16 bytes instead.
template <class F>
struct QStage {
  static constexpr uint32_t kLabels = kStageWords + (sizeof(typename F::T) == 8 ? 0u : 4u);
  static constexpr uint32_t kMinAt = sizeof(typename F::T) == 8 ? kMergeCap : kStageWords;
  // hubs of B on the long path (u32 words)
  static constexpr uint32_t kHubs = sizeof(typename F::T) == 8 ? kLongHubs : kStageWords;
};

template <class F>
__global__ void __launch_bounds__(kQThreads) k_query(F fmt, uint32_t n,
                                                     const unsigned long long *off,
                                                     const typename F::T *lab, uint64_t nq,
                                                     const uint32_t *qs, const uint32_t *qt,
                                                     const int8_t *qk, int32_t *out,
                                                     uint32_t merge_cap, uint32_t lds_hubs) {
  using LT = typename F::T;
  __shared__ LT stage[QStage<F>::kLabels];
  uint32_t *wave_min = reinterpret_cast<uint32_t *>(stage + QStage<F>::kMinAt);
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const uint64_t G = gridDim.x;
  uint64_t q = blockIdx.x;
  if (q >= nq) return;  // uniform over the workgroup

  // ---- prologue: job of q (synchronous), ranges of q+G, (s,t,k) of q+2G
  uint32_t pv = load_params(lane, q, qs, qt, qk);
  uint32_t s = lane_u32(pv, 0), t = lane_u32(pv, 1);
  int k = (int)lane_u32(pv, 2);
  uint64_t rv = needs_join(n, s, t, k) ? load_ranges(lane, off, s, t) : 0;
  QJob<LT> cur = make_job(n, lab, s, t, k, lane_u64(rv, 0), lane_u64(rv, 1), lane_u64(rv, 2),
                         lane_u64(rv, 3));
  pv = q + G < nq ? load_params(lane, q + G, qs, qt, qk) : 0;
  uint32_t s1 = lane_u32(pv, 0), t1 = lane_u32(pv, 1);
  int k1 = q + G < nq ? (int)lane_u32(pv, 2) : -1;
  if (needs_join(n, s1, t1, k1)) rv = load_ranges(lane, off, s1, t1);
  if (q + 2 * G < nq) pv = load_params(lane, q + 2 * G, qs, qt, qk);
  // labels of q
  LT v[kLoadPerLane];
  if (merge_job(cur, merge_cap)) {
    const uint32_t total = cur.ls + cur.lt;
#pragma unroll
    for (uint32_t u = 0; u < kLoadPerLane; ++u) {
      const uint32_t i = u * kQThreads + tid;
      const uint32_t ic = i < total ? i : total - 1;  // lanes past the end reload the last label
      v[u] = ic < cur.ls ? cur.A[ic] : cur.B[ic - cur.ls];
    }
  }

  for (; q < nq; q += G) {
    // In flight now: labels of q (v), ranges of q+G (rv, lanes 0-3) and
    // (s,t,k) of q+2G (pv, lanes 0-2).
    const bool is_merge = merge_job(cur, merge_cap);
    const uint32_t total = cur.ls + cur.lt;
    const uint32_t ls = cur.ls;
    // hash of A's hub groups: u16 indexes into the staged A, two per word,
    // 2^tbits >= 2|A| slots (<= 4096 since |A| <= 2047), right after A
    const uint32_t tbits = ls ? max(6u, 32u - (uint32_t)__builtin_clz(2 * ls - 1)) : 6u;
    const uint32_t tmask = (1u << tbits) - 1u;
    uint32_t *T = reinterpret_cast<uint32_t *>(stage + ls);
    LT vc[kLoadPerLane];  // this query's labels; v takes the next one's
    if (is_merge) {
#pragma unroll
      for (uint32_t u = 0; u < kLoadPerLane; ++u) {
        vc[u] = v[u];
        const uint32_t i = u * kQThreads + tid;
        if (i < ls) stage[i] = vc[u];
      }
      for (uint32_t j = tid; j < (1u << (tbits - 1)); j += kQThreads) T[j] = kInf32;
    }
    __syncthreads();
    // ---- prefetch: labels of q+G, ranges of q+2G, (s,t,k) of q+3G
    const bool has1 = q + G < nq;
    const QJob<LT> nxt = make_job(n, lab, s1, t1, k1, lane_u64(rv, 0), lane_u64(rv, 1),
                                 lane_u64(rv, 2), lane_u64(rv, 3));
    const bool has2 = q + 2 * G < nq;
    const uint32_t s2 = lane_u32(pv, 0), t2 = lane_u32(pv, 1);
    const int k2 = has2 ? (int)lane_u32(pv, 2) : -1;
    if (has1 && merge_job(nxt, merge_cap)) {
      const uint32_t ntot = nxt.ls + nxt.lt;
#pragma unroll
      for (uint32_t u = 0; u < kLoadPerLane; ++u) {
        const uint32_t i = u * kQThreads + tid;
        const uint32_t ic = i < ntot ? i : ntot - 1;
        v[u] = ic < nxt.ls ? nxt.A[ic] : nxt.B[ic - nxt.ls];
      }
    }
    if (needs_join(n, s2, t2, k2)) rv = load_ranges(lane, off, s2, t2);
    if (q + 3 * G < nq) pv = load_params(lane, q + 3 * G, qs, qt, qk);

    // ---- join of q
    uint32_t best = kInf32;
    const uint32_t kk = cur.kk;
    if (is_merge) {
      const LT *SA = stage;
      // insert the group heads of A (hubs are unique among heads)
      for (uint32_t i = tid; i < ls; i += kQThreads) {
        const uint32_t w = fmt.hub(SA[i]);
        if (i > 0 && fmt.hub(SA[i - 1]) == w) continue;
        uint32_t slot = (w * 0x9E3779B1u) >> (32 - tbits);
        for (;;) {
          uint32_t *wp = &T[slot >> 1];
          const uint32_t sh = (slot & 1u) << 4;
          uint32_t word = *wp;
          bool done = false;
          while (((word >> sh) & 0xFFFFu) == 0xFFFFu) {
            const uint32_t got = atomicCAS(wp, word, (word & ~(0xFFFFu << sh)) | (i << sh));
            if (got == word) { done = true; break; }
            word = got;
          }
          if (done) break;
          slot = (slot + 1) & tmask;
        }
      }
      __syncthreads();
      // each lane probes its labels of B (held in registers) and walks A's
      // run of that hub (<= K+1 labels, hops ascending) while h1 + h2 <= k
#pragma unroll
      for (uint32_t u = 0; u < kLoadPerLane; ++u) {
        const uint32_t i = u * kQThreads + tid;
        if (i < ls || i >= total) continue;
        const LT b = vc[u];
        const uint32_t w = fmt.hub(b), h2 = fmt.hop(b), d2 = fmt.dist(b);
        if (h2 > kk) continue;
        uint32_t slot = (w * 0x9E3779B1u) >> (32 - tbits);
        for (;;) {
          const uint32_t e = (T[slot >> 1] >> ((slot & 1u) << 4)) & 0xFFFFu;
          if (e == 0xFFFFu) break;
          if (fmt.hub(SA[e]) == w) {
            for (uint32_t p = e; p < ls; ++p) {
              const LT a = SA[p];
              if (fmt.hub(a) != w || fmt.hop(a) + h2 > kk) break;
              const uint32_t dsum = fmt.dist(a) + d2;
              best = dsum < best ? dsum : best;
            }
            break;
          }
          slot = (slot + 1) & tmask;
        }
      }
    } else if (cur.fixed < 0) {
      // long path: LDS skip index of B's hubs (stride 1 up to lds_hubs)
      const LT *A = cur.A, *B = cur.B;
      const uint32_t ls = cur.ls, lt = cur.lt;
      uint32_t *hubs = reinterpret_cast<uint32_t *>(stage);
      const uint32_t stride = (lt + lds_hubs - 1) / lds_hubs;
      const uint32_t nsm = (lt + stride - 1) / stride;
      for (uint32_t i = tid; i < nsm; i += kQThreads) hubs[i] = fmt.hub(B[(uint64_t)i * stride]);
      __syncthreads();
      for (uint32_t j = tid; j < ls; j += kQThreads) {
        const LT a = A[j];
        const uint32_t w = fmt.hub(a), h1 = fmt.hop(a), d1 = fmt.dist(a);
        if (h1 > kk) continue;
        // first sample >= w, then the first position >= w inside its block
        const uint32_t si = lower_bound_by([&](uint32_t x) { return hubs[x]; }, nsm, w);
        const uint32_t lo = si == 0 ? 0 : (si - 1) * stride + 1;
        const uint32_t hi = si < nsm ? si * stride : lt;
        uint32_t p = lo + lower_bound_by([&](uint32_t x) { return fmt.hub(B[lo + x]); }, hi - lo, w);
        for (; p < lt; ++p) {
          const LT b = B[p];
          if (fmt.hub(b) != w || h1 + fmt.hop(b) > kk) break;  // hops ascend within a hub
          const uint32_t dsum = d1 + fmt.dist(b);
          best = dsum < best ? dsum : best;
        }
      }
    }
    best = wave_min_u32(best);
    if (lane == 0) wave_min[wv] = best;
    __syncthreads();
    if (tid == 0) {
      uint32_t m = wave_min[0];
      for (uint32_t x = 1; x < kQWaves; ++x) m = wave_min[x] < m ? wave_min[x] : m;
      out[q] = cur.fixed >= 0 ? cur.fixed : (m == kInf32 ? INT32_MAX : (int32_t)m);
    }
    // advance the pipeline
    cur = nxt;
    s1 = s2; t1 = t2; k1 = k2;
  }
}

// Workgroups resident at once on a device (the persistent grid), cached per
// kernel form.
template <class F>
static uint32_t query_grid(int dev) {
  static std::mutex mu;
  static int cached_dev = -1;
  static uint32_t cached = 0;
  std::lock_guard<std::mutex> g(mu);
  if (dev == cached_dev && cached) return cached;
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_query<F>, kQThreads, 0) !=
          hipSuccess ||
      per_cu <= 0)
    per_cu = 4;
  cached = (uint32_t)(cus * per_cu);
  cached_dev = dev;
  return cached;
}

__global__ void __launch_bounds__(256) k_pack_query32(const unsigned long long *lab, uint64_t total,
                                                      uint32_t *out, uint32_t hsh, uint32_t db) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long l = lab[i];
    out[i] = (label_hub(l) << hsh) | (label_hop(l) << db) | label_dist(l);
  }
}

static uint32_t bits_for(uint64_t x) {  // bits that hold 0..x, at least 1
  uint32_t b = 1;
  while (b < 64 && (x >> b)) ++b;
  return b;
}

// Makes the index's 32-bit query copy on stream st when the fields fit 32
// bits (hub bits of n - 1, hop bits of K, dist bits of the index's distance
// bound), once per label set, and waits for it; tuning knob HSDL_QUERY32=0
// keeps the 8-byte labels.
static std::mutex g_q32_mu;
static int ensure_query32(const hsdl_index *idx, hipStream_t st) {
  std::lock_guard<std::mutex> g(g_q32_mu);
  if (idx->q32.load(std::memory_order_acquire)) return HSDL_OK;
  const uint32_t hb = bits_for(idx->n ? idx->n - 1u : 0u), pb = bits_for(idx->K),
                 db = bits_for(idx->dist_max);
  if (hb + pb + db > 32) {
    idx->q32.store(2, std::memory_order_release);
    return HSDL_OK;
  }
  // the copy is an optimisation: without room for it, queries read the
  // 8-byte labels
  if (idx->lab32.alloc(std::max<uint64_t>(idx->total, 1) * 4) != HSDL_OK) {
    idx->q32.store(2, std::memory_order_release);
    return HSDL_OK;
  }
  const uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (idx->total + 255) / 256 + 1);
  hipLaunchKernelGGL(k_pack_query32, dim3(grid), dim3(256), 0, st,
                     idx->lab.as<unsigned long long>(), idx->total, idx->lab32.as<uint32_t>(),
                     pb + db, db);
  HSDL_HIP(hipGetLastError());
  // queries on other streams may read the copy as soon as it is marked made
  HSDL_HIP(hipStreamSynchronize(st));
  idx->q32_hop_bits = pb;
  idx->q32_dist_bits = db;
  idx->q32.store(1, std::memory_order_release);
  return HSDL_OK;
}

int prepare_query(const hsdl_index *idx) {
  if (!env_u32("HSDL_QUERY32", 1, 0, 1) || idx->q32.load(std::memory_order_acquire))
    return HSDL_OK;
  return ensure_query32(idx, nullptr);
}

int launch_query(const hsdl_index *idx, size_t nq, const uint32_t *d_s, const uint32_t *d_t,
                 const int8_t *d_k, int32_t *d_out, hipStream_t st, bool may_alloc) {
  if (nq == 0) return HSDL_OK;
  // tuning knobs: labels staged for the merge path (<= 4094; 0 sends every
  // query down the long path) and LDS skip-index capacity of the long path
  const uint32_t merge_cap = env_u32("HSDL_QUERY_MERGE_CAP", kMergeCap, 0, kMergeCap);
  const uint32_t lds_hubs = env_u32("HSDL_QUERY_LDS_HUBS", kLongHubs, 1, kLongHubs);
  const bool want32 = env_u32("HSDL_QUERY32", 1, 0, 1) != 0;
  if (want32 && may_alloc && !idx->q32.load(std::memory_order_acquire)) {
    // a stream under graph capture takes no allocation or synchronisation
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HSDL_HIP(hipStreamIsCapturing(st, &cs));
    if (cs == hipStreamCaptureStatusNone) {
      const int rc = ensure_query32(idx, st);
      if (rc) return rc;
    }
  }
  if (want32 && idx->q32.load(std::memory_order_acquire) == 1) {
    const uint32_t db = idx->q32_dist_bits, pb = idx->q32_hop_bits;
    const Lab32 fmt{db, pb + db, (1u << pb) - 1u, (1u << db) - 1u};
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nq, query_grid<Lab32>(idx->device));
    hipLaunchKernelGGL(k_query<Lab32>, dim3(grid), dim3(kQThreads), 0, st, fmt, idx->n,
                       idx->off.as<unsigned long long>(), idx->lab32.as<uint32_t>(), (uint64_t)nq,
                       d_s, d_t, d_k, d_out, merge_cap,
                       std::min(lds_hubs, QStage<Lab32>::kHubs));
  } else {
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nq, query_grid<Lab64>(idx->device));
    hipLaunchKernelGGL(k_query<Lab64>, dim3(grid), dim3(kQThreads), 0, st, Lab64{}, idx->n,
                       idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                       (uint64_t)nq, d_s, d_t, d_k, d_out, merge_cap,
                       std::min(lds_hubs, QStage<Lab64>::kHubs));
  }
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

static int query_body(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
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

int hsdl_query(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
               const int8_t *k, int32_t *out) {
  return hsdl::guarded([&] { return query_body(idx, n, s, t, k, out); });
}

}  // extern "C"

// ------------------------------------------------------- no-allocation query
// hsdl_query_noalloc stages the requests and answers in `scratch`, a
// file-scope device array that every GPU the library runs on holds until the
// process ends, instead of allocating device buffers per call: for callers
// that cannot allocate.  Layout for n <= kNoAllocMaxQueries requests:
//   [s: u32 x max][t: u32 x max][out: i32 x max][k: i8 x max]
// Nothing in it must start at zero (the query kernel writes every answer).
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU race and must be serialised by the caller.
namespace {
constexpr size_t kNoAllocMaxQueries = (size_t)1 << 20;
constexpr size_t kNoAllocBytes = kNoAllocMaxQueries * 13;
static_assert(HSDL_NOALLOC_MAX_QUERIES == kNoAllocMaxQueries, "header and library disagree");
constexpr int kNoAllocDevices = 64;
}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
// Device address of `scratch` per device: the host-side name is only a
// stand-in, so host code uses these (looked up once per device, with that
// device current).
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocDevices] = {};

int scratch_for_device(int dev, uint8_t **out) {
  if (dev < 0 || dev >= kNoAllocDevices) return fail(HSDL_ENODEV, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<uint8_t *>(g_scratch_ptr[dev]);
  return HSDL_OK;
}
}  // namespace

extern "C" {

static int query_noalloc_body(const hsdl_index *idx, size_t n, const uint32_t *s,
                              const uint32_t *t, const int8_t *k, int32_t *out) {
  if (!idx) return fail(HSDL_EINVAL, "null index");
  if (n > kNoAllocMaxQueries)
    return fail(HSDL_EINVAL, "hsdl_query_noalloc takes at most HSDL_NOALLOC_MAX_QUERIES (2^20) "
                             "queries per call");
  if (n == 0) return HSDL_OK;
  if (!s || !t || !k || !out) return fail(HSDL_EINVAL, "null query arrays");
  for (size_t i = 0; i < n; ++i)
    if (s[i] >= idx->n || t[i] >= idx->n) return fail(HSDL_EINVAL, "query vertex out of range");
  int dev = 0;
  int rc = use_device(idx->device, &dev);
  if (rc) return rc;
  uint8_t *base = nullptr;
  if ((rc = scratch_for_device(dev, &base))) return rc;
  uint32_t *ds = reinterpret_cast<uint32_t *>(base);
  uint32_t *dt = ds + kNoAllocMaxQueries;
  int32_t *dout = reinterpret_cast<int32_t *>(dt + kNoAllocMaxQueries);
  int8_t *dk = reinterpret_cast<int8_t *>(dout + kNoAllocMaxQueries);
  HSDL_HIP(hipMemcpy(ds, s, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dt, t, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dk, k, n, hipMemcpyHostToDevice));
  if ((rc = launch_query(idx, n, ds, dt, dk, dout, nullptr, false))) return rc;
  HSDL_HIP(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
  return HSDL_OK;
}

int hsdl_query_noalloc(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, int32_t *out) {
  return hsdl::guarded([&] { return query_noalloc_body(idx, n, s, t, k, out); });
}

}  // extern "C"
