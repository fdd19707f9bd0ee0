// clean.hip -- canonical repair of a sorted label index on gfx950.
//
// Predicate of the reference's CPU clean, hop_constrained_clean_L
// (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:422-511):
// walk L(v) in (hub, hop, dist) order with kept := {}; label (u,h,d) is kept iff
//   min{ d1 + d2 : (w,h1,d1) in L(u), (w,h2,d2) in kept, h1 + h2 <= h } > d.
// The reference runs it with a thread pool, so L(u) may be read before or
// after u itself was cleaned; these kernels read the input L(u), and the
// oracle's tests require both readings to give the same output.
//
// Two passes, exact in every case:
//  1. k_clean_all: the predicate with `kept` replaced by every earlier label
//     of v.  That set contains the sequential kept set, so its minima are
//     never larger: when it removes nothing from L(v), the sequential walk
//     keeps every label of v as well.  Pass 1 has no order dependence, so one
//     wavefront takes all hub groups of v at once: L(v)'s hubs sit in LDS,
//     64-label windows of L(v) yield up to 64 group heads (ballot on hub
//     changes), the heads' lists L(u) are load-balanced over the lanes by a
//     register prefix sum, and every element (w,h1,d1) of L(u) that could
//     still remove a group label (h1 + 1 <= the group's largest hop and
//     d1 <= its largest dist; hop 0 belongs to the self label only) binary-
//     searches w in the LDS hubs (in global memory for lists longer than the
//     LDS copy) and tests its pairs against the group's labels.
//  2. k_clean_seq: vertices where pass 1 would remove something are walked
//     group by group in order with keep bits in LDS, exactly as the
//     reference does.
// The reference's GPU clean (Hybrid_Generation_Clean/src/label_clean.cu:73-164)
// uses a dense per-vertex hash of all labels instead; the kept-set form is
// the normative one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.hpp"

namespace hsdl {

constexpr uint32_t kCleanWaves = 4;            // waves per workgroup
constexpr uint32_t kSeqBitsPerWave = 65536;    // max labels per vertex (pass 2)
constexpr uint32_t kSeqWordsPerWave = kSeqBitsPerWave / 32;
constexpr uint32_t kAllHubsPerWave = 2048;     // L(v) hubs staged in LDS (pass 1)

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lower bound of hub w in lab[lo, hi)
__device__ inline uint32_t lb_hub(const unsigned long long *lab, uint32_t lo, uint32_t hi,
                                  uint32_t w) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (label_hub(lab[mid]) < w) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t lb_u32(const uint32_t *h, uint32_t lo, uint32_t hi, uint32_t w) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (h[mid] < w) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- pass 1
// L(v)'s hubs are found through an LDS skip index: samp[i] = hub of label
// i*stride (stride 1 = every label for lists up to kAllHubsPerWave).  A lower
// bound searches LDS for the block and then at most log2(stride) labels in
// global memory.
__device__ inline uint32_t lb_skip(const uint32_t *samp, uint32_t stride,
                                   const unsigned long long *Lv, uint32_t end, uint32_t w) {
  const uint32_t ns = (end + stride - 1) / stride;  // samples below `end`
  const uint32_t si = lb_u32(samp, 0, ns, w);
  const uint32_t lo = si == 0 ? 0 : (si - 1) * stride + 1;
  const uint32_t hi = si < ns ? si * stride : end;
  return lb_hub(Lv, lo, hi, w);
}

// Returns true when the all-earlier predicate removes some label of v.
__device__ bool any_removal(const unsigned long long *off, const unsigned long long *lab,
                            const unsigned long long *Lv, uint32_t len, const uint32_t *samp,
                            uint32_t stride) {
  const int lane = lane_id();
  bool removed = false;
  for (uint32_t p0 = 0; p0 < len && !removed; p0 += kWave) {
    // group heads in this window of L(v)
    const uint32_t p = p0 + lane;
    const bool head = p < len && (p == 0 || label_hub(Lv[p]) != label_hub(Lv[p - 1]));
    uint32_t u = 0, glen = 0, maxh = 0, maxd = 0;
    unsigned long long ub = 0;
    uint32_t ulen = 0;
    if (head) {
      u = label_hub(Lv[p]);
      while (p + glen < len && label_hub(Lv[p + glen]) == u) {
        const unsigned long long g = Lv[p + glen];
        maxh = max(maxh, label_hop(g));
        maxd = max(maxd, label_dist(g));
        ++glen;
      }
      ub = off[u];
      ulen = (uint32_t)(off[u + 1] - ub);
    }
    const uint32_t incl = wave_incl_scan_u32(ulen);
    const uint32_t excl = incl - ulen;
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    bool hit = false;
    for (uint32_t jb = 0; jb < total; jb += kWave) {
      const uint32_t j = jb + lane;
      int k = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const uint32_t e = __shfl(excl, k + step, kWave);
        if (e <= j) k += step;
      }
      const uint32_t ku = __shfl(u, k, kWave);
      const uint32_t kglen = __shfl(glen, k, kWave);
      const uint32_t kmaxh = __shfl(maxh, k, kWave);
      const uint32_t kmaxd = __shfl(maxd, k, kWave);
      const uint32_t kex = __shfl(excl, k, kWave);
      const unsigned long long kub = __shfl(ub, k, kWave);
      const uint32_t gs = p0 + (uint32_t)k;  // group start in L(v)
      if (j >= total) continue;
      const unsigned long long l = lab[kub + (j - kex)];
      const uint32_t w = label_hub(l), h1 = label_hop(l), d1 = label_dist(l);
      if (w < ku) {
        // a pair through w has hop >= h1 + 1 (w != v) and dist >= d1
        if (h1 + 1 > kmaxh || d1 > kmaxd) continue;
        for (uint32_t q = lb_skip(samp, stride, Lv, gs, w); q < gs; ++q) {
          const unsigned long long e2 = Lv[q];
          if (label_hub(e2) != w) break;
          const uint32_t x = h1 + label_hop(e2), sd = d1 + label_dist(e2);
          for (uint32_t t = 0; t < kglen; ++t) {
            const unsigned long long g = Lv[gs + t];
            if (x <= label_hop(g) && sd <= label_dist(g)) hit = true;
          }
        }
      } else if (w == ku) {
        // through hub u: earlier labels of the group against later ones
        for (uint32_t t1 = 0; t1 < kglen; ++t1) {
          const unsigned long long g1 = Lv[gs + t1];
          for (uint32_t t2 = t1 + 1; t2 < kglen; ++t2) {
            const unsigned long long g2 = Lv[gs + t2];
            if (h1 + label_hop(g1) <= label_hop(g2) && d1 + label_dist(g1) <= label_dist(g2))
              hit = true;
          }
        }
      }
    }
    removed = __ballot(hit) != 0;
  }
  return removed;
}

__global__ void __launch_bounds__(256) k_clean_all(uint32_t n, const unsigned long long *off,
                                                   const unsigned long long *lab, uint8_t *keep,
                                                   uint32_t *new_len, uint32_t *redo,
                                                   uint32_t *redo_cnt, uint32_t lds_hubs) {
  __shared__ uint32_t hubs_all[kCleanWaves * kAllHubsPerWave];
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t *hubs = hubs_all + wv * kAllHubsPerWave;
  const uint32_t wave = blockIdx.x * kCleanWaves + wv;
  const uint32_t nwaves = gridDim.x * kCleanWaves;
  for (uint32_t v = wave; v < n; v += nwaves) {
    const unsigned long long vb = off[v];
    const uint32_t len = (uint32_t)(off[v + 1] - vb);
    const unsigned long long *Lv = lab + vb;
    if (len == 0) {
      if (lane == 0) new_len[v] = 0;
      continue;
    }
    const uint32_t stride = (len + lds_hubs - 1) / lds_hubs;
    for (uint32_t i = lane; i * stride < len; i += kWave) hubs[i] = label_hub(Lv[(uint64_t)i * stride]);
    wave_lds_sync();
    const bool removed = any_removal(off, lab, Lv, len, hubs, stride);
    wave_lds_sync();
    if (removed) {
      if (lane == 0) redo[atomicAdd(redo_cnt, 1u)] = v;
    } else {
      for (uint32_t i = lane; i < len; i += kWave) keep[vb + i] = 1;
      if (lane == 0) new_len[v] = len;
    }
  }
}

// ---------------------------------------------------------------- pass 2
// The reference's sequential walk for the vertices in `list`.  For a group of
// hub u the 64 lanes stream L(u); each lane binary-searches its label's hub
// w < u in the decided prefix of L(v) and folds the kept matches into per-hop
// minima m[x] (x = h1 + h2).  A wave min-reduction and a prefix-min over x
// give the best pair for every hop budget; the group's labels are then
// decided in hop order, adding the pairs through hub u itself.
__global__ void __launch_bounds__(256) k_clean_seq(uint32_t K, const unsigned long long *off,
                                                   const unsigned long long *lab, uint8_t *keep,
                                                   uint32_t *new_len, const uint32_t *list,
                                                   const uint32_t *list_cnt, uint32_t *bad) {
  __shared__ uint32_t bits_all[kCleanWaves * kSeqWordsPerWave];
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t *bits = bits_all + wv * kSeqWordsPerWave;
  const uint32_t K1 = K + 1;
  const uint32_t wave = blockIdx.x * kCleanWaves + wv;
  const uint32_t nwaves = gridDim.x * kCleanWaves;
  const uint32_t cnt = *list_cnt;
  for (uint32_t li = wave; li < cnt; li += nwaves) {
    const uint32_t v = list[li];
    const unsigned long long vb = off[v];
    const uint32_t len = (uint32_t)(off[v + 1] - vb);
    const unsigned long long *Lv = lab + vb;
    if (len > kSeqBitsPerWave) {
      if (lane == 0) atomicOr(bad, 1u);
      continue;
    }
    for (uint32_t i = lane; i < (len + 31) / 32; i += kWave) bits[i] = 0;
    wave_lds_sync();
    uint32_t kept_total = 0;
    uint32_t gb = 0;
    while (gb < len) {
      const uint32_t u = label_hub(Lv[gb]);
      uint32_t ge = gb + 1;
      while (ge < len && label_hub(Lv[ge]) == u) ++ge;  // <= K+1 labels
      uint32_t m[kMaxK + 1];
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x) m[x] = kInf32;
      const unsigned long long ub = off[u];
      const uint32_t ulen = (uint32_t)(off[u + 1] - ub);
      const unsigned long long *Lu = lab + ub;
      if (gb > 0) {
        for (uint32_t j = lane; j < ulen; j += kWave) {
          const unsigned long long l = Lu[j];
          const uint32_t w = label_hub(l);
          if (w >= u) continue;
          const uint32_t h1 = label_hop(l), d1 = label_dist(l);
          for (uint32_t p = lb_hub(Lv, 0, gb, w); p < gb; ++p) {
            const unsigned long long q = Lv[p];
            if (label_hub(q) != w) break;
            if (!((bits[p >> 5] >> (p & 31)) & 1u)) continue;
            const uint32_t x = h1 + label_hop(q);
            if (x <= K) {
              const uint32_t s = d1 + label_dist(q);
#pragma unroll
              for (uint32_t xx = 0; xx <= kMaxK; ++xx)
                if (xx == x) m[xx] = min(m[xx], s);
            }
          }
        }
      }
      uint32_t M[kMaxK + 1];
      uint32_t run = kInf32;
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x) {
        const uint32_t r = (x < K1) ? wave_min_u32(m[x]) : kInf32;
        run = min(run, r);
        M[x] = run;
      }
      // decide the group's labels in order (all lanes compute the same)
      uint32_t newbits = 0;  // relative to gb; groups hold <= 16 labels
      for (uint32_t i = gb; i < ge; ++i) {
        const unsigned long long l = Lv[i];
        const uint32_t h = label_hop(l), d = label_dist(l);
        uint32_t best = kInf32;
#pragma unroll
        for (uint32_t x = 0; x <= kMaxK; ++x)
          if (x == h) best = M[x];
        // pairs through hub u: L(u)'s hub-u labels x kept labels of this group
        for (uint32_t k2 = gb; k2 < i; ++k2) {
          if (!((newbits >> (k2 - gb)) & 1u)) continue;
          const unsigned long long q = Lv[k2];
          const uint32_t h2 = label_hop(q), d2 = label_dist(q);
          for (uint32_t t = ulen; t > 0; --t) {
            const unsigned long long r = Lu[t - 1];
            if (label_hub(r) != u) break;
            if (label_hop(r) + h2 <= h) best = min(best, label_dist(r) + d2);
          }
        }
        if (best > d) newbits |= 1u << (i - gb);
      }
      if (lane == 0) {
        for (uint32_t i = gb; i < ge; ++i)
          if ((newbits >> (i - gb)) & 1u) atomicOr(&bits[i >> 5], 1u << (i & 31));
      }
      kept_total += __popc(newbits);
      wave_lds_sync();
      gb = ge;
    }
    for (uint32_t i = lane; i < len; i += kWave) keep[vb + i] = (bits[i >> 5] >> (i & 31)) & 1u;
    if (lane == 0) new_len[v] = kept_total;
    wave_lds_sync();
  }
}

__global__ void __launch_bounds__(256) k_compact(uint32_t n, const unsigned long long *off,
                                                 const unsigned long long *lab,
                                                 const uint8_t *keep,
                                                 const unsigned long long *noff,
                                                 unsigned long long *nlab) {
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t v = wave; v < n; v += nwaves) {
    const unsigned long long b = off[v], e = off[v + 1];
    unsigned long long dst = noff[v];
    for (unsigned long long i = b; i < e; i += kWave) {
      const bool valid = i + lane < e;
      const bool k = valid && keep[i + lane];
      const unsigned long long mk = __ballot(k);
      if (k) nlab[dst + __popcll(mk & lt)] = lab[i + lane];
      dst += __popcll(mk);
    }
  }
}

int clean_index(hsdl_index *idx) {
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  const uint32_t n = idx->n;
  const uint64_t t0 = now_ns();
  hipStream_t st;
  HSDL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct G {
    hipStream_t s;
    ~G() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf keep, new_len, noff, nlab, redo, ctr;
  if ((rc = keep.alloc(std::max<uint64_t>(idx->total, 1))) || (rc = new_len.alloc(n * 4ull)) ||
      (rc = noff.alloc((n + 1ull) * 8)) || (rc = redo.alloc(n * 4ull + 4)) ||
      (rc = ctr.alloc(8)))
    return rc;
  HSDL_HIP(hipMemsetAsync(ctr.p, 0, 8, st));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, idx->device);
  const int grid = std::max(1, cus * 8);
  uint32_t *cnt = ctr.as<uint32_t>();
  hipLaunchKernelGGL(k_clean_all, dim3(grid), dim3(kCleanWaves * kWave), 0, st, n,
                     idx->off.as<unsigned long long>(),
                     idx->lab.as<unsigned long long>(), keep.as<uint8_t>(),
                     new_len.as<uint32_t>(), redo.as<uint32_t>(), cnt,
                     env_u32("HSDL_CLEAN_LDS_HUBS", kAllHubsPerWave, 1, kAllHubsPerWave));
  hipLaunchKernelGGL(k_clean_seq, dim3(grid), dim3(kCleanWaves * kWave), 0, st, (uint32_t)idx->K,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     keep.as<uint8_t>(), new_len.as<uint32_t>(), redo.as<uint32_t>(), cnt,
                     cnt + 1);
  HSDL_HIP(hipGetLastError());
  uint32_t hc[2] = {0, 0};
  HSDL_HIP(hipMemcpyAsync(hc, ctr.p, 8, hipMemcpyDeviceToHost, st));
  HSDL_HIP(hipStreamSynchronize(st));
  if (hc[1]) return fail(HSDL_ERANGE, "a vertex holds more than 65536 labels; clean not supported");
  if ((rc = scan_counts_u32(new_len.as<uint32_t>(), noff.as<uint64_t>(), n, st))) return rc;
  uint64_t total = 0;
  HSDL_HIP(hipMemcpy(&total, noff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost));
  if ((rc = nlab.alloc(std::max<uint64_t>(total, 1) * 8))) return rc;
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, st, n, idx->off.as<unsigned long long>(),
                     idx->lab.as<unsigned long long>(), keep.as<uint8_t>(),
                     noff.as<unsigned long long>(), nlab.as<unsigned long long>());
  HSDL_HIP(hipGetLastError());
  HSDL_HIP(hipStreamSynchronize(st));
  idx->off = std::move(noff);
  idx->lab = std::move(nlab);
  idx->total = total;
  idx->cleaned = true;
  idx->stats.labels_cleaned = total;
  idx->stats.clean_passes = hc[0] ? 2 : 1;
  idx->stats.clean_seq_vertices = hc[0];
  idx->stats.t_clean_s = (now_ns() - t0) * 1e-9;
  return HSDL_OK;
}

}  // namespace hsdl

extern "C" int hsdl_clean(hsdl_index *idx) {
  if (!idx) return hsdl::fail(HSDL_EINVAL, "null index");
  return hsdl::clean_index(idx);
}
