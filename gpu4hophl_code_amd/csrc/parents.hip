// parents.hip -- parent vertices of the labels and hop-constrained
// shortest-path extraction.
//
// Parents.  The reference stores, with every label (hub, h, d) of vertex y,
// the vertex before y on the hub's path (Hybrid_Generation_Clean/src/
// label_gen.cu:531-541 keeps the first frontier vertex its scan meets at the
// minimum, which depends on the scan order).  Here the parent is fixed by a
// rule: the smallest frontier vertex of round h that reaches y at distance d.
// The frontier of round h is the set of vertices holding a round-(h-1) label
// of the hub, so after generation the parent is a function of the finished
// label set and the graph:
//   hop 0:  y itself (the self label)
//   hop 1:  the hub (the round-1 frontier is the hub alone)
//   hop h:  min { x in N(y) : (hub, h-1, d - w(x,y)) in L(x) }
// k_derive_parents evaluates it on the uncleaned index (the clean may remove
// the (hub, h-1) label of x); the clean then carries the parents of the
// labels it keeps.  The generation kernels are untouched.
//
// Paths.  hop_constrained_extract_shortest_path
// (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:224-319): merge
// L(s) and L(t) like the distance query, take the first pair in merge order
// (hub, then position in L(s), then position in L(t)) at the smallest dist
// sum with hop sum <= k, emit (s, parent) and (t, parent) for the pair's
// labels where the parent is not the endpoint itself, move the endpoints,
// lower k by one per edge and repeat until s == t or no pair qualifies.
// k_paths runs it with one wavefront per query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.hpp"

namespace hsdl {

// first position in [lo, hi) of the sorted keys L whose key >= key
__device__ inline unsigned long long lower_bound_key(const unsigned long long *L,
                                                     unsigned long long lo, unsigned long long hi,
                                                     unsigned long long key) {
  while (lo < hi) {
    const unsigned long long mid = (lo + hi) >> 1;
    if (L[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// One wavefront per vertex y; lanes take 64 labels of L(y) at a time and
// walk y's id-sorted row together until every lane has found its parent
// (the first match in row order is the smallest id).  A label whose rule
// finds no neighbour (labels not made by K-round generation) counts in
// *missing and gets the vertex itself.
__global__ void __launch_bounds__(256) k_derive_parents(uint32_t n,
                                                        const unsigned long long *__restrict__ off,
                                                        const unsigned long long *__restrict__ lab,
                                                        const unsigned long long *__restrict__ goff,
                                                        const uint2 *__restrict__ adj,
                                                        uint32_t *__restrict__ par,
                                                        unsigned long long *missing) {
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  unsigned long long lost = 0;
  for (uint32_t y = wave; y < n; y += nwaves) {
    const unsigned long long b = off[y], e = off[y + 1];
    const unsigned long long rb = goff[y], re = goff[y + 1];
    for (unsigned long long i0 = b; i0 < e; i0 += kWave) {
      const unsigned long long i = i0 + (unsigned long long)lane;
      const bool valid = i < e;
      const unsigned long long l = valid ? lab[i] : 0ull;
      const uint32_t hub = label_hub(l), hop = label_hop(l), d = label_dist(l);
      uint32_t p = y;
      bool need = false;
      if (valid) {
        if (hop == 1) p = hub;
        else if (hop > 1) need = true;
      }
      for (unsigned long long r = rb; r < re && __ballot(need); ++r) {
        const uint2 ed = adj[r];
        if (need && ed.y <= d) {
          const unsigned long long key = pack_label(hub, hop - 1, d - ed.y);
          const unsigned long long xe = off[ed.x + 1];
          const unsigned long long pos = lower_bound_key(lab, off[ed.x], xe, key);
          if (pos < xe && lab[pos] == key) {
            p = ed.x;
            need = false;
          }
        }
      }
      if (need) ++lost;
      if (valid) par[i] = p;
    }
  }
  lost = wave_sum_u64(lost);
  if (lane == 0 && lost) atomicAdd(missing, lost);
}

// Wave-wide lexicographic minimum of (k1, k2); every lane gets the winner.
__device__ inline void wave_argmin2(unsigned long long &k1, unsigned long long &k2) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long o1 = __shfl_xor(k1, o, kWave);
    const unsigned long long o2 = __shfl_xor(k2, o, kWave);
    if (o1 < k1 || (o1 == k1 && o2 < k2)) {
      k1 = o1;
      k2 = o2;
    }
  }
}

// One wavefront per query.  Each step: lanes take labels j of L(t), find the
// hub's run in L(s) by binary search and scan it in hop order (stopping at
// the first hop sum above k, like the reference's inner loop); the wave then
// takes the minimum of (dist sum, hub, position in L(s), position in L(t)),
// the pair the reference's merge meets first at the minimum.
__global__ void __launch_bounds__(256) k_paths(size_t nq, const unsigned long long *__restrict__ off,
                                               const unsigned long long *__restrict__ lab,
                                               const uint32_t *__restrict__ par,
                                               const uint32_t *__restrict__ qs,
                                               const uint32_t *__restrict__ qt,
                                               const int8_t *__restrict__ qk, uint32_t cap,
                                               uint32_t *__restrict__ edges,
                                               uint32_t *__restrict__ counts) {
  const int lane = lane_id();
  const size_t wave = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) / kWave;
  const size_t nwaves = (gridDim.x * (size_t)blockDim.x) / kWave;
  for (size_t q = wave; q < nq; q += nwaves) {
    uint32_t s = qs[q], t = qt[q];
    int32_t k = qk[q];
    uint32_t cnt = 0;
    uint32_t *out = edges + 2 * (size_t)cap * q;
    // every step that emits lowers k; a step that cannot emit ends the walk
    for (int32_t step = 0; step <= 256 && s != t; ++step) {
      const unsigned long long sb = off[s], se = off[s + 1];
      const unsigned long long tb = off[t], te = off[t + 1];
      unsigned long long k1 = ~0ull, k2 = ~0ull;
      for (unsigned long long j0 = tb; j0 < te; j0 += kWave) {
        const unsigned long long j = j0 + (unsigned long long)lane;
        if (j < te) {
          const unsigned long long lt = lab[j];
          const uint32_t hub = label_hub(lt), hj = label_hop(lt), dj = label_dist(lt);
          if ((int32_t)hj <= k) {
            unsigned long long i = lower_bound_key(lab, sb, se, (unsigned long long)hub << 32);
            for (; i < se; ++i) {
              const unsigned long long ls = lab[i];
              if (label_hub(ls) != hub || (int32_t)(label_hop(ls) + hj) > k) break;
              const unsigned long long c1 =
                  ((unsigned long long)(label_dist(ls) + dj) << 32) | hub;
              const unsigned long long c2 = ((i - sb) << 32) | (j - tb);
              if (c1 < k1 || (c1 == k1 && c2 < k2)) {
                k1 = c1;
                k2 = c2;
              }
            }
          }
        }
      }
      wave_argmin2(k1, k2);
      if (k1 == ~0ull) break;  // nothing connects within k hops
      const uint32_t p1 = par[sb + (k2 >> 32)];
      const uint32_t p2 = par[tb + (k2 & 0xFFFFFFFFull)];
      if (p1 == s && p2 == t) break;
      if (p1 != s) {
        if (lane == 0 && cnt < cap) {
          out[2 * cnt] = s;
          out[2 * cnt + 1] = p1;
        }
        ++cnt;
        s = p1;
        --k;
      }
      if (p2 != t) {
        if (lane == 0 && cnt < cap) {
          out[2 * cnt] = t;
          out[2 * cnt + 1] = p2;
        }
        ++cnt;
        t = p2;
        --k;
      }
    }
    // slots past the path's end read 0
    for (uint32_t e = 2 * min(cnt, cap) + (uint32_t)lane; e < 2 * cap; e += kWave) out[e] = 0;
    if (lane == 0) counts[q] = cnt;
  }
}

// Derives idx->par from the graph (device copy) and the uncleaned sorted
// labels; called by the generation before it returns the index.
int derive_parents(hsdl_index *idx, const hsdl_graph *g, hipStream_t st) {
  if (!g->d_off || !g->d_adj || g->n != idx->n)
    return fail(HSDL_EINVAL, "parents need the index's graph resident on its device");
  int rc = idx->par.alloc(std::max<uint64_t>(idx->total, 1) * 4);
  if (rc) return rc;
  DevBuf missing;
  if ((rc = missing.alloc(8))) return rc;
  HSDL_HIP(hipMemsetAsync(missing.p, 0, 8, st));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, idx->device);
  hipLaunchKernelGGL(k_derive_parents, dim3(cus * 8), dim3(256), 0, st, idx->n,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     static_cast<const unsigned long long *>(g->d_off),
                     static_cast<const uint2 *>(g->d_adj), idx->par.as<uint32_t>(),
                     missing.as<unsigned long long>());
  HSDL_HIP(hipGetLastError());
  unsigned long long lost = 0;
  HSDL_HIP(hipMemcpyAsync(&lost, missing.p, 8, hipMemcpyDeviceToHost, st));
  HSDL_HIP(hipStreamSynchronize(st));
  if (lost) {
    idx->par.release();
    return fail(HSDL_EDEVICE, "parent derivation found labels no neighbour explains");
  }
  idx->has_parents = true;
  return HSDL_OK;
}

}  // namespace hsdl

using namespace hsdl;

extern "C" {

static int export_parents_body(const hsdl_index *idx, uint32_t *parent) {
  if (!idx || !parent) return fail(HSDL_EINVAL, "null argument");
  if (!idx->has_parents) return fail(HSDL_EINVAL, "index has no parents (build with opts.parents = 1)");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  if (idx->total)
    HSDL_HIP(hipMemcpy(parent, idx->par.p, idx->total * 4, hipMemcpyDeviceToHost));
  return HSDL_OK;
}

int hsdl_export_parents(const hsdl_index *idx, uint32_t *parent) {
  return hsdl::guarded([&] { return export_parents_body(idx, parent); });
}

static int index_set_parents_body(hsdl_index *idx, const uint32_t *parent) {
  if (!idx || (!parent && idx->total)) return fail(HSDL_EINVAL, "null argument");
  for (uint64_t i = 0; i < idx->total; ++i)
    if (parent[i] >= idx->n) return fail(HSDL_EINVAL, "parent id out of range");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  DevBuf p;
  if ((rc = p.alloc(std::max<uint64_t>(idx->total, 1) * 4))) return rc;
  if (idx->total) HSDL_HIP(hipMemcpy(p.p, parent, idx->total * 4, hipMemcpyHostToDevice));
  idx->par = std::move(p);
  idx->has_parents = true;
  return HSDL_OK;
}

int hsdl_index_set_parents(hsdl_index *idx, const uint32_t *parent) {
  return hsdl::guarded([&] { return index_set_parents_body(idx, parent); });
}

static int extract_paths_body(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, uint32_t cap, uint32_t *edges, uint32_t *counts) {
  if (!idx) return fail(HSDL_EINVAL, "null index");
  if (n == 0) return HSDL_OK;
  if (!s || !t || !k || !counts || (cap && !edges)) return fail(HSDL_EINVAL, "null path arrays");
  if (!idx->has_parents) return fail(HSDL_EINVAL, "index has no parents (build with opts.parents = 1)");
  for (size_t i = 0; i < n; ++i)
    if (s[i] >= idx->n || t[i] >= idx->n) return fail(HSDL_EINVAL, "query vertex out of range");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  const size_t ebytes = std::max<size_t>((size_t)2 * cap * n * 4, 4);
  DevBuf ds, dt, dk, de, dc;
  if ((rc = ds.alloc(n * 4)) || (rc = dt.alloc(n * 4)) || (rc = dk.alloc(n)) ||
      (rc = de.alloc(ebytes)) || (rc = dc.alloc(n * 4)))
    return rc;
  HSDL_HIP(hipMemcpy(ds.p, s, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dt.p, t, n * 4, hipMemcpyHostToDevice));
  HSDL_HIP(hipMemcpy(dk.p, k, n, hipMemcpyHostToDevice));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, idx->device);
  const size_t blocks = std::min<size_t>((n + 3) / 4, (size_t)cus * 16);
  hipLaunchKernelGGL(k_paths, dim3((uint32_t)blocks), dim3(256), 0, nullptr, n,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     idx->par.as<uint32_t>(), ds.as<uint32_t>(), dt.as<uint32_t>(),
                     dk.as<int8_t>(), cap, de.as<uint32_t>(), dc.as<uint32_t>());
  HSDL_HIP(hipGetLastError());
  HSDL_HIP(hipDeviceSynchronize());
  if (cap) HSDL_HIP(hipMemcpy(edges, de.p, (size_t)2 * cap * n * 4, hipMemcpyDeviceToHost));
  HSDL_HIP(hipMemcpy(counts, dc.p, n * 4, hipMemcpyDeviceToHost));
  return HSDL_OK;
}

int hsdl_extract_paths(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, uint32_t cap, uint32_t *edges, uint32_t *counts) {
  return hsdl::guarded([&] { return extract_paths_body(idx, n, s, t, k, cap, edges, counts); });
}

}  // extern "C"
