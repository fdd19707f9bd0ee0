// groups.hip -- vertex groups for group builds.
//
// The reference's hybrid driver generates labels one vertex group at a time
// (Hybrid_Generation_Clean/src/test.cu:361-449); its groups come from
// include/vgroup/: label propagation with a community size cap
// (CDLP_group.cuh:20-107 over CDLP/GPU_Community_Detection.cuh:104-210), a
// seeded shuffle (Random_group.cuh:10-33), or ids in order.  Any grouping
// gives the same cleaned labels; a grouping only decides how many redundant
// labels the clean removes.
//
// CDLP here is deterministic (the reference's community-size test races
// between threads): every round, vertex v takes the label most frequent
// among its neighbours' labels and its own, which counts once (ties: the
// smallest), skipping labels whose community already held max_size vertices
// after the previous round, unless it is v's own; vertices without
// neighbours keep their own id.  A round is
// a gather of the neighbours' labels, a segmented radix sort (rocPRIM) and a
// per-vertex run scan.  Communities are then cut into chunks of max_size in
// id order, merged greedily (the largest group with the largest one that
// still fits beside it) and ordered by total degree, largest first, as the
// reference's driver does.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <set>
#include <utility>
#include <vector>

#include "device.hpp"

namespace hsdl {

// prop[e] = label of edge e's neighbour
__global__ void k_lp_gather(uint64_t nnz, const uint32_t *col, const uint32_t *label,
                            uint32_t *prop) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (uint64_t)gridDim.x * blockDim.x)
    prop[e] = label[col[e]];
}

// One thread per vertex over its sorted neighbour labels: the most frequent
// label whose community (size after the last round) has room, ties to the
// smallest; the vertex's own label counts one vote more and always has room
// (so two vertices joined by an edge settle on one label instead of trading
// labels every round).  *changed counts the vertices whose label changes
// (one atomic per wave).
__global__ void k_lp_mode(uint32_t n, const unsigned long long *off, const uint32_t *sorted,
                          const uint32_t *label, const uint32_t *size, uint32_t max_size,
                          uint32_t *next, uint32_t *changed) {
  const uint32_t stride = gridDim.x * blockDim.x;
  // whole waves iterate together (the ballot sees every lane)
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1u)); base < n;
       base += stride) {
    const uint32_t v = base + (threadIdx.x & (kWave - 1u));
    bool moved = false;
    if (v < n) {
      const unsigned long long e0 = off[v], e1 = off[v + 1];
      const uint32_t own = label[v];
      uint32_t best = kInf32, best_cnt = 0;
      bool own_seen = false;
      for (unsigned long long e = e0; e < e1;) {
        const uint32_t l = sorted[e];
        unsigned long long r = e + 1;
        while (r < e1 && sorted[r] == l) ++r;
        const uint32_t c = (uint32_t)(r - e) + (l == own ? 1u : 0u);
        own_seen |= l == own;
        // runs come in ascending label order: a tie keeps the smaller
        if (c > best_cnt && (l == own || size[l] < max_size)) {
          best = l;
          best_cnt = c;
        }
        e = r;
      }
      if (!own_seen && (1u > best_cnt || (1u == best_cnt && own < best))) best = own;
      next[v] = best;
      moved = best != own;
    }
    const unsigned long long m = __ballot(moved);
    if ((threadIdx.x & (kWave - 1u)) == 0 && m) atomicAdd(changed, (uint32_t)__popcll(m));
  }
}

__global__ void k_lp_count(uint32_t n, const uint32_t *label, uint32_t *size) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    atomicAdd(&size[label[v]], 1u);
}

__global__ void k_lp_init(uint32_t n, uint32_t *label, uint32_t *size) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    label[v] = v;
    size[v] = 1;
  }
}

// Community labels after `rounds` rounds of capped label propagation, or
// fewer once a round changes no label.
static int cdlp_labels(const hsdl_graph *g, uint32_t max_size, uint32_t rounds,
                       std::vector<uint32_t> &label) {
  const uint32_t n = g->n;
  const uint64_t nnz = g->off[n];
  hipStream_t st = nullptr;
  HSDL_HIP(hipStreamCreate(&st));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf d_off, d_col, d_lab, d_next, d_size, d_prop, d_sorted, d_tmp, d_changed;
  int rc;
  const uint64_t e = std::max<uint64_t>(nnz, 1);
  if ((rc = d_off.alloc((n + 1ull) * 8)) || (rc = d_col.alloc(e * 4)) ||
      (rc = d_lab.alloc(n * 4ull)) || (rc = d_next.alloc(n * 4ull)) ||
      (rc = d_size.alloc(n * 4ull)) || (rc = d_prop.alloc(e * 4)) ||
      (rc = d_sorted.alloc(e * 4)) || (rc = d_changed.alloc(4)))
    return rc;
  HSDL_HIP(hipMemcpyAsync(d_off.p, g->off.data(), (n + 1ull) * 8, hipMemcpyHostToDevice, st));
  if (nnz)
    HSDL_HIP(hipMemcpyAsync(d_col.p, g->col.data(), nnz * 4, hipMemcpyHostToDevice, st));
  const unsigned long long *off = d_off.as<unsigned long long>();
  size_t tmp_bytes = 0;
  HSDL_HIP(rocprim::segmented_radix_sort_keys(nullptr, tmp_bytes, d_prop.as<uint32_t>(),
                                              d_sorted.as<uint32_t>(), (size_t)nnz, n, off,
                                              off + 1, 0, 32, st));
  if ((rc = d_tmp.alloc(std::max<size_t>(tmp_bytes, 16)))) return rc;
  const int grid = 1024;
  hipLaunchKernelGGL(k_lp_init, dim3(grid), dim3(256), 0, st, n, d_lab.as<uint32_t>(),
                     d_size.as<uint32_t>());
  for (uint32_t r = 0; r < rounds && nnz; ++r) {
    hipLaunchKernelGGL(k_lp_gather, dim3(grid), dim3(256), 0, st, nnz, d_col.as<uint32_t>(),
                       d_lab.as<uint32_t>(), d_prop.as<uint32_t>());
    HSDL_HIP(rocprim::segmented_radix_sort_keys(d_tmp.p, tmp_bytes, d_prop.as<uint32_t>(),
                                                d_sorted.as<uint32_t>(), (size_t)nnz, n, off,
                                                off + 1, 0, 32, st));
    HSDL_HIP(hipMemsetAsync(d_changed.p, 0, 4, st));
    hipLaunchKernelGGL(k_lp_mode, dim3(grid), dim3(256), 0, st, n, off,
                       d_sorted.as<uint32_t>(), d_lab.as<uint32_t>(), d_size.as<uint32_t>(),
                       max_size, d_next.as<uint32_t>(), d_changed.as<uint32_t>());
    HSDL_HIP(hipMemsetAsync(d_size.p, 0, n * 4ull, st));
    hipLaunchKernelGGL(k_lp_count, dim3(grid), dim3(256), 0, st, n, d_next.as<uint32_t>(),
                       d_size.as<uint32_t>());
    std::swap(d_lab.p, d_next.p);
    uint32_t changed = 0;
    HSDL_HIP(hipMemcpyAsync(&changed, d_changed.p, 4, hipMemcpyDeviceToHost, st));
    HSDL_HIP(hipStreamSynchronize(st));
    if (changed == 0) break;
  }
  HSDL_HIP(hipGetLastError());
  label.resize(n);
  HSDL_HIP(hipMemcpyAsync(label.data(), d_lab.p, n * 4ull, hipMemcpyDeviceToHost, st));
  HSDL_HIP(hipStreamSynchronize(st));
  return HSDL_OK;
}

// Greedy merge (CDLP_group.cuh:54-80): repeatedly take the largest group and
// join it with the largest group that still fits beside it within max_size
// (the group ordered just below it when that is itself); a group that can
// take no partner is final.  Groups are ordered by (size, id); a merged
// group keeps the id of its larger part.
static std::vector<std::vector<uint32_t>> merge_groups(std::vector<std::vector<uint32_t>> groups,
                                                       uint32_t max_size) {
  std::set<std::pair<uint64_t, uint32_t>> order;  // (size, id)
  for (uint32_t i = 0; i < groups.size(); ++i) order.insert({groups[i].size(), i});
  std::vector<std::vector<uint32_t>> out;
  while (!order.empty()) {
    auto it1 = std::prev(order.end());
    const uint64_t room = max_size > it1->first ? max_size - it1->first : 0;
    // largest (size, id) with size <= room
    auto ub = order.upper_bound({room, UINT32_MAX});
    auto it2 = order.end();
    if (ub != order.begin()) {
      it2 = std::prev(ub);
      if (it2 == it1) it2 = it2 == order.begin() ? order.end() : std::prev(it2);
    }
    if (it2 != order.end() && it1->first + it2->first <= max_size) {
      const uint32_t a = it1->second, b = it2->second;
      groups[a].insert(groups[a].end(), groups[b].begin(), groups[b].end());
      groups[b].clear();
      order.erase(it1);
      order.erase(it2);
      order.insert({groups[a].size(), a});
    } else {
      out.push_back(std::move(groups[it1->second]));
      order.erase(it1);
    }
  }
  return out;
}

static int groups_body(const hsdl_graph *g, uint32_t method, uint32_t max_size,
                       uint32_t rounds, uint64_t seed, int32_t device, uint32_t *group_of,
                       uint32_t *ngroups) {
  if (!g || !group_of || !ngroups) return fail(HSDL_EINVAL, "null argument");
  if (max_size == 0) return fail(HSDL_EINVAL, "max_size must be at least 1");
  if (method > HSDL_GROUP_CDLP) return fail(HSDL_EINVAL, "unknown grouping method");
  const uint32_t n = g->n;
  std::vector<std::vector<uint32_t>> groups;
  if (method == HSDL_GROUP_CONTIGUOUS || method == HSDL_GROUP_RANDOM) {
    std::vector<uint32_t> ids(n);
    std::iota(ids.begin(), ids.end(), 0u);
    if (method == HSDL_GROUP_RANDOM) {
      std::mt19937_64 rng(seed);
      std::shuffle(ids.begin(), ids.end(), rng);
    }
    for (uint64_t i = 0; i < n; i += max_size)
      groups.emplace_back(ids.begin() + i, ids.begin() + std::min<uint64_t>(n, i + max_size));
  } else {
    int rc = use_device(device, nullptr);
    if (rc) return rc;
    std::vector<uint32_t> label;
    if ((rc = cdlp_labels(g, max_size, rounds, label))) return rc;
    // communities in order of their smallest member, cut to max_size
    std::vector<uint32_t> first(n, UINT32_MAX);
    std::vector<std::vector<uint32_t>> comm;
    for (uint32_t v = 0; v < n; ++v) {
      const uint32_t l = label[v];
      if (first[l] == UINT32_MAX) {
        first[l] = (uint32_t)comm.size();
        comm.emplace_back();
      }
      comm[first[l]].push_back(v);
    }
    for (auto &c : comm)
      for (size_t i = 0; i < c.size(); i += max_size)
        groups.emplace_back(c.begin() + i, c.begin() + std::min(c.size(), i + (size_t)max_size));
    groups = merge_groups(std::move(groups), max_size);
    // largest total degree first (ties: the order the merge emitted them)
    std::vector<std::pair<int64_t, uint32_t>> key(groups.size());
    for (uint32_t i = 0; i < groups.size(); ++i) {
      int64_t deg = 0;
      for (uint32_t v : groups[i]) deg -= (int64_t)(g->off[v + 1] - g->off[v]);
      key[i] = {deg, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<std::vector<uint32_t>> sorted(groups.size());
    for (uint32_t i = 0; i < key.size(); ++i) sorted[i] = std::move(groups[key[i].second]);
    groups = std::move(sorted);
  }
  for (uint32_t i = 0; i < groups.size(); ++i)
    for (uint32_t v : groups[i]) group_of[v] = i;
  *ngroups = (uint32_t)groups.size();
  return HSDL_OK;
}

}  // namespace hsdl

extern "C" int hsdl_graph_groups(const hsdl_graph *g, uint32_t method, uint32_t max_size,
                                 uint32_t rounds, uint64_t seed, int32_t device,
                                 uint32_t *group_of, uint32_t *ngroups) {
  return hsdl::guarded(
      [&] { return hsdl::groups_body(g, method, max_size, rounds, seed, device, group_of, ngroups); });
}
