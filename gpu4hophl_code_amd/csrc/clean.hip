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
//  1. k_clean_tile: the predicate with `kept` replaced by every earlier label
//     of v.  That set contains the sequential kept set, so its minima are
//     never larger: when it removes nothing from L(v), the sequential walk
//     keeps every label of v as well.  Pass 1 has no order dependence.  One
//     workgroup (16 wavefronts) takes a vertex v: it builds a hash tile of
//     L(v) (hub -> per-hop prefix minimum distance; in LDS, or in the
//     workgroup's slice of a global arena for long lists), then the waves take
//     16-label windows of L(v) in turn, find the windows' group heads (hub u),
//     load-balance the heads' lists L(u) over the lanes by a register prefix
//     sum, and every element (w,h1,d1) of L(u) with w < u that could still
//     remove a group label (h1 + 1 <= the group's largest hop and d1 <= its
//     largest dist) looks w up once in the tile: a group label (u,h,d) falls
//     when d1 + tile[w][h - h1] <= d.  Pairs through u itself (earlier labels
//     of the same group) are checked directly.
//  2. k_clean_seq: vertices where pass 1 would remove something are walked
//     group by group in order with keep bits in LDS, exactly as the
//     reference does.
// A K-round, rank-ordered generation already satisfies the predicate (its
// acceptance test is the all-earlier test over the same final lists), so on
// this engine's own output pass 1 removes nothing and pass 2 never runs; the
// clean still verifies every label.
// The reference's GPU clean (Hybrid_Generation_Clean/src/label_clean.cu:73-164)
// uses a dense per-vertex hash of all labels instead; the kept-set form is
// the normative one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "device.hpp"

namespace hsdl {

constexpr uint32_t kCleanWaves = 4;            // waves per workgroup (pass 2)
constexpr uint32_t kSeqBitsPerWave = 65536;    // max labels per vertex (pass 2)
constexpr uint32_t kSeqWordsPerWave = kSeqBitsPerWave / 32;
constexpr uint32_t kTileThreads = 1024;        // pass 1: sixteen waves per vertex
constexpr uint32_t kTileWindow = 4;            // labels of L(v) per work unit (default)
constexpr uint32_t kMaxWindow = 64;            // largest window
// LDS tile budget: two workgroups per CU share its 160 KB
constexpr uint32_t kTileLdsBytes = 78 * 1024 + 16;
// Indexed LDS tiles (graphs of at most 2^20 vertices): a key slot holds
// hub << kIdxBits | row, and the hub's values for hop budgets 1..K sit in a
// compact row array (one row per distinct hub, not per slot; budget 0 holds
// only v's own hub, which pairs only through a head u > v and which the
// tile leaves out: a list with a hub above its vertex, possible only in
// imported label sets, takes a per-slot tile).
// So the LDS holds 4-byte key slots at load factor <= 1/8 for lists of up to
// about 2,000 hubs, and lists of up to 4,094 hubs at all, instead of 16-byte
// slots at up to 2/3 and lists of up to 2,730 (longer ones went to the
// arena, at 5x the time per label): a wave's probe chain is its longest
// lane's, and chains grow fast with the load factor.  Rows < 2^kIdxBits - 1
// keep a key slot from ever reading as empty (all ones).
constexpr uint32_t kIdxBits = 12;
constexpr uint32_t kIdxRows = (1u << kIdxBits) - 1;
constexpr uint32_t kGroupFast = 4;             // group sizes with the unrolled check

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

// ---------------------------------------------------------------- pass 1
// The heads' lists L(u) are streamed from the index's 8-byte labels, or from
// a 32-bit copy of them (hub | hop | dist fields) when the fields fit one
// word: the first pass reads far more list labels than anything else and
// runs near the rate HBM sustains, so half the bytes is what counts.
struct Stream64 {
  using T = unsigned long long;
  const unsigned long long *lab;
  __device__ const T *list(unsigned long long b) const { return lab + b; }
  // elements of L(u) of hop <= maxhop: the whole list (not hop-sorted)
  __device__ uint32_t upto(uint32_t, uint32_t, uint32_t len) const { return len; }
  __device__ void dec(T l, uint32_t &w, uint32_t &h, uint32_t &d) const {
    w = label_hub(l);
    h = label_hop(l);
    d = label_dist(l);
  }
};
// The 32-bit copy keeps each list stably sorted by hop (hub order within a
// hop), with hcnt[u * K + x] = elements of L(u) of hop <= x (x < K), when
// hcnt is set: an element pairs with the tile through hop h1 + 1 or more
// when w != v, so a group whose deepest label has hop maxh can use only
// L(u)'s elements of hop <= maxh - 1 -- a prefix of the sorted list, which
// the stream stops at (the self label, hop 0, is its first element).  Only a
// head u above v can hold w = v (pairing with v's self label, hop 0): its
// stream goes to hop maxh.
struct Stream32 {
  using T = uint32_t;
  const uint32_t *lab;
  uint32_t hop_bits, dist_bits;
  const uint32_t *hcnt;
  uint32_t K;
  __device__ const T *list(unsigned long long b) const { return lab + b; }
  __device__ uint32_t upto(uint32_t u, uint32_t maxhop, uint32_t len) const {
    return hcnt && maxhop < K ? hcnt[(uint64_t)u * K + maxhop] : len;
  }
  __device__ void dec(T l, uint32_t &w, uint32_t &h, uint32_t &d) const {
    w = l >> (hop_bits + dist_bits);
    h = __builtin_amdgcn_ubfe(l, dist_bits, hop_bits);
    d = l & ((1u << dist_bits) - 1u);
  }
};

__device__ inline uint32_t pack32(unsigned long long l, uint32_t hop_bits, uint32_t dist_bits) {
  return (label_hub(l) << (hop_bits + dist_bits)) | (label_hop(l) << dist_bits) | label_dist(l);
}

__global__ void k_pack32(uint64_t total, const unsigned long long *lab, uint32_t *out,
                         uint32_t hop_bits, uint32_t dist_bits) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = pack32(lab[i], hop_bits, dist_bits);
}

// The 32-bit copy with every list stably sorted by hop, and the per-hop
// prefix counts (Stream32): a wave per vertex counts its labels' hops, then
// scatters them in list order into their hop's run (positions by ballot).
__global__ void __launch_bounds__(256) k_pack32_hop(uint32_t n, uint32_t K,
                                                    const unsigned long long *off,
                                                    const unsigned long long *lab, uint32_t *out,
                                                    uint32_t *hcnt, uint32_t hop_bits,
                                                    uint32_t dist_bits) {
  const uint32_t lane = (uint32_t)lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  const uint32_t K1 = K + 1;
  for (uint32_t v = wave; v < n; v += nwaves) {
    const unsigned long long b = off[v];
    const uint32_t len = (uint32_t)(off[v + 1] - b);
    uint32_t cnt[kMaxK + 1];
#pragma unroll
    for (uint32_t x = 0; x <= kMaxK; ++x) cnt[x] = 0;
    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t j = j0 + lane;
      const uint32_t hop = j < len ? label_hop(lab[b + j]) : kInf32;
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x)
        if (x < K1) cnt[x] += (uint32_t)__popcll(__ballot(hop == x));
    }
    uint32_t pos[kMaxK + 1];
    uint32_t run = 0;
#pragma unroll
    for (uint32_t x = 0; x <= kMaxK; ++x) {
      pos[x] = run;
      run += x < K1 ? cnt[x] : 0u;
      // hcnt[v * K + x]: labels of hop <= x, for x = 0..K-1
      if (x < K && lane == 0) hcnt[(uint64_t)v * K + x] = run;
    }
    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t j = j0 + lane;
      unsigned long long l = 0;
      uint32_t hop = kInf32;
      if (j < len) {
        l = lab[b + j];
        hop = label_hop(l);
      }
      uint32_t dst = 0;
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x) {
        if (x >= K1) continue;
        const unsigned long long m = __ballot(hop == x);
        if (hop == x) dst = pos[x] + (uint32_t)__popcll(m & lt_mask);
        pos[x] += (uint32_t)__popcll(m);
      }
      if (j < len) out[b + dst] = pack32(l, hop_bits, dist_bits);
    }
  }
}

// True when the all-earlier predicate removes a label of a group whose head
// lies in L(v)[p0, p0 + win).  The wave walks the window's heads in order
// (uniform control flow: the group's labels are uniform scalar loads) and
// streams each head's list L(u) 128 elements per step with coalesced loads;
// only the filter and the tile probe are per lane.  The two step buffers
// alternate (a, b) and loads are clamped to the list (lanes past the end are
// masked), so a step's probes run while the next step's loads are in flight.
// KP / VP: the tile's address space (LDS or the global arena).
template <class V, bool kIdx, class S, class KP, class VP>
__device__ bool window_removes(uint32_t v, uint32_t K1, const unsigned long long *__restrict__ off,
                               const S &str, const unsigned long long *__restrict__ Lv, uint32_t len,
                               uint32_t p0, uint32_t win, KP keys, VP vals, uint32_t cap,
                               uint32_t last_row, uint64_t &streamed) {
  constexpr bool kArena = !std::is_same<KP, lds_ptr<uint32_t>>::value;
  // hub w's key slot: sets `found` and returns the index of its value row
  // (an empty slot's is any in-bounds row: the caller masks it)
  auto find = [&](uint32_t w, bool &found) -> uint32_t {
    uint32_t slot = hash_slot(w, cap);
    uint32_t k = keys[slot];
    if (kIdx) {
      while ((k >> kIdxBits) != w && k != kInf32) {
        slot = (slot + 1) & (cap - 1);
        k = keys[slot];
      }
      // an empty slot reads hub 2^20 - 1, which ranks below every head u
      // that the pair test also asks for (w < u)
      found = (k >> kIdxBits) == w;
      return mul_u24(min(k & kIdxRows, last_row), K1 - 1);
    }
    while (k != w && k != kInf32) {
      slot = (slot + 1) & (cap - 1);
      k = keys[slot];
    }
    found = k == w;
    return mul_u24(slot, K1);
  };
  // the value of row `base` for a group label of hop g and an element of hop
  // h1 (budget x = g - h1), and whether the budget can pair (reach): per-slot
  // rows hold budgets 0..K, indexed rows 1..K (budget 0 holds only v's hub)
  auto value_at = [&](uint32_t base, uint32_t g, uint32_t h1, bool &reach) -> uint32_t {
    if (kIdx) {
      reach = g > h1;
      return vals[base + (reach ? g - 1u - h1 : 0u)];
    }
    reach = g >= h1;
    return vals[base + (reach ? g - h1 : 0u)];
  };
  const uint32_t lane = (uint32_t)lane_id();
  bool hit = false;
  const uint32_t pend = min(p0 + win, len);
  for (uint32_t p = p0; p < pend; ++p) {
    const uint32_t u = label_hub(Lv[p]);
    if ((p > 0 && label_hub(Lv[p - 1]) == u) || u == v) continue;  // not a head / self label
    uint32_t glen = 0, maxh = 0, maxd = 0;
    while (p + glen < len && label_hub(Lv[p + glen]) == u) {
      const unsigned long long g = Lv[p + glen];
      maxh = max(maxh, label_hop(g));
      maxd = max(maxd, label_dist(g));
      ++glen;
    }
    const unsigned long long ub = off[u];
    // an element (w,h1,d1) of L(u) pairs with the tile through hop >= h1 + 1
    // when w != v; w = v (possible only for u > v, in imported label sets
    // whose hubs may rank below their vertex) pairs with v's self label
    const uint32_t hmin = u > v ? 0u : 1u;
    const uint32_t ulen = str.upto(u, maxh - hmin, (uint32_t)(off[u + 1] - ub));
    if (ulen == 0) continue;
    const typename S::T *Lu = str.list(ub);
    // the group's hops and distances, decoded once per head into scalars for
    // groups of up to kGroupFast labels (a group holds at most K + 1, one per
    // hop): the per-element loop over them is then unrolled straight-line
    // code instead of a scalar load, a wait and a decode per group label
    const bool fast = glen <= kGroupFast;
    uint32_t gh[kGroupFast], gd[kGroupFast];
#pragma unroll
    for (uint32_t t = 0; t < kGroupFast; ++t) {
      const unsigned long long g = (fast && t < glen) ? Lv[p + t] : 0ull;
      gh[t] = label_hop(g);
      gd[t] = label_dist(g);
    }
    streamed += ulen;
    const uint32_t last = ulen - 1;
    // one element (w,h1,d1) of L(u) against the group.  The pass issues
    // VALU instructions near the rate a CU sustains, so the check is
    // straight-line code but for the rare collision chain:
    // * a lane past the list's end holds a copy of its last element, which
    //   gives the same answer as the element itself: no mask;
    // * in an LDS tile the lookup runs for every lane: the chain ends at w
    //   or at an empty slot (whose values count only when k == w), and the
    //   shortcuts d1 <= maxd and h1 + hmin <= maxh are left to the pair test,
    //   which implies them (d1 + m <= gd with m >= 0; budget gh - h1 = 0
    //   holds only v's self label, hop-0 labels being self labels); an arena
    //   tile's lookups are global loads, which the shortcuts still save;
    // * a 16-bit empty value (0xFFFF) is above every label distance (small
    //   values: the largest distance < 0xFFFF), so only 32-bit values test
    //   for it.
    auto check = [&](typename S::T l) {
      uint32_t w, h1, d1;
      str.dec(l, w, h1, d1);
      if (w == u) {
        // through hub u: earlier labels of the group against later ones
        for (uint32_t t1 = 0; t1 < glen; ++t1) {
          const unsigned long long g1 = Lv[p + t1];
          for (uint32_t t2 = t1 + 1; t2 < glen; ++t2) {
            const unsigned long long g2 = Lv[p + t2];
            if (h1 + label_hop(g1) <= label_hop(g2) && d1 + label_dist(g1) <= label_dist(g2))
              hit = true;
          }
        }
      }
      if (kArena && !((w < u) & (h1 + hmin <= maxh) & (d1 <= maxd))) return;
      bool found;
      const uint32_t base = find(w, found);
      const bool pair = found & (w < u);
      auto fits = [&](uint32_t m, uint32_t gd) -> bool {
        if (sizeof(V) == 2) return d1 + m <= gd;
        return (m != TileVal<V>::kEmpty) & (d1 + m <= gd);
      };
      if (fast) {
#pragma unroll
        for (uint32_t t = 0; t < kGroupFast; ++t) {
          if (t >= glen) break;
          // a label shallower than h1 reads the row's first value and is
          // masked; at equal hops only the hop-0 tile label (v's) pairs
          bool reach;
          const uint32_t m = value_at(base, gh[t], h1, reach);
          hit |= pair & reach & fits(m, gd[t]);
        }
      } else {
        for (uint32_t t = 0; t < glen; ++t) {
          const unsigned long long g = Lv[p + t];
          const uint32_t hg = label_hop(g);
          bool reach;
          const uint32_t m = value_at(base, hg, h1, reach);
          if (pair && reach && fits(m, label_dist(g))) hit = true;
        }
      }
    };
    // a group of one label (most groups): its hop and distance are the
    // group's maxima, no pair runs through u inside the group, and the
    // per-label loop and its branches go
    auto check1 = [&](typename S::T l) {
      uint32_t w, h1, d1;
      str.dec(l, w, h1, d1);
      if (kArena && !((w < u) & (h1 + hmin <= maxh) & (d1 <= maxd))) return;
      bool found, reach;
      const uint32_t base = find(w, found);
      const uint32_t m = value_at(base, gh[0], h1, reach);
      const bool near = sizeof(V) == 2 ? d1 + m <= gd[0]
                                       : (m != TileVal<V>::kEmpty) & (d1 + m <= gd[0]);
      hit |= found & (w < u) & reach & near;
    };
    auto ld = [&](uint32_t j) { return Lu[min(j, last)]; };
    auto stream = [&](auto &&chk) {
      typename S::T a0 = ld(lane), a1 = ld(kWave + lane);
      for (uint32_t jb = 0; jb < ulen; jb += 4 * kWave) {
        const typename S::T b0 = ld(jb + 2 * kWave + lane), b1 = ld(jb + 3 * kWave + lane);
        chk(a0);
        chk(a1);
        if (jb + 2 * kWave >= ulen) break;
        a0 = ld(jb + 4 * kWave + lane);
        a1 = ld(jb + 5 * kWave + lane);
        chk(b0);
        chk(b1);
      }
    };
    if (glen == 1)
      stream(check1);
    else
      stream(check);
  }
  return __ballot(hit) != 0;
}

// Builds the tile of L(v) (keys, vals in LDS or in the workgroup's arena
// slice) and runs the windows of L(v) over it; true when some window removes
// a label.  Whole workgroup.
template <class V, bool kIdx, class S, class KP, class VP>
__device__ __attribute__((always_inline)) inline bool clean_vertex(
    uint32_t v, uint32_t K1, const unsigned long long *__restrict__ off, const S &str,
    const unsigned long long *__restrict__ Lv,
    uint32_t len, KP keys, VP vals, uint32_t cap, uint32_t rows, bool in_arena, uint32_t win,
    uint32_t &sh_win, uint32_t &sh_removed, uint32_t &sh_row, uint64_t &streamed) {
  const uint32_t tid = threadIdx.x, lane = (uint32_t)lane_id();
  for (uint32_t i = tid; i < cap; i += kTileThreads) keys[i] = kInf32;
  if (tid == 0) sh_row = 0;
  __syncthreads();
  // one thread per group: claim the hub's slot (and, indexed, a value row),
  // write its per-hop prefix minima
  for (uint32_t i = tid; i < len; i += kTileThreads) {
    const uint32_t w = label_hub(Lv[i]);
    if (i > 0 && label_hub(Lv[i - 1]) == w) continue;
    if (kIdx && w == v) continue;  // v's own hub: no head u > v pairs with it here
    uint32_t slot = hash_slot(w, cap);
    if (kIdx) {
      const uint32_t row = atomicAdd(&sh_row, 1u);
      while (atomicCAS((uint32_t *)&keys[slot], kInf32, (w << kIdxBits) | row) != kInf32)
        slot = (slot + 1) & (cap - 1);
      slot = row;  // the values go to the row
    } else {
      while (atomicCAS((uint32_t *)&keys[slot], kInf32, w) != kInf32) slot = (slot + 1) & (cap - 1);
    }
    uint32_t m[kMaxK + 1];
#pragma unroll
    for (uint32_t x = 0; x <= kMaxK; ++x) m[x] = kInf32;
    for (uint32_t q = i; q < len && label_hub(Lv[q]) == w; ++q) {
      const uint32_t h = label_hop(Lv[q]), d = label_dist(Lv[q]);
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x)
        if (x == h) m[x] = min(m[x], d);
    }
    uint32_t run = kInf32;
#pragma unroll
    for (uint32_t x = 0; x <= kMaxK; ++x) {
      run = min(run, m[x]);
      const V val = (V)min(run, (uint32_t)TileVal<V>::kEmpty);
      if (kIdx) {
        if (x >= 1 && x < K1) vals[slot * (K1 - 1) + (x - 1)] = val;  // budgets 1..K
      } else {
        if (x < K1) vals[slot * K1 + x] = val;
      }
    }
  }
  // an arena tile was written through L2: drop stale L1 lines before reading
  if (in_arena) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (in_arena) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  for (;;) {
    uint32_t p0 = 0;
    if (lane == 0) p0 = atomicAdd(&sh_win, win);
    p0 = __builtin_amdgcn_readfirstlane(__shfl(p0, 0, kWave));
    if (p0 >= len) break;
    if (window_removes<V, kIdx>(v, K1, off, str, Lv, len, p0, win, keys, vals, cap, rows - 1,
                                streamed)) {
      if (lane == 0) sh_removed = 1;
      break;
    }
  }
  __syncthreads();
  return sh_removed != 0;
}

// ctr[0]: redo count, ctr[2]: next vertex, ctr[4..5]: labels of L(u) read (u64).  gtile: per-workgroup arena slices
// of gwords words for tiles above lds_slots slots.
template <class V, class S>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_clean_tile(uint32_t n, uint32_t K, const unsigned long long *__restrict__ off,
             const unsigned long long *__restrict__ lab, S str,
             uint8_t *keep, uint32_t *new_len, uint32_t *redo, uint32_t *ctr, uint32_t lds_slots,
             uint32_t *gtile, uint64_t gwords, uint32_t win, uint32_t indexed) {
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t sh_v, sh_win, sh_heads, sh_removed, sh_row;
  // LDS bytes for a tile (the launch gives lds_slots slots of key + values)
  const uint32_t lds_tile_bytes = lds_slots * (4u + (K + 1) * (uint32_t)sizeof(V));
  const uint32_t K1 = K + 1;
  const uint32_t tid = threadIdx.x, lane = (uint32_t)lane_id();
  uint32_t *gkeys = gtile + (uint64_t)blockIdx.x * gwords;
  uint64_t streamed = 0;  // labels of L(u) lists this wave read
  for (;;) {
    if (tid == 0) {
      sh_v = atomicAdd(&ctr[2], 1u);
      sh_win = 0;
      sh_heads = 0;
      sh_removed = 0;
    }
    __syncthreads();
    // wave-uniform values are made explicit, so the list offsets and hubs are
    // scalar loads and the loops over them carry no per-lane masks
    const uint32_t v = __builtin_amdgcn_readfirstlane(sh_v);
    if (v >= n) break;
    const unsigned long long vb = off[v];
    const uint32_t len = (uint32_t)(off[v + 1] - vb);
    const unsigned long long *Lv = lab + vb;
    if (len > 1) {
      // distinct hubs of L(v) size the tile
      uint32_t mine = 0;
      for (uint32_t i = tid; i < len; i += kTileThreads)
        mine += (i == 0 || label_hub(Lv[i]) != label_hub(Lv[i - 1])) ? 1u : 0u;
      mine = wave_sum_u32(mine);
      if (lane == 0 && mine) atomicAdd(&sh_heads, mine);
      __syncthreads();
      const uint32_t heads = sh_heads;
      const uint32_t cap = tile_cap_for(heads);
      // indexed: key slots at load factor <= 1/8 where the rows leave room
      uint32_t icap = 0;
      // (lists are hub-sorted: the last hub is the largest)
      if (indexed && heads <= kIdxRows && label_hub(Lv[len - 1]) <= v) {
        const uint32_t row_bytes = (heads * K * (uint32_t)sizeof(V) + 15u) & ~15u;
        if (row_bytes + 4u * cap <= lds_tile_bytes) {
          icap = pow2_at_least(8ull * heads + 4);
          while (row_bytes + 4u * icap > lds_tile_bytes) icap >>= 1;
        }
      }
      if (icap) {
        clean_vertex<V, true>(v, K1, off, str, Lv, len, (lds_ptr<uint32_t>)lds,
                              (lds_ptr<V>)(lds + icap), icap, heads, false, win, sh_win,
                              sh_removed, sh_row, streamed);
      } else if (cap <= lds_slots) {
        const uint32_t lcap = lds_tile_cap(heads, lds_slots);
        clean_vertex<V, false>(v, K1, off, str, Lv, len, (lds_ptr<uint32_t>)lds,
                               (lds_ptr<V>)(lds + lcap), lcap, lcap, false, win, sh_win,
                               sh_removed, sh_row, streamed);
      } else {
        clean_vertex<V, false>(v, K1, off, str, Lv, len, (glb_ptr<uint32_t>)gkeys,
                               (glb_ptr<V>)(gkeys + cap), cap, cap, true, win, sh_win,
                               sh_removed, sh_row, streamed);
      }
    }
    if (sh_removed) {
      if (tid == 0) redo[atomicAdd(&ctr[0], 1u)] = v;
    } else {
      for (uint32_t i = tid; i < len; i += kTileThreads) keep[vb + i] = 1;
      if (tid == 0) new_len[v] = len;
    }
    __syncthreads();  // shared state and the tile are reused for the next vertex
  }
  if (lane == 0 && streamed)
    atomicAdd(reinterpret_cast<unsigned long long *>(ctr) + 2, (unsigned long long)streamed);
}

// Largest list length and largest label distance of an index (out[0], out[1]).
__global__ void __launch_bounds__(256) k_index_maxima(uint32_t n, const unsigned long long *off,
                                                      const unsigned long long *lab,
                                                      uint32_t *out) {
  uint32_t ml = 0, md = 0;
  const uint32_t nth = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += nth) {
    const unsigned long long b = off[v], e = off[v + 1];
    ml = max(ml, (uint32_t)(e - b));
  }
  const unsigned long long total = off[n];
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < total;
       i += nth)
    md = max(md, label_dist(lab[i]));
  ml = wave_max_u32(ml);
  md = wave_max_u32(md);
  if (lane_id() == 0) {
    atomicMax(&out[0], ml);
    atomicMax(&out[1], md);
  }
}

// ---------------------------------------------------------------- pass 2
// The reference's sequential walk for the vertices in `list`.  For a group of
// hub u the 64 lanes stream L(u); each lane binary-searches its label's hub
// w < u in the decided prefix of L(v) and folds the kept matches into per-hop
// minima m[x] (x = h1 + h2).  A wave min-reduction and a prefix-min over x
// give the best pair for every hop budget; the group's labels are then
// decided in hop order, adding the pairs through hub u itself.
// kLong = false: lists of up to kSeqBitsPerWave labels, keep bits in LDS;
// longer lists are left to the kLong = true launch, whose keep bits live in
// the wave's slice of `gbits` (gwords words): read and written with
// agent-scope atomics, so every lane sees lane 0's updates without relying
// on the vector L1.  A hub group of more than 32 labels (impossible with one
// label per (hub, hop) and K <= 15, which import and generation guarantee)
// sets bad bit 2 instead of overflowing the group's bit mask.
template <bool kLong>
__global__ void __launch_bounds__(256) k_clean_seq(uint32_t K, const unsigned long long *off,
                                                   const unsigned long long *lab, uint8_t *keep,
                                                   uint32_t *new_len, const uint32_t *list,
                                                   const uint32_t *list_cnt, uint32_t *bad,
                                                   uint32_t *gbits, uint32_t gwords) {
  __shared__ uint32_t bits_all[kLong ? 1 : kCleanWaves * kSeqWordsPerWave];
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  const uint32_t K1 = K + 1;
  const uint32_t wave = blockIdx.x * kCleanWaves + wv;
  const uint32_t nwaves = gridDim.x * kCleanWaves;
  uint32_t *const gb_slice = kLong ? gbits + (uint64_t)wave * gwords : nullptr;
  auto bit_word = [&](uint32_t i) -> uint32_t {
    if constexpr (kLong)
      return __hip_atomic_load(gb_slice + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      return bits_all[wv * kSeqWordsPerWave + i];
  };
  auto bits_sync = [&]() {
    if constexpr (kLong) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
      __builtin_amdgcn_wave_barrier();
    } else {
      wave_lds_sync();
    }
  };
  const uint32_t cnt = *list_cnt;
  for (uint32_t li = wave; li < cnt; li += nwaves) {
    const uint32_t v = list[li];
    const unsigned long long vb = off[v];
    const uint32_t len = (uint32_t)(off[v + 1] - vb);
    const unsigned long long *Lv = lab + vb;
    if (kLong ? (len <= kSeqBitsPerWave) : (len > kSeqBitsPerWave)) continue;  // the other launch's
    if (kLong && (len + 31) / 32 > gwords) {
      if (lane == 0) atomicOr(bad, 1u);
      continue;
    }
    for (uint32_t i = lane; i < (len + 31) / 32; i += kWave) {
      if constexpr (kLong)
        __hip_atomic_store(gb_slice + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else
        bits_all[wv * kSeqWordsPerWave + i] = 0;
    }
    bits_sync();
    uint32_t kept_total = 0;
    uint32_t gb = 0;
    while (gb < len) {
      const uint32_t u = label_hub(Lv[gb]);
      uint32_t ge = gb + 1;
      while (ge < len && label_hub(Lv[ge]) == u) ++ge;  // <= K+1 labels
      if (ge - gb > 32) {  // see above: cannot happen on valid input
        if (lane == 0) atomicOr(bad, 2u);
        break;
      }
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
            if (!((bit_word(p >> 5) >> (p & 31)) & 1u)) continue;
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
      // L(u)'s own labels (hub u): found by search, since a list may also
      // hold hubs ranked below its vertex (imported label sets)
      const uint32_t ub0 = lb_hub(Lu, 0, ulen, u);
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
          for (uint32_t t = ub0; t < ulen; ++t) {
            const unsigned long long r = Lu[t];
            if (label_hub(r) != u) break;
            if (label_hop(r) + h2 <= h) best = min(best, label_dist(r) + d2);
          }
        }
        if (best > d) newbits |= 1u << (i - gb);
      }
      if (lane == 0) {
        for (uint32_t i = gb; i < ge; ++i)
          if ((newbits >> (i - gb)) & 1u) {
            if constexpr (kLong)
              atomicOr(gb_slice + (i >> 5), 1u << (i & 31));
            else
              atomicOr(&bits_all[wv * kSeqWordsPerWave + (i >> 5)], 1u << (i & 31));
          }
      }
      kept_total += __popc(newbits);
      bits_sync();
      gb = ge;
    }
    for (uint32_t i = lane; i < len; i += kWave) keep[vb + i] = (bit_word(i >> 5) >> (i & 31)) & 1u;
    if (lane == 0) new_len[v] = kept_total;
    bits_sync();
  }
}

// Kept labels (and their parents, when par is not null) into the new CSR.
__global__ void __launch_bounds__(256) k_compact(uint32_t n, const unsigned long long *off,
                                                 const unsigned long long *lab,
                                                 const uint8_t *keep,
                                                 const unsigned long long *noff,
                                                 unsigned long long *nlab,
                                                 const uint32_t *par, uint32_t *npar) {
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
      if (k) {
        nlab[dst + __popcll(mk & lt)] = lab[i + lane];
        if (par) npar[dst + __popcll(mk & lt)] = par[i + lane];
      }
      dst += __popcll(mk);
    }
  }
}

// The clean's working buffers (keep flags, counts, the 32-bit copy and its
// hop counts, the tile arena) and the index arrays a clean replaces are kept
// per device for the next clean, sized for the largest so far: freeing and
// allocating gigabytes inside every clean now and then cost a tenth of a
// second there.  hsdl_release_scratch frees them.
enum CleanScratchId : int {
  kCsKeep, kCsNewLen, kCsRedo, kCsGtile, kCsLab32, kCsHcnt, kCsOff, kCsLab, kCsCount
};
constexpr int kCsDevices = 64;
struct CleanScratch {
  void *p[kCsCount] = {};
  size_t bytes[kCsCount] = {};
};
static CleanScratch g_clean_scratch[kCsDevices];
static std::mutex g_clean_scratch_mu;

// b <- a kept buffer of at least n bytes when there is one, else a new one
static int take_scratch(int dev, int id, DevBuf &b, size_t n) {
  if (dev >= 0 && dev < kCsDevices) {
    std::lock_guard<std::mutex> lk(g_clean_scratch_mu);
    CleanScratch &c = g_clean_scratch[dev];
    if (c.p[id] && c.bytes[id] >= n) {
      b.release();
      b.p = c.p[id];
      b.bytes = c.bytes[id];
      c.p[id] = nullptr;
      c.bytes[id] = 0;
      return HSDL_OK;
    }
    if (c.p[id]) {
      (void)hipFree(c.p[id]);
      c.p[id] = nullptr;
      c.bytes[id] = 0;
    }
  }
  return b.alloc(n);
}
// b goes back to the device's kept buffers (or is freed when the slot is taken)
static void give_scratch(int dev, int id, DevBuf &b) {
  if (!b.p || dev < 0 || dev >= kCsDevices) return;
  std::lock_guard<std::mutex> lk(g_clean_scratch_mu);
  CleanScratch &c = g_clean_scratch[dev];
  if (c.p[id]) return;  // b frees itself
  c.p[id] = b.p;
  c.bytes[id] = b.bytes;
  b.p = nullptr;
  b.bytes = 0;
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
  DevBuf keep, new_len, noff, nlab, redo, ctr, gtile, lab32, hcnt;
  const int dev = idx->device;
  // every buffer taken from the kept ones goes back when the clean ends
  struct Give {
    int dev;
    DevBuf *b[6];
    int id[6];
    ~Give() {
      for (int i = 0; i < 6; ++i) give_scratch(dev, id[i], *b[i]);
    }
  } give{dev, {&keep, &new_len, &redo, &gtile, &lab32, &hcnt},
         {kCsKeep, kCsNewLen, kCsRedo, kCsGtile, kCsLab32, kCsHcnt}};
  if ((rc = take_scratch(dev, kCsKeep, keep, std::max<uint64_t>(idx->total, 1))) ||
      (rc = take_scratch(dev, kCsNewLen, new_len, n * 4ull)) ||
      (rc = take_scratch(dev, kCsOff, noff, (n + 1ull) * 8)) ||
      (rc = take_scratch(dev, kCsRedo, redo, n * 4ull + 4)) || (rc = ctr.alloc(32)))
    return rc;
  HSDL_HIP(hipMemsetAsync(ctr.p, 0, 32, st));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, idx->device);
  const int grid = std::max(1, cus * 8);
  uint32_t *cnt = ctr.as<uint32_t>();
  // list-length and distance maxima size the global tile arena and pick the
  // tile's value width
  DevBuf maxima;
  if ((rc = maxima.alloc(8))) return rc;
  HSDL_HIP(hipMemsetAsync(maxima.p, 0, 8, st));
  hipLaunchKernelGGL(k_index_maxima, dim3(cus * 4), dim3(256), 0, st, n,
                     idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                     maxima.as<uint32_t>());
  HSDL_HIP(hipGetLastError());
  uint32_t mx[2] = {0, 0};
  HSDL_HIP(hipMemcpyAsync(mx, maxima.p, 8, hipMemcpyDeviceToHost, st));
  HSDL_HIP(hipStreamSynchronize(st));
  const uint32_t K1 = (uint32_t)idx->K + 1;
  const bool small_vals = mx[1] < 0xFFFFu;
  // indexed LDS tiles need hub << kIdxBits to fit a word (tuning knob
  // HSDL_CLEAN_INDEXED=0: per-slot values only)
  const uint32_t indexed =
      (uint64_t)n <= (1ull << (32 - kIdxBits)) ? env_u32("HSDL_CLEAN_INDEXED", 1, 0, 1) : 0u;
  const uint32_t slot_bytes = 4 + K1 * (small_vals ? 2 : 4);
  // tuning knob HSDL_CLEAN_LDS_SLOTS caps the LDS tile (small values send
  // every vertex to the global arena)
  uint32_t lds_slots = (kTileLdsBytes - 16) / slot_bytes;
  lds_slots = env_u32("HSDL_CLEAN_LDS_SLOTS", lds_slots, 1, lds_slots);
  const size_t lds_bytes = (size_t)lds_slots * slot_bytes + 16;
  const int tgrid = std::max(1, cus * 2);
  // tuning knob HSDL_CLEAN_WINDOW: labels of L(v) per wave work unit
  const uint32_t win = env_u32("HSDL_CLEAN_WINDOW", kTileWindow, 1, kMaxWindow);
  const uint32_t gcap = tile_cap_for(mx[0]);
  const uint64_t gwords = small_vals ? tile_words<uint16_t>(gcap, K1) : tile_words<uint32_t>(gcap, K1);
  if ((rc = take_scratch(dev, kCsGtile, gtile, (uint64_t)tgrid * gwords * 4))) return rc;
  // 32-bit copy of the labels for the heads' list streams, when hub, hop and
  // distance fields fit one word (tuning knob HSDL_STREAM32=0 keeps the
  // 8-byte stream; so does a failed allocation of the copy)
  uint32_t hub_bits = 1, hop_bits = 1, dist_bits = 1;
  while ((1ull << hub_bits) <= (uint64_t)(n ? n - 1 : 0)) ++hub_bits;
  while ((1ull << hop_bits) <= (uint64_t)idx->K) ++hop_bits;
  while ((1ull << dist_bits) <= (uint64_t)mx[1]) ++dist_bits;
  bool use32 = env_u32("HSDL_STREAM32", 1, 0, 1) && idx->total > 0 &&
               hub_bits + hop_bits + dist_bits <= 32;
  // tuning knob HSDL_CLEAN_HOP_SORT=0 keeps the 32-bit copy in the index's
  // (hub, hop, dist) order and streams whole lists
  const bool hop_sort = use32 && env_u32("HSDL_CLEAN_HOP_SORT", 1, 0, 1);
  if (use32) {
    if (take_scratch(dev, kCsLab32, lab32, idx->total * 4) == HSDL_OK) {
      if (hop_sort) {
        if ((rc = take_scratch(dev, kCsHcnt, hcnt, (uint64_t)n * idx->K * 4))) return rc;
        hipLaunchKernelGGL(k_pack32_hop, dim3(cus * 8), dim3(256), 0, st, n, (uint32_t)idx->K,
                           idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(),
                           lab32.as<uint32_t>(), hcnt.as<uint32_t>(), hop_bits, dist_bits);
      } else {
        hipLaunchKernelGGL(k_pack32, dim3(cus * 8), dim3(256), 0, st, (uint64_t)idx->total,
                           idx->lab.as<unsigned long long>(), lab32.as<uint32_t>(), hop_bits,
                           dist_bits);
      }
      HSDL_HIP(hipGetLastError());
    } else {
      (void)hipGetLastError();
      use32 = false;
    }
  }
  const Stream64 s64{idx->lab.as<unsigned long long>()};
  const Stream32 s32{lab32.as<uint32_t>(), hop_bits, dist_bits,
                     use32 && hop_sort ? hcnt.as<uint32_t>() : nullptr, (uint32_t)idx->K};
  auto launch = [&](auto kern, auto str) {
    hipLaunchKernelGGL(kern, dim3(tgrid), dim3(kTileThreads), lds_bytes, st, n, (uint32_t)idx->K,
                       idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>(), str,
                       keep.as<uint8_t>(), new_len.as<uint32_t>(), redo.as<uint32_t>(), cnt,
                       lds_slots, gtile.as<uint32_t>(), gwords, win, indexed);
  };
  const uint64_t t_kern = now_ns();  // buffers and the 32-bit copy's set-up end here
  if (small_vals) {
    if (use32) launch(k_clean_tile<uint16_t, Stream32>, s32);
    else launch(k_clean_tile<uint16_t, Stream64>, s64);
  } else {
    if (use32) launch(k_clean_tile<uint32_t, Stream32>, s32);
    else launch(k_clean_tile<uint32_t, Stream64>, s64);
  }
  HSDL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_clean_seq<false>, dim3(grid), dim3(kCleanWaves * kWave), 0, st,
                     (uint32_t)idx->K, idx->off.as<unsigned long long>(),
                     idx->lab.as<unsigned long long>(), keep.as<uint8_t>(),
                     new_len.as<uint32_t>(), redo.as<uint32_t>(), cnt, cnt + 1, nullptr, 0u);
  HSDL_HIP(hipGetLastError());
  DevBuf long_bits;
  if (mx[0] > kSeqBitsPerWave) {
    // vertices with more labels than an LDS bit set holds: keep bits in a
    // global slice per wave (few waves: such vertices are rare)
    const uint32_t lgrid = 64;
    const uint32_t words = (mx[0] + 31) / 32;
    if ((rc = long_bits.alloc((uint64_t)lgrid * kCleanWaves * words * 4))) return rc;
    hipLaunchKernelGGL(k_clean_seq<true>, dim3(lgrid), dim3(kCleanWaves * kWave), 0, st,
                       (uint32_t)idx->K, idx->off.as<unsigned long long>(),
                       idx->lab.as<unsigned long long>(), keep.as<uint8_t>(),
                       new_len.as<uint32_t>(), redo.as<uint32_t>(), cnt, cnt + 1,
                       long_bits.as<uint32_t>(), words);
    HSDL_HIP(hipGetLastError());
  }
  uint32_t hc[6] = {0, 0, 0, 0, 0, 0};
  HSDL_HIP(hipMemcpyAsync(hc, ctr.p, 24, hipMemcpyDeviceToHost, st));
  HSDL_HIP(hipStreamSynchronize(st));
  const uint64_t t_passes = now_ns();
  if (hc[1] & 2u) return fail(HSDL_EINVAL, "a vertex holds more than 32 labels of one hub");
  if (hc[1]) return fail(HSDL_EDEVICE, "clean: keep-bit scratch smaller than a list");
  if ((rc = scan_counts_u32(new_len.as<uint32_t>(), noff.as<uint64_t>(), n, st))) return rc;
  uint64_t total = 0;
  HSDL_HIP(hipMemcpy(&total, noff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost));
  if ((rc = take_scratch(dev, kCsLab, nlab, std::max<uint64_t>(total, 1) * 8))) return rc;
  DevBuf npar;
  if (idx->has_parents && (rc = npar.alloc(std::max<uint64_t>(total, 1) * 4))) return rc;
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, st, n, idx->off.as<unsigned long long>(),
                     idx->lab.as<unsigned long long>(), keep.as<uint8_t>(),
                     noff.as<unsigned long long>(), nlab.as<unsigned long long>(),
                     idx->has_parents ? idx->par.as<uint32_t>() : nullptr,
                     idx->has_parents ? npar.as<uint32_t>() : nullptr);
  HSDL_HIP(hipGetLastError());
  HSDL_HIP(hipStreamSynchronize(st));
  // the replaced arrays are kept for the next clean's new ones
  std::swap(idx->off, noff);
  std::swap(idx->lab, nlab);
  idx->lab32.release();
  idx->q32.store(0, std::memory_order_release);
  give_scratch(dev, kCsOff, noff);
  give_scratch(dev, kCsLab, nlab);
  if (idx->has_parents) idx->par = std::move(npar);
  idx->total = total;
  idx->cleaned = true;
  idx->stats.labels_cleaned = total;
  idx->stats.clean_passes = hc[0] ? 2 : 1;
  idx->stats.clean_seq_vertices = hc[0];
  idx->stats.clean_label_reads = (uint64_t)hc[4] | ((uint64_t)hc[5] << 32);
  idx->stats.clean_label_bytes = use32 ? 4 : 8;
  const uint64_t t_end = now_ns();
  idx->stats.t_clean_s = (t_end - t0) * 1e-9;
  if (getenv("HSDL_VERBOSE"))
    fprintf(stderr, "[hsdl] clean: %.3f s set-up, %.3f s passes, %.3f s compaction\n",
            (t_kern - t0) * 1e-9, (t_passes - t_kern) * 1e-9, (t_end - t_passes) * 1e-9);
  return HSDL_OK;
}

}  // namespace hsdl

static int clean_body(hsdl_index *idx) {
  if (!idx) return hsdl::fail(HSDL_EINVAL, "null index");
  return hsdl::clean_index(idx);
}

extern "C" int hsdl_clean(hsdl_index *idx) {
  return hsdl::guarded([&] { return clean_body(idx); });
}

static int release_scratch_body(int32_t device) {
  std::lock_guard<std::mutex> lk(hsdl::g_clean_scratch_mu);
  int cur = -1;
  (void)hipGetDevice(&cur);
  for (int d = 0; d < hsdl::kCsDevices; ++d) {
    if (device >= 0 && d != device) continue;
    hsdl::CleanScratch &c = hsdl::g_clean_scratch[d];
    bool any = false;
    for (int i = 0; i < hsdl::kCsCount; ++i) any |= c.p[i] != nullptr;
    if (!any) continue;
    (void)hipSetDevice(d);
    for (int i = 0; i < hsdl::kCsCount; ++i) {
      if (c.p[i]) (void)hipFree(c.p[i]);
      c.p[i] = nullptr;
      c.bytes[i] = 0;
    }
  }
  if (cur >= 0) (void)hipSetDevice(cur);
  return HSDL_OK;
}

extern "C" int hsdl_release_scratch(int32_t device) {
  return hsdl::guarded([&] { return release_scratch_body(device); });
}
