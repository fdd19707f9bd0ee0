// gen.hip -- K-round hop-bounded label generation on gfx950.
//
// Algorithm (the set the reference's GPU path and its CPU twin produce):
// Hybrid_Generation_Clean/src/label_gen.cu:477-593 (frontier expansion,
// min-dedupe, rank pruning y > s), :87-110 (accept iff candidate < query),
// :113-150 (append), round driver :603-837; CPU twin
// Hybrid_Generation_Clean/include/HBPLL/hop_constrained_two_hop_labels_generation.h:136-243.
//
//   L(v) = {(v,0,0)}, F_0(s) = {(s,0)}
//   round h = 1..K, every hub s:
//     C(y) = min over (x,dx) in F_{h-1}(s), (y,w) in N(x), y > s of dx + w
//     accept (s,h,C(y)) into L(y) and F_h(s) iff
//       C(y) < min{ da+db : (a,ha,da) in L(y), (a,hb,db) in L(s), ha+hb <= h }
//
// Schedule.  Hubs are processed in batches [s0, s0+B) in id order; a batch
// runs all K rounds before the next starts.  This gives the reference's set:
// in hub s's round-h query a common hub a < s pairs only labels with
// ha, hb <= h-1 (hb = 0 would need a = s, ha = 0 would need a = y > s), so it
// suffices that every hub a < s has finished round h-1, which the batch order
// guarantees; labels of hubs > s never pair (L(s) holds hubs <= s only).
// B = 1 is the plain one-hub-per-launch form.
//
// Per round, all on one stream, no host synchronisation:
//   round_prep  -> zero per-hub counters
//   expand      -> the frontier is grouped by vertex (the labels vertex x
//                  gained in the last round, of many hubs); a wave takes 64
//                  groups, binary-searches each x's id-sorted adjacency for
//                  the first neighbour above the batch's first hub and walks
//                  the (edge, label) pairs edge-major, 64 at a time, so the
//                  lanes of one edge reach the neighbour's slots of nearby
//                  hub slots (slot windows, slot_at) in a few sectors;
//                  candidates are deduplicated by atomicMin on a dense
//                  per-(hub,vertex) slot that keeps the batch's earlier
//                  rounds' minima (tagged per batch attempt); the first
//                  toucher that undercuts them appends (hub,vertex) to a flat
//                  list
//   count       -> per-hub candidate counts of the flat list (LDS hash per
//                  2048-entry chunk: one global atomic per chunk and hub)
//   chunk_scan  -> per-hub offsets and work items of `tile` candidates
//   group       -> flat candidates scattered into per-hub runs (hubs tallied
//                  per 2048-entry chunk in an LDS hash: one global
//                  reservation per chunk and hub)
//   prune       -> persistent: a workgroup of 16 waves takes a work item,
//                  stages the hub's labels L(s) in an LDS hash tile (hub ->
//                  prefix-min dist per hop budget), then the waves take its
//                  candidates one at a time, stream L(y) 128 labels per step
//                  with the next step in flight and stop at the first witness
//   reserve     -> grows label segments that the round's accepted labels
//                  overflow (bump allocation in a device pool; no locks)
//   append      -> writes accepted labels
//   finalize    -> publishes segment lengths; each vertex's new labels
//                  become its frontier group of the next round
// Capacities (label pool, tile arena, candidate and frontier lists) start
// small; a batch that overflows one raises a flag, the later kernels of the
// batch stand down, and the host rolls the batch back, grows the buffer and
// runs the batch again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "device.hpp"

namespace hsdl {

struct Item {  // accepted label of a round: hub slot, vertex, distance
  uint32_t b;
  uint32_t y;
  uint32_t d;
};

// The next round's frontier, grouped by vertex: the labels vertex y gained
// in a round are the cnt labels of its list at pool offset off (appended
// together, so contiguous), and they are expanded together over y's edges.
struct FGroup {
  uint32_t y;
  uint32_t cnt;
  unsigned long long off;
};

enum : uint32_t {
  kOvfPool = 1u,   // label pool
  kOvfArena = 2u,  // global tile arena
  kOvfCand = 4u,   // flat candidate list
  kOvfFront = 8u,  // next frontier list
};

struct Ctrl {
  unsigned long long pool_top;
  unsigned long long cand_total;
  unsigned long long item_total;
  unsigned long long edge_total;
  unsigned long long label_reads;  // labels of L(y) streamed by the prune
  unsigned long long tile_reads;   // labels of L(s) staged into tiles
  unsigned long long row_slots;    // dense slots row items scanned
  uint32_t fcnt[2];
  uint32_t ccnt;  // flat candidates this round
  uint32_t touched_cnt;
  uint32_t work_ctr;
  uint32_t total_chunks;
  uint32_t overflow;
  uint32_t maxlen;
  uint32_t expand_ctr;  // next chunk of 64 frontier groups (k_expand)
  uint32_t gcnt;        // frontier groups of the running round
  uint32_t cand_round;  // candidates the running round's expansion opened
  uint32_t heavy_cnt;   // pieces pushed to the heavy list this round
  uint32_t heavy_ctr;   // next heavy piece (k_expand_heavy)
  uint32_t row_mode;    // this round's prune takes row items (k_prune_rows)
  uint32_t row_ctr[8];  // next row item of each XCD's queue
  // the running batch attempt (k_set_batch): the kernels take these instead
  // of their arguments' s0 / B / tag, so one captured batch replays for any
  uint32_t bs0, bB, btag;
};

// A piece of a large frontier group: its cnt labels (at pool offset off)
// over ecnt consecutive edges from beg, relaxed by a whole workgroup.
struct HeavyItem {
  unsigned long long off;
  unsigned long long beg;
  uint32_t cnt;
  uint32_t ecnt;
};

struct GenArgs {
  uint32_t n;
  uint32_t K;
  uint32_t s0;         // the batch's first entry of the hub sequence
  uint32_t B;
  uint32_t h;
  uint32_t cur;        // frontier buffer read this round
  uint32_t tile;       // candidates per prune work item
  uint32_t row_policy; // prune row items: 0 never, 1 when candidates are dense, 2 always
  uint32_t row_density;  // "dense": candidates >= batch slots / row_density
  uint32_t row_run;      // y-blocks per row item (one tile build per item)
  uint32_t row_block;    // vertices per y-block (a multiple of 1024)
  uint32_t claim_max;    // candidates a prune wave claims at once, at most
  uint32_t idx_tiles;    // LDS tiles of more labels than this are indexed (kInf32: none)
  uint32_t heavy_pairs;  // frontier groups with more pairs go to heavy pieces
  uint32_t piece_pairs;  // pairs per heavy piece (whole edges)
  uint32_t lds_slots;  // LDS tile capacity (slots)
  // hub ids of a group build, ascending (hub slot b of the batch is
  // hubs[s0 + b]); nullptr: every vertex is a hub (slot b is hub s0 + b)
  const uint32_t *hubs;
  const unsigned long long *off;
  const uint2 *adj;
  unsigned long long *pool;
  unsigned long long pool_cap;
  // the pool's labels again as 32-bit words (hub | hop | dist fields, widths
  // below), kept only when the fields fit; the prune streams these
  uint32_t *pool32;
  uint32_t p32_hop_bits;
  uint32_t p32_dist_bits;
  unsigned long long *seg_start;
  uint32_t *seg_len;
  uint32_t *seg_cap;
  // two-zone lists: a vertex's labels [0, sorted_len) are sorted by hop
  // (stably) with hcnt[y * hcnt_stride + x - 1] of them of hop <= x
  // (x = 1..K-1); labels appended since follow in append order.  nullptr:
  // hop zones off (every list is read whole)
  uint32_t *sorted_len;
  uint32_t *hcnt;
  uint32_t hcnt_stride;  // K - 1
  uint32_t *add_cnt;
  uint32_t *touched;
  FGroup *fgrp;               // frontier groups (n entries)
  uint32_t *hub_slot;         // group builds: hub slot of each hub of the batch
  // dense slots, tagged (slot_fresh), in hub windows of
  // 2^slot_lw slots: [Bmax >> slot_lw][n][2^slot_lw] (slot_at)
  uint32_t *cand_d;
  uint32_t slot_lw;
  uint32_t tag;               // epoch of this batch attempt (high bits of a slot)
  uint32_t slot_shift;        // tag << slot_shift | dist << slot_rbits | round
  uint32_t slot_rbits;        // bits of the round field (rounds 1..K)
  unsigned long long *cflat;  // (b << 32 | y), ccap entries
  uint32_t *cgroup;           // y grouped by hub, ccap entries
  uint32_t ccap;
  uint32_t fcap;
  uint32_t *hub_ccnt;
  uint32_t *hub_off;
  uint32_t *hub_cur;
  uint32_t *chunk_off;
  Item *F0;
  Item *F1;
  uint32_t *frank;  // per item of the next frontier: its rank among its vertex's new labels
  HeavyItem *heavy;
  uint32_t heavy_cap;
  uint32_t *gtile;
  unsigned long long gtile_cap;  // u32 words
  Ctrl *ctrl;
};

__device__ inline Item *frontier(const GenArgs &a, uint32_t which) { return which ? a.F1 : a.F0; }

// The batch attempt's first hub slot, size and slot tag, from the control
// block (wave-uniform).
__device__ inline void with_batch(GenArgs &a) {
  a.s0 = __builtin_amdgcn_readfirstlane(a.ctrl->bs0);
  a.B = __builtin_amdgcn_readfirstlane(a.ctrl->bB);
  a.tag = __builtin_amdgcn_readfirstlane(a.ctrl->btag);
}

__global__ void k_set_batch(Ctrl *c, uint32_t s0, uint32_t B, uint32_t tag) {
  c->bs0 = s0;
  c->bB = B;
  c->btag = tag;
}

// Dense slot of (hub slot b, vertex y): the slots of a window of 2^slot_lw
// consecutive hub slots lie side by side for every vertex, so the expansion
// of a frontier group (one vertex's new labels of many hubs) reaches a
// neighbour's slots in a few memory sectors instead of one sector per label.
__device__ inline uint64_t slot_at(const GenArgs &a, uint32_t b, uint32_t y) {
  return ((((uint64_t)(b >> a.slot_lw)) * a.n + y) << a.slot_lw) |
         (b & ((1u << a.slot_lw) - 1u));
}

// vertex id of hub slot b of the running batch
__device__ inline uint32_t hub_of(const GenArgs &a, uint32_t b) {
  return a.hubs ? a.hubs[a.s0 + b] : a.s0 + b;
}

// The positions of a label list a round-h query can use, as one sequence
// j = 0..len-1: the sorted zone's run [a0, a0 + na), then the append zone
// from b0 on.  Every position lies below the list's length.
struct ListView {
  uint32_t a0, na, b0, len;
  __device__ uint32_t at(uint32_t j) const { return j < na ? a0 + j : b0 + (j - na); }
};

// labels of hop <= x in y's sorted zone (the self label is the only hop-0
// label and sits at position 0)
__device__ inline uint32_t hop_prefix(const GenArgs &a, uint32_t y, uint32_t x) {
  return x == 0 ? 1u : a.hcnt[(uint64_t)y * a.hcnt_stride + (x - 1)];
}

// L(y) for candidate (s, y) of round h.  A witness pairs (a, ha, .) in L(y)
// with (a, hb, .) in L(s), ha + hb <= h: a = s pairs only s's own labels of
// earlier rounds of this batch (append zone) with s's self label; a = y would
// need a label of hub y > s in L(s); any other a has ha, hb >= 1.  So the
// query needs y's labels of hop 1..h-1 and its append zone -- and none at
// h = 1.
__device__ inline ListView cand_view(const GenArgs &a, uint32_t len, uint32_t sorted,
                                     uint32_t pre) {
  if (a.h == 1) return ListView{0, 0, 0, 0};
  if (!a.hcnt) return ListView{0, len, len, len};
  return ListView{1, pre - 1, sorted, (pre - 1) + (len - sorted)};
}
// L(s) for the hub tile of round h: labels of hop <= h - 1 (the L(y) side
// brings at least one hop) with the self label, and the append zone.
__device__ inline ListView tile_view(const GenArgs &a, uint32_t s) {
  const uint32_t len = a.seg_len[s];
  if (!a.hcnt) return ListView{0, len, len, len};
  const uint32_t sorted = a.sorted_len[s], pre = hop_prefix(a, s, a.h - 1);
  return ListView{0, pre, sorted, pre + (len - sorted)};
}

// 32-bit labels of the pool copy: hub << (hop_bits + dist_bits) | hop <<
// dist_bits | dist.
__device__ inline uint32_t pack32(const GenArgs &a, uint32_t hub, uint32_t hop, uint32_t dist) {
  return (hub << (a.p32_hop_bits + a.p32_dist_bits)) | (hop << a.p32_dist_bits) | dist;
}
__device__ inline void unpack(const GenArgs &a, uint32_t l, uint32_t &hub, uint32_t &hop,
                              uint32_t &dist) {
  hub = l >> (a.p32_hop_bits + a.p32_dist_bits);
  hop = __builtin_amdgcn_ubfe(l, a.p32_dist_bits, a.p32_hop_bits);
  dist = l & ((1u << a.p32_dist_bits) - 1u);
}
__device__ inline void unpack(const GenArgs &, unsigned long long l, uint32_t &hub,
                              uint32_t &hop, uint32_t &dist) {
  hub = label_hub(l);
  hop = label_hop(l);
  dist = label_dist(l);
}
// the pool a prune streams: LT = uint32_t (pool32) or unsigned long long
// Label i of the pool as (hub, hop, dist): a build that keeps the 32-bit
// copy keeps no 8-byte pool (its labels are the copy's, widened where the
// index needs them).
__device__ inline void pool_label(const GenArgs &a, unsigned long long i, uint32_t &hub,
                                  uint32_t &hop, uint32_t &dist) {
  if (a.pool32)
    unpack(a, a.pool32[i], hub, hop, dist);
  else
    unpack(a, a.pool[i], hub, hop, dist);
}
template <class LT> __device__ inline const LT *stream_pool(const GenArgs &a);
template <> __device__ inline const uint32_t *stream_pool<uint32_t>(const GenArgs &a) {
  return a.pool32;
}
template <> __device__ inline const unsigned long long *stream_pool<unsigned long long>(
    const GenArgs &a) {
  return a.pool;
}

// Dense dedupe slots keep, per (hub slot, vertex), the smallest candidate
// distance of every round of the current batch attempt, with the round that
// set it:  tag << shift | d << rbits | h.
// Older attempts carry larger tags (tags count down; a cleared slot is all
// ones), so any value of the running attempt is below them.  A round-h
// candidate whose distance is not below an earlier round's candidate for the
// same (s, y) is always pruned -- the earlier one was either accepted (then
// (s,0,0) in L(s) with that label is a witness) or pruned by a witness that
// still applies with h >= its hop count -- so the expansion drops it: a
// round-h value undercuts an earlier round's exactly when its distance is
// smaller (at equal distance the later round compares larger), and a slot
// opens for round h when its old value is from another attempt or round.
// The prune leaves the slot as it is.
__device__ inline uint32_t slot_fresh(const GenArgs &a, uint32_t d) {
  return (a.tag << a.slot_shift) | (d << a.slot_rbits) | a.h;
}
__device__ inline uint32_t slot_dist(const GenArgs &a, uint32_t v) {
  return (v & ((1u << a.slot_shift) - 1u)) >> a.slot_rbits;
}
// a candidate of the running round (attempt and round match)
__device__ inline bool slot_current(const GenArgs &a, uint32_t v) {
  return (v >> a.slot_shift) == a.tag && (v & ((1u << a.slot_rbits) - 1u)) == a.h;
}
// first candidate of the running round at a slot whose previous value was `old`
__device__ inline bool slot_opens(const GenArgs &a, uint32_t old) { return !slot_current(a, old); }

// ------------------------------------------------------------------ init
__global__ void k_init_labels(GenArgs a, uint32_t c0) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < a.n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    a.seg_start[v] = v * c0;
    a.seg_len[v] = 1;
    a.seg_cap[v] = c0;
    a.add_cnt[v] = 0;
    if (a.pool) a.pool[v * c0] = pack_label((uint32_t)v, 0, 0);
    if (a.pool32) a.pool32[v * c0] = pack32(a, (uint32_t)v, 0, 0);
    if (a.hcnt) {
      a.sorted_len[v] = 1;
      for (uint32_t x = 0; x < a.hcnt_stride; ++x) a.hcnt[v * a.hcnt_stride + x] = 1;
    }
  }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t val) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = val;
}

// The first frontier: every hub of the batch with its self label (position
// 0 of its list), one group each.
__global__ void k_batch_begin(GenArgs a) {
  with_batch(a);
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < a.B; b += gridDim.x * blockDim.x) {
    const uint32_t s = hub_of(a, b);
    a.fgrp[b] = FGroup{s, 1u, a.seg_start[s]};
    if (a.hubs) a.hub_slot[s] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctrl->fcnt[a.cur] = a.B;
    a.ctrl->gcnt = a.B;
  }
}

__global__ void k_round_prep(GenArgs a) {
  with_batch(a);
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < a.B; b += gridDim.x * blockDim.x) {
    a.hub_ccnt[b] = 0;
    a.hub_cur[b] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctrl->fcnt[a.cur ^ 1] = 0;
    a.ctrl->ccnt = 0;
    a.ctrl->touched_cnt = 0;
    a.ctrl->work_ctr = 0;
    a.ctrl->expand_ctr = 0;
    a.ctrl->heavy_cnt = 0;
    a.ctrl->heavy_ctr = 0;
    a.ctrl->cand_round = 0;
  }
}

// ---------------------------------------------------------------- expand
// XCDs of an MI355X (each has its own L2); workgroups are dealt to them
// round-robin by id.
constexpr uint32_t kXcds = 8;

// first index in adj[beg, end) whose neighbour id exceeds s
__device__ inline unsigned long long first_above(const uint2 *adj, unsigned long long beg,
                                                 unsigned long long end, uint32_t s) {
  while (beg < end) {
    const unsigned long long mid = (beg + end) >> 1;
    if (adj[mid].x <= s) beg = mid + 1; else end = mid;
  }
  return beg;
}

// Frontier groups whose (label, edge) pairs above the batch's first hub
// exceed a.heavy_pairs are cut into pieces of about a.piece_pairs pairs
// (whole edges) for k_expand_heavy, so no wave is left alone with a large
// group while the rest of the GPU idles.

// Flat-list space is reserved per wave in blocks of kListBlock entries (one
// atomic on the list counter per block instead of one per step); unused
// tails are filled with empty entries that the counting and grouping skip.
constexpr uint32_t kListBlock = 512;
constexpr unsigned long long kEmptyEntry = ~0ull;

struct ListRes {
  uint32_t base = 0, used = kListBlock;  // starts exhausted
  uint32_t cands = 0;                    // candidates opened, not yet added to cand_round
};

// Writes empty entries over [base+used, base+kListBlock) (whole wave).
__device__ inline void list_pad(const GenArgs &a, const ListRes &r) {
  for (uint32_t i = r.used + (uint32_t)lane_id(); i < kListBlock; i += kWave)
    if (r.base + i < a.ccap) a.cflat[r.base + i] = kEmptyEntry;
}

// End of a kernel: pad the open block and add the wave's candidate count.
__device__ inline void list_close(const GenArgs &a, const ListRes &r) {
  list_pad(a, r);
  if (lane_id() == 0 && r.cands) atomicAdd(&a.ctrl->cand_round, r.cands);
}

// A frontier label at pool offset `at`: its hub slot, hub and distance.
// `ok` is false for a hub outside the running batch, which a list never
// holds at these offsets; the check keeps every slot access in bounds.
__device__ inline void frontier_label(const GenArgs &a, unsigned long long at, uint32_t &b,
                                      uint32_t &hub, uint32_t &d, bool &ok) {
  uint32_t hop;
  if (a.pool32)
    unpack(a, a.pool32[at], hub, hop, d);
  else
    unpack(a, a.pool[at], hub, hop, d);
  if (a.hubs) {
    b = hub < a.n ? a.hub_slot[hub] : kInf32;
    ok = b < a.B && a.hubs[a.s0 + b] == hub;
  } else {
    b = hub - a.s0;
    ok = hub >= a.s0 && b < a.B;
  }
}

// Relaxes edge ed = (vertex, weight) for hub slot kb (hub kh) at distance
// kd, when valid and the edge's vertex ranks below the hub: atomicMin on the
// tagged slot; first touches that undercut earlier rounds go to the flat list
// (the wave's reserved block).  Called by all lanes of a wave; returns
// whether this lane relaxed an edge.  (The callers load the edge and the
// frontier label independently, so the two loads are in flight together.)
__device__ inline bool relax_edge(const GenArgs &a, bool valid, uint32_t kb, uint32_t kh,
                                  uint32_t kd, uint2 ed, unsigned long long lt_mask,
                                  ListRes &res) {
  const int lane = lane_id();
  bool first = false, relaxed = false;
  uint32_t y = 0;
  if (valid) {
    y = ed.x;
    if (y > kh) {  // rank pruning: hub kh labels only vertices ranked below it
      relaxed = true;
      const uint32_t nv = slot_fresh(a, kd + ed.y);
      uint32_t *slot = &a.cand_d[slot_at(a, kb, y)];
      // A slot only falls while the expansion runs (atomicMin; clears happen
      // between batches), so a read at or below nv, even a stale one, means
      // the atomic would neither change the slot nor open it: most
      // relaxations end with this read, which costs less than a returning
      // atomic.
      if (nv < *slot) {
        const uint32_t old = atomicMin(slot, nv);
        first = nv < old && slot_opens(a, old);
      }
    }
  }
  const unsigned long long firsts = __ballot(first);
  if (firsts) {
    const uint32_t cnt = (uint32_t)__popcll(firsts);
    if (res.used + cnt > kListBlock) {
      list_pad(a, res);
      uint32_t nb = 0;
      if (lane == 0) nb = atomicAdd(&a.ctrl->ccnt, kListBlock);
      res.base = __shfl(nb, 0, kWave);
      res.used = 0;
    }
    if (first) {
      const uint32_t pos = res.base + res.used + (uint32_t)__popcll(firsts & lt_mask);
      if (pos < a.ccap)
        a.cflat[pos] = ((unsigned long long)kb << 32) | y;
      else
        atomicOr(&a.ctrl->overflow, kOvfCand);
    }
    res.used += cnt;
    res.cands += cnt;
  }
  return relaxed;
}

// Round h's expansion.  A wave takes 64 frontier groups (vertex x and the
// labels x gained in round h-1, of many hubs) and walks the pairs (edge of x
// above the batch's first hub, label of the group) edge-major: consecutive
// lanes take the labels of one group on one edge, so they share the edge's
// load and reach the neighbour's slots of nearby hub slots, side by side in
// the slot windows.  Lanes whose label's hub ranks above the neighbour do no
// work (rank pruning).
__global__ void __launch_bounds__(256) k_expand(GenArgs a) {
  with_batch(a);
  if (a.ctrl->overflow) return;  // the batch is rolled back and run again
  const FGroup *G = a.fgrp;
  const uint32_t gcnt = min(a.ctrl->gcnt, a.n);
  const uint32_t lo_hub = hub_of(a, 0);
  const int lane = lane_id();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t edges_seen = 0;
  ListRes res;
  const uint32_t wave_id = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  for (bool first = true;; first = false) {
    // chunks of 64 groups: a wave's first chunk is its own id and later ones
    // come from a counter (waves with nothing to do never touch it; dealing
    // the chunks to XCDs in contiguous ranges instead cost every wave its
    // claims even in small rounds: the 1024x1024 grid's expansion took
    // 0.258 s that way against 0.042 s)
    uint32_t chunk = wave_id;
    if (!first) {
      if (lane == 0) chunk = nwaves + atomicAdd(&a.ctrl->expand_ctr, 1u);
      chunk = __shfl(chunk, 0, kWave);
    }
    const uint64_t base = (uint64_t)chunk * kWave;
    if (base >= gcnt) break;
    const uint64_t i = base + lane;
    uint32_t gc = 0, pairs = 0;
    unsigned long long goff = 0, beg = 0, deg = 0;
    bool kept = false;  // a large group this wave expands itself (heavy list full)
    if (i < gcnt) {
      const FGroup gr = G[i];
      gc = gr.cnt;
      goff = gr.off;
      const unsigned long long e0 = a.off[gr.y], e1 = a.off[gr.y + 1];
      beg = first_above(a.adj, e0, e1, lo_hub);
      deg = gc ? e1 - beg : 0;
      const unsigned long long pr = (unsigned long long)gc * deg;
      if (pr > a.heavy_pairs) {
        const uint32_t ep = max(1u, a.piece_pairs / gc);  // edges per piece
        const unsigned long long np = (deg + ep - 1) / ep;
        uint32_t pos = kInf32;
        if (np < a.heavy_cap) pos = atomicAdd(&a.ctrl->heavy_cnt, (uint32_t)np);
        if (pos != kInf32 && pos + np <= a.heavy_cap) {
          for (uint32_t j = 0; j < (uint32_t)np; ++j) {
            const unsigned long long eb = (unsigned long long)j * ep;
            a.heavy[pos + j] =
                HeavyItem{goff, beg + eb, gc, (uint32_t)min((unsigned long long)ep, deg - eb)};
          }
        } else {
          // no room: this wave keeps the group; reserved slots below the
          // cap are written as empty pieces
          if (pos != kInf32)
            for (uint32_t j = 0; pos + j < a.heavy_cap && j < (uint32_t)np; ++j)
              a.heavy[pos + j] = HeavyItem{0ull, 0ull, 0u, 0u};
          kept = true;
        }
      } else {
        pairs = (uint32_t)pr;
      }
    }
    const uint32_t incl = wave_incl_scan_u32(pairs);
    const uint32_t excl = incl - pairs;
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    for (uint32_t jb = 0; jb < total; jb += kWave) {
      const uint32_t j = jb + lane;
      // largest k with excl_k <= j (groups without pairs are skipped over)
      int k = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const uint32_t e = __shfl(excl, k + step, kWave);
        if (e <= j) k += step;
      }
      const uint32_t kc = max(1u, (uint32_t)__shfl(gc, k, kWave));
      const uint32_t kex = __shfl(excl, k, kWave);
      const unsigned long long kbeg = __shfl(beg, k, kWave);
      const unsigned long long koff = __shfl(goff, k, kWave);
      const uint32_t p = j - kex;
      const uint32_t q = p / kc, r = p - q * kc;  // edge q, label r of group k
      uint32_t b = 0, hub = 0, d = 0;
      bool ok = false;
      uint2 ed = make_uint2(0u, 0u);
      if (j < total) {
        ed = a.adj[kbeg + q];
        frontier_label(a, koff + r, b, hub, d, ok);
      }
      edges_seen += relax_edge(a, j < total && ok, b, hub, d, ed, lt_mask, res);
    }
    // large groups the heavy list had no room for, one at a time
    unsigned long long keep = __ballot(kept);
    while (keep) {
      const int l = __ffsll((long long)keep) - 1;
      keep &= keep - 1;
      const uint32_t kc = __shfl(gc, l, kWave);
      const unsigned long long kbeg = __shfl(beg, l, kWave);
      const unsigned long long koff = __shfl(goff, l, kWave);
      const unsigned long long kpairs = (unsigned long long)kc * __shfl(deg, l, kWave);
      for (unsigned long long pb = 0; pb < kpairs; pb += kWave) {
        const unsigned long long p = pb + lane;
        const unsigned long long q = p / kc;
        uint32_t b = 0, hub = 0, d = 0;
        bool ok = false;
        uint2 ed = make_uint2(0u, 0u);
        if (p < kpairs) {
          ed = a.adj[kbeg + q];
          frontier_label(a, koff + (p - q * kc), b, hub, d, ok);
        }
        edges_seen += relax_edge(a, p < kpairs && ok, b, hub, d, ed, lt_mask, res);
      }
    }
  }
  list_close(a, res);
  edges_seen = wave_sum_u64(edges_seen);
  if (lane == 0 && edges_seen) atomicAdd(&a.ctrl->edge_total, (unsigned long long)edges_seen);
}

// Heavy pieces: a workgroup takes one piece at a time and spreads its pairs
// over its 256 lanes, edge-major like k_expand.
__global__ void __launch_bounds__(256) k_expand_heavy(GenArgs a) {
  with_batch(a);
  __shared__ uint32_t sh_h;
  if (a.ctrl->overflow) return;
  const int lane = lane_id();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t wv = threadIdx.x / kWave;
  const uint32_t cnt = min(a.ctrl->heavy_cnt, a.heavy_cap);
  if (blockIdx.x >= cnt) return;  // uniform: most launches find no pieces
  ListRes res;
  uint64_t edges_seen = 0;
  for (bool first = true;; first = false) {
    // a workgroup's first piece is its own id and later ones come from a
    // counter
    if (threadIdx.x == 0)
      sh_h = first ? blockIdx.x : gridDim.x + atomicAdd(&a.ctrl->heavy_ctr, 1u);
    __syncthreads();
    const uint32_t h = sh_h;
    __syncthreads();
    if (h >= cnt) break;
    const HeavyItem it = a.heavy[h];
    // pairs <= max(piece_pairs, cnt) < 2^32; whole waves iterate together so
    // the ballots see every lane
    const uint32_t kc = max(1u, it.cnt);
    const uint32_t pairs = it.cnt * it.ecnt;
    for (uint32_t pb = wv * kWave; pb < pairs; pb += 256) {
      const uint32_t p = pb + (uint32_t)lane;
      const uint32_t q = p / kc;
      uint32_t b = 0, hub = 0, d = 0;
      bool ok = false;
      uint2 ed = make_uint2(0u, 0u);
      if (p < pairs) {
        ed = a.adj[it.beg + q];
        frontier_label(a, it.off + (p - q * kc), b, hub, d, ok);
      }
      edges_seen += relax_edge(a, p < pairs && ok, b, hub, d, ed, lt_mask, res);
    }
  }
  list_close(a, res);
  edges_seen = wave_sum_u64(edges_seen);
  if (lane == 0 && edges_seen) atomicAdd(&a.ctrl->edge_total, (unsigned long long)edges_seen);
}

// ------------------------------------------------------------ chunk scan
// A round's prune takes row items (k_prune_rows) when its candidates fill at
// least 1/row_density of the batch's dense slots (default 32: the row scan
// then costs at most 128 B per candidate, against about 1.5 KB of list reads).
// Row items: XCDs of an MI355X (each has its own L2).
// (vertices per row block: a.row_block, a multiple of the 1024 a workgroup
// scans at a time)

// Row items when the round's candidates fill at least 1/row_density of the
// batch's slot rows (or always / never by the policy), and only while an XCD
// queue's item count fits its 32-bit counter.  A batch of one hub has no list
// reads to share: the plain form never takes row items unless forced.
__device__ inline bool rows_chosen(const GenArgs &a) {
  const uint32_t cands = a.ctrl->cand_round;
  const uint64_t queue_items =
      (uint64_t)a.B * (((uint64_t)a.n + (uint64_t)a.row_block * kXcds - 1) /
                       ((uint64_t)a.row_block * kXcds));
  return cands && queue_items < (1ull << 31) &&
         (a.row_policy == 2 || (a.row_policy == 1 && a.B > 1 &&
                                (uint64_t)cands * a.row_density >= (uint64_t)a.B * a.n));
}

// Per-hub candidate counts of a round that takes list items: a workgroup
// takes 2048 consecutive entries of the flat list, counts their hubs in an
// LDS hash and adds each hub's count with one global atomic.
constexpr uint32_t kGroupThreads = 256;
constexpr uint32_t kGroupPer = 8;                                 // entries per thread
constexpr uint32_t kGroupChunk = kGroupThreads * kGroupPer;       // 2048
constexpr uint32_t kGroupSlots = 2 * kGroupChunk;                 // load factor <= 1/2
static_assert(kGroupSlots == 4096, "hubs hash to 12 bits");

// LDS hash slot of hub slot b (inserted when absent)
__device__ inline uint32_t group_slot(uint32_t *hkey, uint32_t b) {
  uint32_t sl = (b * 0x9E3779B1u) >> 20;  // 12 bits: kGroupSlots = 4096
  for (;;) {
    const uint32_t prev = atomicCAS(&hkey[sl], kInf32, b);
    if (prev == kInf32 || prev == b) return sl;
    sl = (sl + 1) & (kGroupSlots - 1);
  }
}

__global__ void __launch_bounds__(kGroupThreads) k_count(GenArgs a) {
  with_batch(a);
  __shared__ uint32_t hkey[kGroupSlots];
  __shared__ uint32_t hcnt[kGroupSlots];
  if (a.ctrl->overflow || rows_chosen(a)) return;
  // reserved blocks may reach past the list's end; only empty entries lie there
  const uint32_t cnt = min(a.ctrl->ccnt, a.ccap);
  const uint32_t tid = threadIdx.x;
  for (uint64_t c0 = (uint64_t)blockIdx.x * kGroupChunk; c0 < cnt;
       c0 += (uint64_t)gridDim.x * kGroupChunk) {
    for (uint32_t i = tid; i < kGroupSlots; i += kGroupThreads) {
      hkey[i] = kInf32;
      hcnt[i] = 0;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kGroupPer; ++r) {
      const uint64_t i = c0 + r * kGroupThreads + tid;
      const unsigned long long e = i < cnt ? a.cflat[i] : kEmptyEntry;
      if (e != kEmptyEntry) atomicAdd(&hcnt[group_slot(hkey, (uint32_t)(e >> 32))], 1u);
    }
    __syncthreads();
    for (uint32_t i = tid; i < kGroupSlots; i += kGroupThreads)
      if (hcnt[i]) atomicAdd(&a.hub_ccnt[hkey[i]], hcnt[i]);
    __syncthreads();
  }
}

// One workgroup: hub_off = exclusive scan of hub_ccnt; chunk_off[b] =
// sum_{b'<b} ceil(hub_ccnt[b'] / tile).  An overflowed batch gets no work.
// Each wave scans a contiguous range 64 entries per step (coalesced loads,
// a wave scan per step, the carry in a register) after the waves' totals
// give it its start.
__global__ void __launch_bounds__(1024) k_chunk_scan(GenArgs a) {
  with_batch(a);
  __shared__ uint32_t wsum[1024 / kWave], wcsum[1024 / kWave];
  const uint32_t t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const uint32_t nw = blockDim.x / kWave;
  const uint32_t B = a.B, tile = a.tile;
  const uint32_t per_w = ((B + nw - 1) / nw + kWave - 1) & ~(kWave - 1);
  const uint32_t lo = min(B, w * per_w), hi = min(B, lo + per_w);
  uint32_t s = 0, cs = 0;
  for (uint32_t b = lo + lane; b < hi; b += kWave) {
    const uint32_t c = a.hub_ccnt[b];
    s += (c + tile - 1) / tile;
    cs += c;
  }
  s = wave_sum_u32(s);
  cs = wave_sum_u32(cs);
  if (lane == 0) {
    wsum[w] = s;
    wcsum[w] = cs;
  }
  __syncthreads();
  uint32_t run = 0, runc = 0;
  for (uint32_t x = 0; x < w; ++x) {
    run += wsum[x];
    runc += wcsum[x];
  }
  for (uint32_t b0 = lo; b0 < hi; b0 += kWave) {
    const uint32_t b = b0 + lane;
    const uint32_t c = b < hi ? a.hub_ccnt[b] : 0u;
    const uint32_t k = (c + tile - 1) / tile;
    const uint32_t ik = wave_incl_scan_u32(k), ic = wave_incl_scan_u32(c);
    if (b < hi) {
      a.chunk_off[b] = run + ik - k;
      a.hub_off[b] = runc + ic - c;
    }
    run += (uint32_t)__shfl((int)ik, kWave - 1, kWave);
    runc += (uint32_t)__shfl((int)ic, kWave - 1, kWave);
  }
  if (t < 8) a.ctrl->row_ctr[t] = 0;
  if (t == blockDim.x - 1) {
    // the last wave's carry ends at the totals (its range, possibly empty,
    // is the last)
    const bool dead = a.ctrl->overflow != 0;
    const bool rows = !dead && rows_chosen(a);
    a.chunk_off[B] = run;
    a.hub_off[B] = runc;
    a.ctrl->row_mode = rows ? 1u : 0u;
    a.ctrl->total_chunks = (dead || rows) ? 0u : run;
    a.ctrl->work_ctr = 0;
    a.ctrl->cand_total += a.ctrl->cand_round;
    a.ctrl->item_total += min(a.ctrl->fcnt[a.cur], a.fcap);
  }
}

// Scatter of the flat candidates into per-hub runs.  A workgroup takes 2048
// consecutive entries, counts their hubs in an LDS hash, reserves each hub's
// space with one global atomic and writes the entries: one global atomic per
// (chunk, hub) instead of one per (wave, hub), so hubs with millions of
// candidates do not serialize the scatter on their counters.

__global__ void __launch_bounds__(kGroupThreads) k_group(GenArgs a) {
  with_batch(a);
  __shared__ uint32_t hkey[kGroupSlots];
  __shared__ uint32_t hcnt[kGroupSlots];
  __shared__ uint32_t hbase[kGroupSlots];
  if (a.ctrl->overflow || a.ctrl->row_mode) return;
  // reserved blocks may reach past the list's end; only empty entries lie there
  const uint32_t cnt = min(a.ctrl->ccnt, a.ccap);
  const uint32_t tid = threadIdx.x;
  for (uint64_t c0 = (uint64_t)blockIdx.x * kGroupChunk; c0 < cnt;
       c0 += (uint64_t)gridDim.x * kGroupChunk) {
    for (uint32_t i = tid; i < kGroupSlots; i += kGroupThreads) {
      hkey[i] = kInf32;
      hcnt[i] = 0;
    }
    __syncthreads();
    uint32_t slot[kGroupPer], yv[kGroupPer];
#pragma unroll
    for (uint32_t r = 0; r < kGroupPer; ++r) {
      const uint64_t i = c0 + r * kGroupThreads + tid;
      slot[r] = kInf32;
      yv[r] = 0;
      const unsigned long long e = i < cnt ? a.cflat[i] : kEmptyEntry;
      if (e != kEmptyEntry) {
        yv[r] = (uint32_t)e;
        const uint32_t sl = group_slot(hkey, (uint32_t)(e >> 32));
        slot[r] = sl;
        atomicAdd(&hcnt[sl], 1u);
      }
    }
    __syncthreads();
    for (uint32_t i = tid; i < kGroupSlots; i += kGroupThreads) {
      const uint32_t c = hcnt[i];
      if (c) {
        const uint32_t b = hkey[i];
        hbase[i] = a.hub_off[b] + atomicAdd(&a.hub_cur[b], c);
        hcnt[i] = 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kGroupPer; ++r) {
      if (slot[r] == kInf32) continue;
      const uint32_t rank = atomicAdd(&hcnt[slot[r]], 1u);
      a.cgroup[hbase[slot[r]] + rank] = yv[r];
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- prune
// Sixteen wavefronts share one tile: with a ~60 KB tile two workgroups fit a
// CU, which keeps 32 waves (the maximum) streaming label lists per CU.
constexpr uint32_t kPruneThreads = 1024;

// The tile of L(s) is the hash tile of device.hpp (hub -> per-hop prefix
// minimum distance).

// LDS atomicMin on 16-bit values: CAS on the containing word.
// The word is reached by pointer arithmetic on p, not through an integer, so
// an LDS pointer stays an LDS pointer (ds_cmpst, not a flat atomic).
__device__ inline void atomic_min_v(uint16_t *p, uint32_t v) {
  const uint32_t odd = (uint32_t)(reinterpret_cast<uintptr_t>(p) >> 1) & 1u;
  uint32_t *w = reinterpret_cast<uint32_t *>(p - odd);
  const uint32_t sh = odd ? 16 : 0;
  uint32_t old = *w;
  for (;;) {
    const uint32_t cur = (old >> sh) & 0xFFFFu;
    if (cur <= v) return;
    const uint32_t nw = (old & ~(0xFFFFu << sh)) | (v << sh);
    const uint32_t got = atomicCAS(w, old, nw);
    if (got == old) return;
    old = got;
  }
}
__device__ inline void atomic_min_v(uint32_t *p, uint32_t v) { atomicMin(p, v); }

// A prune tile of round h holds one record per slot: the hub w (the key) and
// K + 1 values, value ha = the smallest distance of w's labels in L(s) of hop
// <= h - ha (empty when ha > h): the value a label (w, ha, da) of L(y) pairs
// with.  Records are 2^rec_shift words (16 bytes for K <= 5 and 16-bit
// values), so a probe reaches the key with one shift-add from the slot and
// the value at ha * sizeof(V) bytes past the key's end.
template <class V> __host__ __device__ inline uint32_t rec_shift(uint32_t K1) {
  const uint32_t w = 1u + (K1 * (uint32_t)sizeof(V) + 3u) / 4u;
  uint32_t sh = 0;
  while ((1u << sh) < w) ++sh;
  return sh;
}
template <class V> __host__ __device__ inline uint32_t rec_words(uint32_t K1) {
  return 1u << rec_shift<V>(K1);
}
template <class V> __host__ __device__ inline uint64_t rec_tile_words(uint32_t cap, uint32_t K1) {
  return (uint64_t)cap * rec_words<V>(K1);
}
// ha * sizeof(V) for label l of hop ha, straight from its hop field
template <class V> __device__ inline uint32_t hop_off(const GenArgs &a, uint32_t l) {
  constexpr uint32_t vs = sizeof(V) == 2 ? 1u : 2u;  // 32-bit values: dist_bits >= 16
  return (l >> (a.p32_dist_bits - vs)) & (((1u << a.p32_hop_bits) - 1u) << vs);
}
template <class V> __device__ inline uint32_t hop_off(const GenArgs &, unsigned long long l) {
  constexpr uint32_t vs = sizeof(V) == 2 ? 1u : 2u;
  return label_hop(l) << vs;
}

template <class V>
__device__ void build_tile(const GenArgs &a, uint32_t s, const ListView &view, uint32_t *rec,
                           uint32_t cap) {
  const uint32_t K1 = a.K + 1;
  const uint32_t stride = rec_words<V>(K1);
  // empty keys and empty values are all ones: the tile is filled 16 bytes
  // per store where it is so aligned (a row item builds a tile for every few
  // dozen candidates, and filling an LDS tile of 4096 slots a 16-bit value
  // at a time cost more than the candidates' probes)
  const uint64_t words = rec_tile_words<V>(cap, K1);
  if ((reinterpret_cast<uintptr_t>(rec) & 15u) == 0) {
    uint4 *q = reinterpret_cast<uint4 *>(rec);
    const uint4 ones = make_uint4(kInf32, kInf32, kInf32, kInf32);
    for (uint32_t i = threadIdx.x; i < words / 4; i += blockDim.x) q[i] = ones;
    for (uint32_t i = (uint32_t)(words & ~3ull) + threadIdx.x; i < words; i += blockDim.x)
      rec[i] = kInf32;
  } else {
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) rec[i] = kInf32;
  }
  __syncthreads();
  const unsigned long long st = a.seg_start[s];
  for (uint32_t i = threadIdx.x; i < view.len; i += blockDim.x) {
    uint32_t hub, hb, db;
    pool_label(a, st + view.at(i), hub, hb, db);
    uint32_t slot = hash_slot(hub, cap);
    for (;;) {
      const uint32_t prev = atomicCAS(&rec[slot * stride], kInf32, hub);
      if (prev == kInf32 || prev == hub) break;
      slot = (slot + 1) & (cap - 1);
    }
    V *vals = reinterpret_cast<V *>(rec + slot * stride + 1);
    for (uint32_t ha = 0; ha + hb <= a.h; ++ha) atomic_min_v(&vals[ha], db);
  }
  __syncthreads();
}


// True when some common hub gives da + db <= dy (candidate pruned); `reads`
// receives the number of labels of L(y) examined (algorithmic: up to the
// 128-label step that holds the witness).  L(y) is streamed 128 labels per
// step (two per lane) with the next step's loads already in flight, so a wave
// waits about one memory round trip per 128 labels, not per 64.  The tile is
// read through address-space pointers (KP / VP, device.hpp), so a probe's
// wait does not also wait for the next step's loads.
//
// The probe is straight-line code but for the rare collision chain, since
// the prune issues VALU instructions near the rate a CU sustains:
// * lanes past the end of the list hold a copy of its last label, which
//   gives the same answer as the label itself, so they need no mask;
// * a label of hop ha > h reads its record's value ha, which is empty;
// * the chain ends at the hub or at an empty slot, whose values are empty,
//   so the value read needs no select: a 16-bit empty value (0xFFFF) is
//   above every candidate distance (small values: K * max weight < 0xFFFF),
//   and a 32-bit one is tested.
template <class V>
__device__ inline bool within(uint32_t da, uint32_t v, uint32_t dy) {
  if (sizeof(V) == 2) return da + v <= dy;
  return v != TileVal<V>::kEmpty && da + v <= dy;
}
// LDS tile: addresses as 32-bit LDS byte offsets, so the value's address is
// the key's plus the hop offset (the record's base is computed once)
template <class V, class LT>
__device__ inline bool probe_label(const GenArgs &a, LT l, uint32_t dy,
                                   lds_ptr<const uint32_t> rec, uint32_t cap, uint32_t sh) {
  uint32_t hub, ha, da;
  unpack(a, l, hub, ha, da);
  const uint32_t base = (uint32_t)(uintptr_t)rec;
  const uint32_t rsh = sh + 2;
  uint32_t slot = hash_slot(hub, cap);
  uint32_t ka = base + (slot << rsh);
  uint32_t k = *(lds_ptr<const uint32_t>)(uintptr_t)ka;
  while (k != hub && k != kInf32) {
    slot = (slot + 1) & (cap - 1);
    ka = base + (slot << rsh);
    k = *(lds_ptr<const uint32_t>)(uintptr_t)ka;
  }
  const uint32_t v = *(lds_ptr<const V>)(uintptr_t)(ka + 4u + hop_off<V>(a, l));
  return within<V>(da, v, dy);
}
// arena tile
template <class V, class LT>
__device__ inline bool probe_label(const GenArgs &a, LT l, uint32_t dy,
                                   glb_ptr<const uint32_t> rec, uint32_t cap, uint32_t sh) {
  uint32_t hub, ha, da;
  unpack(a, l, hub, ha, da);
  uint32_t slot = hash_slot(hub, cap);
  glb_ptr<const uint32_t> r = rec + (slot << sh);
  uint32_t k = *r;
  while (k != hub && k != kInf32) {
    slot = (slot + 1) & (cap - 1);
    r = rec + (slot << sh);
    k = *r;
  }
  const uint32_t v = *(glb_ptr<const V>)((glb_ptr<const char>)(r + 1) + hop_off<V>(a, l));
  return within<V>(da, v, dy);
}

// Indexed LDS tiles, for views of more than kRecSmall labels (graphs of at
// most 2^20 vertices): key slots hold hub << kIdxBits | row (0: empty) and
// the values sit in one row of K + 1 per distinct hub (row 0: all empty), so
// 4-byte key slots run at a lower load factor than 16-byte records in the
// same LDS: a wave's probe chain is its longest lane's, and at scale 20 most
// record tiles ran at load factors of 1/4 to 1/2.  An empty slot reads hub
// 0 and row 0, so a lookup of an absent hub 0 ends at the empty row, and no
// mask is needed (hub 0's own entry has a row >= 1 and lies before any
// empty slot of its chain).
constexpr uint32_t kIdxBits = 12;
constexpr uint32_t kIdxRowMask = (1u << kIdxBits) - 1;
constexpr uint32_t kIdxMaxRows = kIdxRowMask;  // rows 1..4094 for hubs, 0 empty
constexpr uint32_t kRecSmall = 511;            // records reach load factor 1/8 up to here
struct IdxTile {
  lds_ptr<const uint32_t> keys;
  uint32_t rows;     // LDS byte address of row 0
  uint32_t row_bytes;
};
// key slots for an indexed tile of `len` labels in `lds_bytes`, or 0 when it
// does not fit (rows for up to len distinct hubs plus the empty row)
__device__ inline uint32_t idx_tile_cap(uint32_t len, uint32_t row_bytes, uint32_t lds_bytes) {
  if (len + 1 > kIdxMaxRows) return 0;
  const uint32_t rows = ((len + 1) * row_bytes + 15u) & ~15u;
  uint32_t cap = pow2_at_least(8ull * len + 4);
  while (cap >= 4 && 4u * cap + rows > lds_bytes) cap >>= 1;
  return cap >= tile_cap_for(len) ? cap : 0u;
}
template <class V, class LT>
__device__ inline bool probe_label(const GenArgs &a, LT l, uint32_t dy, IdxTile t, uint32_t cap,
                                   uint32_t) {
  uint32_t hub, ha, da;
  unpack(a, l, hub, ha, da);
  uint32_t slot = hash_slot(hub, cap);
  uint32_t k = t.keys[slot];
  while ((k >> kIdxBits) != hub && k != 0u) {
    slot = (slot + 1) & (cap - 1);
    k = t.keys[slot];
  }
  const uint32_t addr = t.rows + mul_u24(k & kIdxRowMask, t.row_bytes) + hop_off<V>(a, l);
  return within<V>(da, *(lds_ptr<const V>)(uintptr_t)addr, dy);
}
// Whole workgroup: keys[0, cap) and rows after them (rows in the same LDS);
// sh_rows: a workgroup counter.  Claims first (a hub's first claimer takes
// the next row and empties it), then the per-hop minima by lookup.
template <class V>
__device__ void build_idx_tile(const GenArgs &a, uint32_t s, const ListView &view, uint32_t *keys,
                               uint32_t cap, uint32_t &sh_rows) {
  const uint32_t K1 = a.K + 1;
  V *rows = reinterpret_cast<V *>(keys + cap);
  for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) keys[i] = 0u;
  if (threadIdx.x == 0) sh_rows = 0;
  for (uint32_t x = threadIdx.x; x < K1; x += blockDim.x) rows[x] = (V)TileVal<V>::kEmpty;
  __syncthreads();
  const unsigned long long st = a.seg_start[s];
  for (uint32_t i = threadIdx.x; i < view.len; i += blockDim.x) {
    uint32_t hub, hb, db;
    pool_label(a, st + view.at(i), hub, hb, db);
    uint32_t slot = hash_slot(hub, cap);
    for (;;) {
      const uint32_t prev = atomicCAS(&keys[slot], 0u, (hub << kIdxBits) | kIdxRowMask);
      if (prev == 0u) {
        const uint32_t row = atomicAdd(&sh_rows, 1u) + 1u;
        for (uint32_t x = 0; x < K1; ++x) rows[row * K1 + x] = (V)TileVal<V>::kEmpty;
        keys[slot] = (hub << kIdxBits) | row;
        break;
      }
      if ((prev >> kIdxBits) == hub) break;
      slot = (slot + 1) & (cap - 1);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < view.len; i += blockDim.x) {
    uint32_t hub, hb, db;
    pool_label(a, st + view.at(i), hub, hb, db);
    uint32_t slot = hash_slot(hub, cap);
    uint32_t k = keys[slot];
    while ((k >> kIdxBits) != hub) {
      slot = (slot + 1) & (cap - 1);
      k = keys[slot];
    }
    V *vals = rows + (k & kIdxRowMask) * K1;
    for (uint32_t ha = 0; ha + hb <= a.h; ++ha) atomic_min_v(&vals[ha], db);
  }
  __syncthreads();
}

// Label `lane` of a list view, clamped to it (a lane past the end reloads
// the last label); an empty view reads the list's position 0, its self label.
template <class LT>
__device__ inline LT first_step(const LT *L, const ListView &v) {
  const uint32_t lane = (uint32_t)lane_id();
  return L[v.len ? v.at(min(lane, v.len - 1)) : 0u];
}

// Witness search over L(y) = L[0, len): true when some common hub gives
// da + db <= dy; f0 holds the list's first 64 labels, loaded by the caller
// ahead of time (first_step).  The prune moves HBM bytes near the rate the
// memory sustains, so loads are speculative only where they pay: witnesses
// mostly sit among the first, top-ranked hubs, so labels [64, 192) are
// loaded only once the first 64 hold none, and only lists still unresolved
// after 192 labels (mostly candidates that will be accepted and read to the
// end) keep the next 128-label step in flight while one is probed.  No load
// sits behind a branch and the step buffers alternate (a, b) instead of
// being copied, so a wait leaves the step in flight alone.
template <class V, class KP, class LT>
__device__ bool witness_search(const GenArgs &a, const LT *L, const ListView &view, LT f0,
                               uint32_t dy, KP rec, uint32_t cap, uint32_t sh,
                               uint32_t &reads) {
  const uint32_t lane = (uint32_t)lane_id();
  const uint32_t len = view.len;
  if (len == 0) {
    reads = 0;
    return false;
  }
  const uint32_t last = len - 1;
  auto ld = [&](uint32_t j) { return L[view.at(min(j, last))]; };
  auto probe2 = [&](LT l0, LT l1, uint32_t j) {
    const bool h0 = probe_label<V>(a, l0, dy, rec, cap, sh);
    const bool h1 = probe_label<V>(a, l1, dy, rec, cap, sh);
    return __ballot(h0 || h1) != 0;
  };
  constexpr uint32_t kStep = 2 * kWave;
  if (__ballot(probe_label<V>(a, f0, dy, rec, cap, sh))) {
    reads = min(len, (uint32_t)kWave);
    return true;
  }
  if (len <= kWave) {
    reads = len;
    return false;
  }
  {
    const LT s0 = ld(kWave + lane), s1 = ld(kWave + kWave + lane);
    if (probe2(s0, s1, kWave)) {
      reads = min(len, kWave + kStep);
      return true;
    }
  }
  constexpr uint32_t kLong = kWave + kStep;  // 192
  if (len <= kLong) {
    reads = len;
    return false;
  }
  LT b0 = ld(kLong + lane), b1 = ld(kLong + kWave + lane);
  for (uint32_t j = kLong; j < len; j += 2 * kStep) {
    // b holds labels [j, j + kStep); a is loaded with [j + kStep, j + 2 kStep)
    const LT a0 = ld(j + kStep + lane), a1 = ld(j + kStep + kWave + lane);
    if (probe2(b0, b1, j)) {
      reads = min(len, j + kStep);
      return true;
    }
    if (j + kStep >= len) break;
    b0 = ld(j + 2 * kStep + lane);
    b1 = ld(j + 2 * kStep + kWave + lane);
    if (probe2(a0, a1, j + kStep)) {
      reads = min(len, j + 2 * kStep);
      return true;
    }
  }
  reads = len;
  return false;
}

// Accepted labels staged in LDS by a prune workgroup (one work item holds at
// most `tile` <= kAccSlots candidates).
constexpr uint32_t kAccSlots = 256;

// Publishes the staged accepted labels (whole workgroup; the caller has
// synchronised after staging): per-vertex add counts, the vertices touched
// for the first time this round gathered in LDS, one reservation each on the
// frontier and the touched list, then the writes.  Empties the staging.
__device__ inline void publish_accepted(const GenArgs &a, Item *Fn, Item *sh_acc,
                                        uint32_t *sh_rank, uint32_t *sh_touch, uint32_t &sh_acc_n,
                                        uint32_t &sh_acc_base, uint32_t &sh_touch_n,
                                        uint32_t &sh_touch_base) {
  const uint32_t nacc = sh_acc_n;
  for (uint32_t i = threadIdx.x; i < nacc; i += blockDim.x) {
    const uint32_t y = sh_acc[i].y;
    // the count before this label is its rank among y's new labels (the
    // append's slot past y's list)
    const uint32_t r = atomicAdd(&a.add_cnt[y], 1u);
    if (r == 0) sh_touch[atomicAdd(&sh_touch_n, 1u)] = y;
    sh_rank[i] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (nacc) {
      sh_acc_base = atomicAdd(&a.ctrl->fcnt[a.cur ^ 1], nacc);
      if (sh_acc_base + nacc > a.fcap) atomicOr(&a.ctrl->overflow, kOvfFront);
    }
    if (sh_touch_n) sh_touch_base = atomicAdd(&a.ctrl->touched_cnt, sh_touch_n);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nacc; i += blockDim.x)
    if (sh_acc_base + i < a.fcap) {
      Fn[sh_acc_base + i] = sh_acc[i];
      a.frank[sh_acc_base + i] = sh_rank[i];
    }
  const uint32_t ntouch = sh_touch_n;
  for (uint32_t i = threadIdx.x; i < ntouch; i += blockDim.x)
    a.touched[sh_touch_base + i] = sh_touch[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    sh_acc_n = 0;
    sh_touch_n = 0;
  }
  __syncthreads();
}

// Candidates a prune wave claims at once, at most, and the divisor that sizes
// a claim from the item's unclaimed candidates (claims shrink toward the
// item's end, where the slowest wave decides when the workgroup moves on).
constexpr uint32_t kClaimSpread = 2 * (kPruneThreads / kWave);

// The waves of a prune workgroup take the item's candidates [c_beg, c_end) of
// hub slot b (L(y) lengths vary widely, so dynamically) and stage the
// accepted ones in LDS.  A wave claims up to a.claim_max candidates at once and
// loads their vertices, then their slot distances and list extents, in one
// round trip per level (lane i: claimed candidate i); the first 64 labels of
// the next claimed candidate's list are loaded before the current one's probes.
// Instantiated once for an LDS tile and once for an arena tile.
template <class V, class LT, class KP>
__device__ __attribute__((always_inline)) inline void prune_candidates(
    const GenArgs &a, uint32_t b, uint32_t c_beg, uint32_t c_end, const uint32_t *cands, KP rec,
    uint32_t cap, uint32_t &sh_next, uint32_t &sh_acc_n, Item *sh_acc,
    unsigned long long &wave_reads) {
  const uint32_t sh = rec_shift<V>(a.K + 1);
  const uint32_t lane = (uint32_t)lane_id();
  for (;;) {
    uint32_t c = 0, k = 0;
    if (lane == 0) {
      const uint32_t seen =
          __hip_atomic_load(&sh_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const uint32_t rem = c_end - min(c_end, c_beg + seen);
      k = min(a.claim_max, max(1u, rem / kClaimSpread));
      c = c_beg + atomicAdd(&sh_next, k);
    }
    c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, kWave));
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, kWave));
    if (c >= c_end) break;
    const uint32_t cn = min(k, c_end - c);
    // lane i: claimed candidate i's vertex, distance, list start and view
    // (computed lane-parallel once; a candidate's values are read out into
    // scalars once, when it is next)
    uint32_t y = 0, dy = 0;
    ListView vw{0, 0, 0, 0};
    unsigned long long st = 0;
    if (lane < cn) {
      y = cands[c + lane];
      dy = slot_dist(a, a.cand_d[slot_at(a, b, y)]);
      st = a.seg_start[y];
      const uint32_t len = a.seg_len[y];
      uint32_t sorted = 0, pre = 1;
      if (a.hcnt && a.h > 1) {
        sorted = a.sorted_len[y];
        pre = hop_prefix(a, y, a.h - 1);
      }
      vw = cand_view(a, len, sorted, pre);
    }
    const LT *pool = stream_pool<LT>(a);
    auto list_of = [&](uint32_t i) {
      const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)st, i);
      const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(st >> 32), i);
      return pool + (((unsigned long long)hi << 32) | lo);
    };
    auto rl = [](uint32_t v, uint32_t i) { return (uint32_t)__builtin_amdgcn_readlane(v, i); };
    auto view_of = [&](uint32_t i) {
      return ListView{rl(vw.a0, i), rl(vw.na, i), rl(vw.b0, i), rl(vw.len, i)};
    };
    const LT *Lc = list_of(0);
    ListView vc = view_of(0);
    LT f0 = first_step(Lc, vc);
    for (uint32_t i = 0; i < cn; ++i) {
      // next candidate's first step (the last one reloads its own, so the
      // loads are unconditional)
      const uint32_t nx = min(i + 1, cn - 1);
      const LT *Ln = list_of(nx);
      const ListView vn = view_of(nx);
      const LT g0 = first_step(Ln, vn);
      const uint32_t yi = __builtin_amdgcn_readlane(y, i), di = __builtin_amdgcn_readlane(dy, i);
      uint32_t reads = 0;
      const bool pruned = witness_search<V>(a, Lc, vc, f0, di, rec, cap, sh, reads);
      wave_reads += reads;
      if (lane == 0) {
        if (!pruned) {
          const uint32_t p = atomicAdd(&sh_acc_n, 1u);
          sh_acc[p] = Item{b, yi, di};
        }
      }
      Lc = Ln;
      vc = vn;
      f0 = g0;
    }
  }
}

// LT: label type of the stream (uint32_t: the pool's 32-bit copy).
template <class V, class LT>
__global__ void __launch_bounds__(kPruneThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_prune(GenArgs a) {
  with_batch(a);
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t sh_w, sh_acc_n, sh_acc_base, sh_next, sh_touch_n, sh_touch_base;
  __shared__ uint32_t sh_item, sh_item_end, sh_rows;
  __shared__ Item sh_acc[kAccSlots];
  __shared__ uint32_t sh_rank[kAccSlots];
  __shared__ uint32_t sh_touch[kAccSlots];
  const int lane = lane_id();
  const uint32_t K1 = a.K + 1;
  const uint32_t lds_bytes = a.lds_slots << (rec_shift<V>(K1) + 2);
  Item *Fn = frontier(a, a.cur ^ 1);
  uint32_t tile_b = kInf32;  // hub slot whose tile is in LDS
  unsigned long long wave_reads = 0, tile_reads = 0;
  const uint64_t gwords = a.gtile_cap / gridDim.x;
  uint32_t *gkeys = a.gtile + (uint64_t)blockIdx.x * gwords;
  const uint32_t total = a.ctrl->total_chunks;
  // Work items are claimed `claim` at a time, enough for about 64 candidates
  // per claim (1..8 items): with one claim per item, rounds of small items
  // (late hubs, low-degree graphs) were bound by the shared counter, while
  // items of 64+ candidates stay one per claim to keep the load balanced.  A
  // workgroup's first claim is its own id, later ones come from the counter.
  const uint64_t cands = a.hub_off[a.B];
  const uint32_t claim =
      cands ? (uint32_t)std::min<uint64_t>(8, std::max<uint64_t>(1, (64ull * total) / cands)) : 1u;
  if (threadIdx.x == 0) {
    sh_item = blockIdx.x * claim;
    sh_item_end = sh_item + claim;
    sh_acc_n = 0;
    sh_touch_n = 0;
  }
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) {
      if (sh_item == sh_item_end) {
        sh_item = gridDim.x * claim + atomicAdd(&a.ctrl->work_ctr, claim);
        sh_item_end = sh_item + claim;
      }
      sh_w = sh_item++;
      sh_next = 0;
    }
    __syncthreads();
    const uint32_t w = sh_w;
    if (w >= total) break;
    // hub slot b with chunk_off[b] <= w < chunk_off[b+1], by a 64-way search:
    // each round the wave samples 64 evenly spaced offsets of [lo, hi) and
    // keeps the last one <= w (three dependent loads for 2^16 hubs instead of
    // sixteen; items of a few candidates each spent most of their time here)
    uint32_t lo = 0, hi = a.B;  // invariant chunk_off[lo] <= w < chunk_off[hi]
    while (hi - lo > 1) {
      const uint32_t step = (hi - lo + kWave - 1) / kWave;
      const uint32_t pos = lo + (uint32_t)lane * step;
      const bool le = pos < hi && a.chunk_off[pos] <= w;
      const unsigned long long m = __ballot(le);  // lane 0 (pos = lo) always holds
      lo += (uint32_t)(63 - __builtin_clzll(m)) * step;
      hi = min(hi, lo + step);
    }
    const uint32_t b = lo;
    const uint32_t s = hub_of(a, b);
    const uint32_t c_beg = (w - a.chunk_off[b]) * a.tile;
    const uint32_t c_end = min(a.hub_ccnt[b], c_beg + a.tile);
    // the staged labels of earlier items are published once this item's
    // could overflow the staging slots (c_end - c_beg <= tile <= kAccSlots)
    if (sh_acc_n + (c_end - c_beg) > kAccSlots)
      publish_accepted(a, Fn, sh_acc, sh_rank, sh_touch, sh_acc_n, sh_acc_base, sh_touch_n, sh_touch_base);
    const uint32_t *cands = a.cgroup + a.hub_off[b];
    const ListView sv = tile_view(a, s);
    const uint32_t cap = tile_cap_for(sv.len);
    const uint32_t icap = sv.len > a.idx_tiles
                              ? idx_tile_cap(sv.len, K1 * (uint32_t)sizeof(V), lds_bytes)
                              : 0u;
    if (icap) {
      if (tile_b != b) {
        build_idx_tile<V>(a, s, sv, lds, icap, sh_rows);
        tile_b = b;
        tile_reads += sv.len;
      }
      const IdxTile it{(lds_ptr<const uint32_t>)lds,
                       (uint32_t)(uintptr_t)(lds_ptr<const uint32_t>)(lds + icap),
                       K1 * (uint32_t)sizeof(V)};
      prune_candidates<V, LT>(a, b, c_beg, c_end, cands, it, icap, sh_next, sh_acc_n, sh_acc,
                              wave_reads);
    } else if (cap <= a.lds_slots) {
      const uint32_t lcap = lds_tile_cap(sv.len, a.lds_slots);
      if (tile_b != b) {
        build_tile<V>(a, s, sv, lds, lcap);
        tile_b = b;
        tile_reads += sv.len;
      }
      prune_candidates<V, LT>(a, b, c_beg, c_end, cands, (lds_ptr<const uint32_t>)lds, lcap,
                              sh_next, sh_acc_n, sh_acc, wave_reads);
    } else {
      if (rec_tile_words<V>(cap, K1) > gwords) {
        // arena too small: the host rolls the batch back, grows the arena
        // and runs the batch again
        if (threadIdx.x == 0) atomicOr(&a.ctrl->overflow, kOvfArena);
        __syncthreads();
        continue;
      }
      build_tile<V>(a, s, sv, gkeys, cap);
      // the tile was written through L2 (atomics); drop stale L1 lines of
      // this workgroup's arena slice before plain loads read it
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      tile_b = kInf32;
      tile_reads += sv.len;
      prune_candidates<V, LT>(a, b, c_beg, c_end, cands, (glb_ptr<const uint32_t>)gkeys, cap,
                              sh_next, sh_acc_n, sh_acc, wave_reads);
    }
    __syncthreads();
  }
  publish_accepted(a, Fn, sh_acc, sh_rank, sh_touch, sh_acc_n, sh_acc_base, sh_touch_n, sh_touch_base);
  if (lane == 0 && wave_reads) atomicAdd(&a.ctrl->label_reads, wave_reads);
  if (threadIdx.x == 0 && tile_reads) atomicAdd(&a.ctrl->tile_reads, tile_reads);
}

// --------------------------------------------------------- prune, by rows
// Row items for rounds whose candidates are dense (k_chunk_scan sets
// row_mode): an item is hub slot b's candidates among a run of row_run
// y-blocks of row_block vertices, found by scanning b's dense slot row there
// (fresh slots of the running attempt are exactly the round's candidates);
// one tile build serves the run.  Each XCD has its own queue of y-blocks
// k = x, x + 8, ..., whose runs come with all B hubs in turn, so the B hubs'
// items of one run are processed at the same time on one XCD and re-read the
// same lists L(y) from its L2 instead of from HBM (hub-ordered items re-read
// a vertex's list once per hub, spread over the whole round).  Workgroups
// are dealt to XCDs round-robin by id; a workgroup whose queue runs dry takes
// items from the other queues in turn.  The decisions do not depend on the order: a round's
// prune reads only labels of earlier rounds and batches.
template <class V, class LT>
__global__ void __launch_bounds__(kPruneThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_prune_rows(GenArgs a) {
  with_batch(a);
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t sh_b, sh_k, sh_kn, sh_q, sh_next, sh_cnt, sh_rows;
  __shared__ uint32_t sh_acc_n, sh_acc_base, sh_touch_n, sh_touch_base;
  __shared__ Item sh_acc[kAccSlots];
  __shared__ uint32_t sh_rank[kAccSlots];
  __shared__ uint32_t sh_touch[kAccSlots];
  __shared__ uint32_t sh_cand[kPruneThreads];
  if (!a.ctrl->row_mode) return;
  const int lane = lane_id();
  const uint32_t K1 = a.K + 1;
  const uint32_t lds_bytes = a.lds_slots << (rec_shift<V>(K1) + 2);
  Item *Fn = frontier(a, a.cur ^ 1);
  uint32_t tile_b = kInf32;  // hub slot whose tile is in LDS
  unsigned long long wave_reads = 0, tile_reads = 0, row_slots = 0;
  const uint64_t gwords = a.gtile_cap / gridDim.x;
  uint32_t *gkeys = a.gtile + (uint64_t)blockIdx.x * gwords;
  const uint32_t nblk = (uint32_t)(((uint64_t)a.n + a.row_block - 1) / a.row_block);
  const uint32_t home = blockIdx.x % kXcds;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  if (threadIdx.x == 0) {
    sh_q = 0;
    sh_acc_n = 0;
    sh_touch_n = 0;
  }
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) {
      sh_b = kInf32;
      while (sh_q < kXcds) {
        // queue x: its blocks k = x + 8i, in runs of row_run blocks; item j
        // is run j / B of hub slot j % B
        const uint32_t x = (home + sh_q) % kXcds;
        const uint32_t nb = x < nblk ? (nblk - x + kXcds - 1) / kXcds : 0u;
        const uint32_t nruns = (nb + a.row_run - 1) / a.row_run;
        const uint64_t items = (uint64_t)nruns * a.B;
        const uint32_t j = items ? atomicAdd(&a.ctrl->row_ctr[x], 1u) : 0u;
        if (j < items) {
          const uint32_t i0 = (j / a.B) * a.row_run;
          sh_k = x + kXcds * i0;            // first block of the run
          sh_kn = min(a.row_run, nb - i0);  // blocks in the run
          sh_b = j % a.B;
          break;
        }
        ++sh_q;
      }
    }
    __syncthreads();
    const uint32_t b = sh_b;
    if (b == kInf32) break;
    const uint32_t k = sh_k;
    const uint32_t s = hub_of(a, b);
    const ListView sv = tile_view(a, s);
    const uint32_t cap = tile_cap_for(sv.len);
    const uint32_t icap = sv.len > a.idx_tiles
                              ? idx_tile_cap(sv.len, K1 * (uint32_t)sizeof(V), lds_bytes)
                              : 0u;
    const bool in_lds = icap || cap <= a.lds_slots;
    const uint32_t lcap = icap ? icap : in_lds ? lds_tile_cap(sv.len, a.lds_slots) : cap;
    if (!in_lds && rec_tile_words<V>(cap, K1) > gwords) {
      // arena too small: the host rolls the batch back, grows the arena and
      // runs the batch again
      if (threadIdx.x == 0) atomicOr(&a.ctrl->overflow, kOvfArena);
      __syncthreads();
      continue;
    }
    bool built = false;
    // the run's blocks k, k + 8, ..., 1024 vertices at a time
    const uint32_t kSubs = a.row_block / kPruneThreads;
    const uint32_t kn = sh_kn;
    for (uint32_t i = 0; i < kn * kSubs; ++i) {
      const uint32_t kb = k + kXcds * (i / kSubs);
      const uint32_t y0 = kb * a.row_block + (i % kSubs) * kPruneThreads;
      const uint32_t y_end = min(a.n, (kb + 1) * a.row_block);
      if (y0 >= y_end) continue;
      row_slots += min(kPruneThreads, y_end - y0);
      // this hub's candidates among [y0, y0 + 1024), compacted into sh_cand
      if (threadIdx.x == 0) sh_cnt = 0;
      __syncthreads();
      const uint32_t y = y0 + threadIdx.x;
      const uint32_t v = y < y_end ? a.cand_d[slot_at(a, b, y)] : kInf32;
      const bool cand = y < y_end && slot_current(a, v);
      const unsigned long long m = __ballot(cand);
      uint32_t base = 0;
      if (lane == 0 && m) base = atomicAdd(&sh_cnt, (uint32_t)__popcll(m));
      base = __shfl(base, 0, kWave);
      if (cand) sh_cand[base + (uint32_t)__popcll(m & lt_mask)] = y;
      __syncthreads();
      const uint32_t cnt = sh_cnt;
      if (cnt == 0) continue;
      if (!built) {
        if (in_lds) {
          if (tile_b != b) {
            if (icap)
              build_idx_tile<V>(a, s, sv, lds, icap, sh_rows);
            else
              build_tile<V>(a, s, sv, lds, lcap);
            tile_b = b;
            tile_reads += sv.len;
          }
        } else {
          build_tile<V>(a, s, sv, gkeys, cap);
          // written through L2 (atomics): drop stale L1 lines of the slice
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          tile_b = kInf32;
          tile_reads += sv.len;
        }
        built = true;
      }
      for (uint32_t c0 = 0; c0 < cnt; c0 += a.tile) {
        const uint32_t c1 = min(cnt, c0 + a.tile);
        // staged labels are published once this chunk's could overflow them
        if (sh_acc_n + (c1 - c0) > kAccSlots)
          publish_accepted(a, Fn, sh_acc, sh_rank, sh_touch, sh_acc_n, sh_acc_base, sh_touch_n,
                           sh_touch_base);
        if (threadIdx.x == 0) sh_next = 0;
        __syncthreads();
        if (icap)
          prune_candidates<V, LT>(a, b, c0, c1, sh_cand,
                                  IdxTile{(lds_ptr<const uint32_t>)lds,
                                          (uint32_t)(uintptr_t)(lds_ptr<const uint32_t>)(lds + icap),
                                          K1 * (uint32_t)sizeof(V)},
                                  icap, sh_next, sh_acc_n, sh_acc, wave_reads);
        else if (in_lds)
          prune_candidates<V, LT>(a, b, c0, c1, sh_cand, (lds_ptr<const uint32_t>)lds, lcap,
                                  sh_next, sh_acc_n, sh_acc, wave_reads);
        else
          prune_candidates<V, LT>(a, b, c0, c1, sh_cand, (glb_ptr<const uint32_t>)gkeys, cap,
                                  sh_next, sh_acc_n, sh_acc, wave_reads);
        __syncthreads();
      }
    }
  }
  publish_accepted(a, Fn, sh_acc, sh_rank, sh_touch, sh_acc_n, sh_acc_base, sh_touch_n, sh_touch_base);
  if (lane == 0 && wave_reads) atomicAdd(&a.ctrl->label_reads, wave_reads);
  if (threadIdx.x == 0 && tile_reads) atomicAdd(&a.ctrl->tile_reads, tile_reads);
  if (threadIdx.x == 0 && row_slots) atomicAdd(&a.ctrl->row_slots, row_slots);
}

// -------------------------------------------------------- reserve / append
__global__ void __launch_bounds__(256) k_reserve(GenArgs a) {
  with_batch(a);
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  const uint32_t cnt = a.ctrl->touched_cnt;
  // 64 touched vertices per wave step (lane i: the step's vertex i), with one
  // pool reservation for all of them that outgrow their segments: in a
  // batch's first rounds most vertices grow, and one reservation per vertex
  // on the single pool counter serialised them
  for (uint32_t t0 = wave * kWave; t0 < cnt; t0 += nwaves * kWave) {
    const uint32_t t = t0 + (uint32_t)lane;
    uint32_t y = 0, len = 0, ncap = 0;
    if (t < cnt) {
      y = a.touched[t];
      len = a.seg_len[y];
      const uint32_t cap = a.seg_cap[y];
      const uint32_t need = len + a.add_cnt[y];
      if (need > cap) ncap = (max(need, cap * 2u) + 7u) & ~7u;
    }
    const unsigned long long grow = __ballot(ncap != 0);
    if (!grow) continue;
    const unsigned long long incl = wave_incl_scan_u64(ncap);
    const unsigned long long total = __shfl(incl, kWave - 1, kWave);
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.ctrl->pool_top, total);
    base = __shfl(base, 0, kWave);
    if (base + total > a.pool_cap) {
      if (lane == 0) atomicOr(&a.ctrl->overflow, kOvfPool);
      continue;
    }
    const unsigned long long ns = base + incl - ncap;
    // each growing vertex's labels move to its new segment, the whole wave
    // copying one vertex at a time
    for (unsigned long long m = grow; m; m &= m - 1) {
      const int j = __ffsll((long long)m) - 1;
      const uint32_t yj = __shfl(y, j, kWave), lj = __shfl(len, j, kWave);
      const unsigned long long nsj = __shfl(ns, j, kWave);
      const unsigned long long os = a.seg_start[yj];
      if (a.pool)
        for (uint32_t i = lane; i < lj; i += kWave) a.pool[nsj + i] = a.pool[os + i];
      if (a.pool32)
        for (uint32_t i = lane; i < lj; i += kWave) a.pool32[nsj + i] = a.pool32[os + i];
    }
    if (ncap) {
      a.seg_start[y] = ns;
      a.seg_cap[y] = ncap;
    }
  }
}

__global__ void __launch_bounds__(256) k_append(GenArgs a) {
  with_batch(a);
  const Item *Fn = frontier(a, a.cur ^ 1);
  const uint32_t cnt = min(a.ctrl->fcnt[a.cur ^ 1], a.fcap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const Item it = Fn[i];
    const uint32_t p = a.seg_len[it.y] + a.frank[i];
    if (p < a.seg_cap[it.y]) {
      const unsigned long long at = a.seg_start[it.y] + p;
      const uint32_t hub = hub_of(a, it.b);
      if (a.pool) a.pool[at] = pack_label(hub, a.h, it.d);
      if (a.pool32) a.pool32[at] = pack32(a, hub, a.h, it.d);
    } else {
      atomicOr(&a.ctrl->overflow, kOvfPool);
    }
  }
}

__global__ void __launch_bounds__(256) k_finalize(GenArgs a) {
  with_batch(a);
  const uint32_t cnt = a.ctrl->touched_cnt;
  // a lost frontier item leaves a counted label unwritten: publish nothing
  // (the host rolls the batch back)
  const bool keep = !(a.ctrl->overflow & kOvfFront);
  uint32_t mx = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x) {
    const uint32_t y = a.touched[t];
    if (keep) {
      // capped: when the pool ran out some appends were dropped
      const uint32_t len = a.seg_len[y];
      const uint32_t nl = min(len + a.add_cnt[y], a.seg_cap[y]);
      a.seg_len[y] = nl;
      mx = max(mx, nl);
      // the next round's frontier group of y: the labels it just gained
      a.fgrp[t] = FGroup{y, nl - len, a.seg_start[y] + len};
    }
    a.add_cnt[y] = 0;
  }
  if (mx) atomicMax(&a.ctrl->maxlen, mx);
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl->gcnt = keep ? min(cnt, a.n) : 0u;
}

// Removes the labels of the batch's hubs (the batch being rolled back): they
// sit at the tail of every segment, after all labels of earlier batches,
// whose hubs are all smaller (hubs run in ascending order); the self label at
// position 0 stays.  The redo runs under a new slot tag, so the batch's dense
// slots need no refill.
__global__ void __launch_bounds__(256) k_rollback(GenArgs a) {
  with_batch(a);
  const uint32_t lo = hub_of(a, 0), hi = hub_of(a, a.B - 1);
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
    uint32_t len = a.seg_len[v];
    const unsigned long long st = a.seg_start[v];
    while (len > 1) {
      uint32_t hub, hop, dist;
      pool_label(a, st + len - 1, hub, hop, dist);
      if (hub < lo || hub > hi) break;
      --len;
    }
    a.seg_len[v] = len;
    a.add_cnt[v] = 0;
  }
}

// --------------------------------------------------------------- export
__global__ void __launch_bounds__(256) k_gather(GenArgs a, const unsigned long long *out_off,
                                                unsigned long long *out) {
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t v = wave; v < a.n; v += nwaves) {
    const unsigned long long src = a.seg_start[v], dst = out_off[v];
    const uint32_t len = a.seg_len[v];
    for (uint32_t i = lane; i < len; i += kWave) {
      uint32_t hub, hop, dist;
      pool_label(a, src + i, hub, hop, dist);
      out[dst + i] = pack_label(hub, hop, dist);
    }
  }
}

// The export's 32-bit path: the lists' 32-bit copies gathered into one
// array (k_gather32), sorted per vertex as 32-bit keys, then widened into
// the index's 8-byte labels (k_widen32); the packing keeps (hub, hop, dist)
// order.
__global__ void __launch_bounds__(256) k_gather32(GenArgs a, const unsigned long long *out_off,
                                                  uint32_t *out) {
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  for (uint32_t v = wave; v < a.n; v += nwaves) {
    const unsigned long long src = a.seg_start[v], dst = out_off[v];
    const uint32_t len = a.seg_len[v];
    for (uint32_t i = lane; i < len; i += kWave) out[dst + i] = a.pool32[src + i];
  }
}
__global__ void __launch_bounds__(256) k_widen32(GenArgs a, const uint32_t *in, uint64_t total,
                                                 unsigned long long *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t hub, hop, dist;
    unpack(a, in[i], hub, hop, dist);
    out[i] = pack_label(hub, hop, dist);
  }
}

// Hop zones: after a batch, every list whose append zone has grown past
// 1/alpha_div of its sorted zone is sorted again, stably by hop, and its
// per-hop prefix counts rewritten -- so a later round's query skips the
// labels of hops it cannot use.  A wave per list: it counts the hops of the
// list (hop 0 is the self label), scatters the labels stably by hop into its
// slice of `tmp` (and the 32-bit copies into tmp32), then copies them back
// over the segment.  Lists longer than a slice (`per_wave` labels) stay as
// they are: the query reads their append zone whole, which is always
// correct.  Runs between batches (never while a batch can still be rolled
// back, whose labels must stay at the tail of every list).
// One list's stable scatter by hop into (t64, t32) (pos[x]: the first slot of
// hop x), then back over the segment (whole wave).  T64 / T32: LDS or
// global pointers.
template <class T64, class T32>
__device__ inline void hop_scatter(const GenArgs &a, unsigned long long st, uint32_t len,
                                   uint32_t *pos, T64 t64, T32 t32, bool p32) {
  const uint32_t lane = (uint32_t)lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const uint32_t K1 = a.K + 1;
  const bool p64 = a.pool != nullptr;  // absent when the build keeps the 32-bit copy
  for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
    const uint32_t j = j0 + lane;
    unsigned long long l = 0;
    uint32_t l32 = 0, hop = kInf32;
    if (j < len) {
      if (p64) {
        l = a.pool[st + j];
        hop = label_hop(l);
      }
      if (p32) {
        l32 = a.pool32[st + j];
        uint32_t hub, d;
        unpack(a, l32, hub, hop, d);
      }
    }
    uint32_t dst = 0;
#pragma unroll
    for (uint32_t x = 0; x <= kMaxK; ++x) {
      if (x >= K1) continue;
      const unsigned long long m = __ballot(hop == x);
      if (hop == x) dst = pos[x] + (uint32_t)__popcll(m & lt_mask);
      pos[x] += (uint32_t)__popcll(m);
    }
    if (j < len) {
      if (p64) t64[dst] = l;
      if (p32) t32[dst] = l32;
    }
  }
  // the scratch was written by other lanes of this wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  for (uint32_t j = lane; j < len; j += kWave) {
    if (p64) a.pool[st + j] = t64[j];
    if (p32) a.pool32[st + j] = t32[j];
  }
  // the next list reuses the scratch
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Lists of up to lds_labels labels are scattered in the wave's LDS slice
// (one read and one write of each label in HBM), longer ones in its slice of
// `tmp` in HBM (two of each).
__global__ void __launch_bounds__(256) k_hop_sort(GenArgs a, unsigned long long *tmp,
                                                  uint32_t *tmp32, uint32_t per_wave,
                                                  uint32_t alpha_div, uint32_t lds_labels) {
  extern __shared__ unsigned long long zl64[];
  const uint32_t lane = (uint32_t)lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / kWave;
  const uint32_t wib = threadIdx.x / kWave;  // wave in block
  const bool p32 = tmp32 != nullptr;
  unsigned long long *t64 = tmp + (uint64_t)wave * per_wave;
  uint32_t *t32 = p32 ? tmp32 + (uint64_t)wave * per_wave : nullptr;
  // LDS: 4 slices of lds_labels u64, then 4 of lds_labels u32; a build
  // with only the 32-bit pool has the u32 slices alone
  const bool p64 = a.pool != nullptr;
  lds_ptr<unsigned long long> l64 = (lds_ptr<unsigned long long>)(zl64 + wib * lds_labels);
  lds_ptr<uint32_t> l32 =
      (lds_ptr<uint32_t>)(reinterpret_cast<uint32_t *>(zl64 + (p64 ? 4 * lds_labels : 0u)) +
                          wib * lds_labels);
  const uint32_t K1 = a.K + 1;
  for (uint64_t base = (uint64_t)wave * kWave; base < a.n; base += (uint64_t)nwaves * kWave) {
    const uint32_t v = (uint32_t)base + lane;
    bool need = false;
    if (v < a.n) {
      const uint32_t len = a.seg_len[v], sorted = a.sorted_len[v];
      need = len > sorted && (uint64_t)(len - sorted) * alpha_div > sorted &&
             (len <= per_wave || len <= lds_labels);
    }
    unsigned long long todo = __ballot(need);
    while (todo) {
      const uint32_t i = (uint32_t)(__ffsll((long long)todo) - 1);
      todo &= todo - 1;
      const uint32_t y = (uint32_t)base + i;
      const unsigned long long st = a.seg_start[y];
      const uint32_t len = a.seg_len[y];
      // 1. labels per hop (wave-uniform counters), from the 32-bit copy
      // when there is one
      uint32_t cnt[kMaxK + 1];
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x) cnt[x] = 0;
      for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
        const uint32_t j = j0 + lane;
        uint32_t hop = kInf32;
        if (j < len) {
          if (p32) {
            uint32_t hub, d;
            unpack(a, a.pool32[st + j], hub, hop, d);
          } else {
            hop = label_hop(a.pool[st + j]);
          }
        }
#pragma unroll
        for (uint32_t x = 0; x <= kMaxK; ++x)
          if (x < K1) cnt[x] += (uint32_t)__popcll(__ballot(hop == x));
      }
      // 2. stable scatter by hop, back over the segment
      uint32_t pos[kMaxK + 1];
      uint32_t run = 0;
#pragma unroll
      for (uint32_t x = 0; x <= kMaxK; ++x) {
        pos[x] = run;
        run += x < K1 ? cnt[x] : 0u;
      }
      if (len <= lds_labels)
        hop_scatter(a, st, len, pos, l64, l32, p32);
      else
        hop_scatter(a, st, len, pos, (glb_ptr<unsigned long long>)t64, (glb_ptr<uint32_t>)t32,
                    p32);
      if (lane == 0) {
        uint32_t c = cnt[0];
        for (uint32_t x = 1; x < K1 - 1; ++x) {
          c += cnt[x];
          a.hcnt[(uint64_t)y * a.hcnt_stride + (x - 1)] = c;
        }
        a.sorted_len[y] = len;
      }
    }
  }
}

// A group build keeps only labels of its own hubs: the self label (v,0,0) at
// position 0 of a vertex outside the group is dropped from its segment
// before the export (in_group: one byte per vertex).
__global__ void __launch_bounds__(256) k_drop_foreign_self(GenArgs a, const uint8_t *in_group) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
    if (in_group[v]) continue;
    a.seg_start[v] += 1;
    a.seg_len[v] -= 1;
  }
}

// ----------------------------------------------------------------- host
static void free_device_graph(hsdl_graph *g) {
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (g->d_device >= 0 && cur != g->d_device) (void)hipSetDevice(g->d_device);
  if (g->d_off) (void)hipFree(g->d_off);
  if (g->d_adj) (void)hipFree(g->d_adj);
  if (g->d_slots) (void)hipFree(g->d_slots);
  if (g->d_pool) (void)hipFree(g->d_pool);
  if (g->d_pool32) (void)hipFree(g->d_pool32);
  for (int i = 0; i < hsdl_graph::kScratch; ++i) {
    if (g->scratch_p[i]) (void)hipFree(g->scratch_p[i]);
    g->scratch_p[i] = nullptr;
    g->scratch_bytes[i] = 0;
  }
  if (cur >= 0 && cur != g->d_device) (void)hipSetDevice(cur);
  g->d_off = g->d_adj = g->d_slots = g->d_pool = g->d_pool32 = nullptr;
  g->d_slots_bytes = g->d_pool_bytes = g->d_pool32_bytes = 0;
  g->slot_key = 0;
  g->d_device = -1;
}

static int upload_graph(hsdl_graph *g, int dev) {
  if (g->d_off && g->d_device == dev) return HSDL_OK;
  g->drop_device_copy();
  const uint32_t n = g->n;
  const uint64_t nnz = g->off[n];
  std::vector<uint2> adj(nnz);
  for (uint64_t e = 0; e < nnz; ++e) adj[e] = make_uint2(g->col[e], (uint32_t)g->w[e]);
  void *doff = nullptr, *dadj = nullptr;
  hipError_t e = hipMalloc(&doff, (n + 1ull) * 8);
  if (e == hipSuccess) e = hipMalloc(&dadj, std::max<uint64_t>(nnz, 1) * 8);
  if (e == hipSuccess) e = hipMemcpy(doff, g->off.data(), (n + 1ull) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(dadj, adj.data(), nnz * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (doff) (void)hipFree(doff);
    if (dadj) (void)hipFree(dadj);
    return hip_fail(e, "graph upload", __FILE__, __LINE__);
  }
  g->d_off = doff;
  g->d_adj = dadj;
  g->d_device = dev;
  g->d_free = free_device_graph;
  return HSDL_OK;
}

// Slots of hsdl_graph::scratch for Gen's buffers.
enum ScratchId : int {
  kScSegStart, kScSegLen, kScSegCap, kScAddCnt, kScTouched, kScFgrp, kScHubSlot,
  kScSortedLen, kScHcnt, kScSortTmp, kScSortTmp32, kScCflat, kScCgroup, kScHubCcnt, kScHubOff,
  kScHubCur, kScChunkOff, kScF0, kScF1, kScGtile, kScCtrl, kScHeavy, kScExpKeys, kScExpSorted,
  kScExpTmp, kScFrank, kScCount
};
static_assert(kScCount <= hsdl_graph::kScratch, "scratch slots");

struct Gen {
  GenArgs a{};
  hsdl_graph *mg = nullptr;  // the graph whose scratch the buffers come from and go back to
  DevBuf pool, pool32, seg_start, seg_len, seg_cap, add_cnt, touched, fgrp, hub_slot;
  DevBuf sorted_len, hcnt, sort_tmp, sort_tmp32;  // hop zones
  uint32_t sort_per_wave = 0;                     // labels per wave slice of sort_tmp
  int sort_grid = 0;
  DevBuf cflat, cgroup, hub_ccnt, hub_off, hub_cur, chunk_off, F0, F1, gtile, ctrl, heavy;
  DevBuf exp_keys, exp_sorted, exp_tmp;  // export: 32-bit keys, sorted keys, sort space
  DevBuf frank;                          // ranks of the next frontier's items
  uint64_t budget = 0;  // device bytes the build may hold
  uint64_t fixed = 0;   // per-vertex state and the dense slots
  int grid = 0;
  int prune_grid = 0;  // resident k_prune workgroups (two per CU)
  size_t lds_bytes = 0;
  bool small_vals = false;  // 16-bit tile values
  uint32_t bmax = 1;        // largest batch (slot rows)
  // captured batch sequences (without / with the row-item prune) and the
  // buffers each was captured with
  bool use_graph = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  GenArgs gkey[2]{};
  uint64_t graph_captures = 0;
  hipStream_t st = nullptr;
  bool verbose = false;
  // kernel timing (events around every expand and prune launch of a batch)
  bool timing = false;
  std::vector<hipEvent_t> ev;
  double t_expand = 0, t_prune = 0;
  uint64_t n_prune = 0;
  double h_grow = 0, h_enqueue = 0, h_wait = 0;  // host-side phase timers (verbose)
  uint64_t held() const {
    return fixed + pool.bytes + pool32.bytes + cflat.bytes + cgroup.bytes + F0.bytes + F1.bytes +
           frank.bytes +
           gtile.bytes + sort_tmp.bytes + sort_tmp32.bytes;
  }
  DevBuf *by_id(int id) {
    DevBuf *t[kScCount] = {&seg_start, &seg_len, &seg_cap, &add_cnt, &touched, &fgrp,
                           &hub_slot, &sorted_len, &hcnt, &sort_tmp, &sort_tmp32, &cflat, &cgroup,
                           &hub_ccnt, &hub_off, &hub_cur, &chunk_off, &F0, &F1, &gtile, &ctrl,
                           &heavy, &exp_keys, &exp_sorted, &exp_tmp, &frank};
    return t[id];
  }
  // Buffer `id` of at least n bytes: the graph's kept one when it is large
  // enough, else a new one (the kept one is freed).
  int take(int id, size_t n) {
    DevBuf &b = *by_id(id);
    if (mg && mg->scratch_p[id] && mg->scratch_bytes[id] >= n) {
      b.release();
      b.p = mg->scratch_p[id];
      b.bytes = mg->scratch_bytes[id];
      mg->scratch_p[id] = nullptr;
      mg->scratch_bytes[id] = 0;
      return HSDL_OK;
    }
    if (mg && mg->scratch_p[id]) {
      (void)hipFree(mg->scratch_p[id]);
      mg->scratch_p[id] = nullptr;
      mg->scratch_bytes[id] = 0;
    }
    return b.alloc(n);
  }
  ~Gen() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (hipGraphExec_t x : gexec)
      if (x) (void)hipGraphExecDestroy(x);
    // the buffers go back to the graph for its next build
    for (int id = 0; mg && id < kScCount; ++id) {
      DevBuf &b = *by_id(id);
      if (!b.p || mg->scratch_p[id]) continue;
      mg->scratch_p[id] = b.p;
      mg->scratch_bytes[id] = b.bytes;
      b.p = nullptr;
      b.bytes = 0;
    }
  }
};

static int grid_for_device(int dev) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus * 4;
}

static int read_ctrl(Gen &g, Ctrl &c) {
  HSDL_HIP(hipMemcpyAsync(&c, g.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, g.st));
  HSDL_HIP(hipStreamSynchronize(g.st));
  return HSDL_OK;
}

// Grows the label pool (at least doubling); segment offsets stay valid
// because the used prefix is copied as is.
static int grow_pool(Gen &g, uint64_t used, uint64_t min_cap) {
  uint64_t ncap = g.a.pool_cap * 2;
  while (ncap < min_cap) ncap *= 2;
  const bool p32 = g.a.pool32 != nullptr;  // then there is no 8-byte pool
  const uint64_t per_label = p32 ? 4 : 8;
  if (g.held() - g.pool.bytes - g.pool32.bytes + ncap * per_label > g.budget)
    return fail(HSDL_ENOMEM, "label pool exceeds the HBM budget");
  DevBuf np, np32;
  int rc = p32 ? np32.alloc(ncap * 4) : np.alloc(ncap * 8);
  if (rc) return rc;
  if (p32)
    HSDL_HIP(hipMemcpyAsync(np32.p, g.pool32.p, used * 4, hipMemcpyDeviceToDevice, g.st));
  else
    HSDL_HIP(hipMemcpyAsync(np.p, g.pool.p, used * 8, hipMemcpyDeviceToDevice, g.st));
  HSDL_HIP(hipStreamSynchronize(g.st));
  if (!p32) {
    g.pool = std::move(np);
    g.a.pool = g.pool.as<unsigned long long>();
  }
  if (p32) {
    g.pool32 = std::move(np32);
    g.a.pool32 = g.pool32.as<uint32_t>();
  }
  g.a.pool_cap = ncap;
  return HSDL_OK;
}

static int grow_gtile(Gen &g) {
  const uint64_t nw = g.a.gtile_cap * 2;
  if (g.held() - g.gtile.bytes + nw * 4 > g.budget)
    return fail(HSDL_ENOMEM, "tile arena exceeds the HBM budget");
  g.gtile.release();
  int rc = g.take(kScGtile, nw * 4);
  if (rc) return rc;
  g.a.gtile = g.gtile.as<uint32_t>();
  g.a.gtile_cap = nw;
  return HSDL_OK;
}

static int alloc_lists(Gen &g, uint32_t ccap, uint32_t fcap) {
  int rc;
  if (ccap != g.a.ccap) {
    g.cflat.release();
    g.cgroup.release();
    if (g.held() + (uint64_t)ccap * 12 > g.budget)
      return fail(HSDL_ENOMEM, "candidate lists exceed the HBM budget");
    if ((rc = g.take(kScCflat, (uint64_t)ccap * 8)) || (rc = g.take(kScCgroup, (uint64_t)ccap * 4)))
      return rc;
    g.a.cflat = g.cflat.as<unsigned long long>();
    g.a.cgroup = g.cgroup.as<uint32_t>();
    g.a.ccap = ccap;
  }
  if (fcap != g.a.fcap) {
    g.F0.release();
    g.F1.release();
    g.frank.release();
    if (g.held() + (uint64_t)fcap * (2 * sizeof(Item) + 4) > g.budget)
      return fail(HSDL_ENOMEM, "frontier lists exceed the HBM budget");
    if ((rc = g.take(kScF0, (uint64_t)fcap * sizeof(Item))) ||
        (rc = g.take(kScF1, (uint64_t)fcap * sizeof(Item))) ||
        (rc = g.take(kScFrank, (uint64_t)fcap * 4)))
      return rc;
    g.a.F0 = g.F0.as<Item>();
    g.a.F1 = g.F1.as<Item>();
    g.a.frank = g.frank.as<uint32_t>();
    g.a.fcap = fcap;
  }
  return HSDL_OK;
}

// Enqueues one batch (K rounds) of hubs [s0, s0+B).
// The launches of one batch attempt (its K rounds) on g.st.  The kernels
// read the batch's s0 / B / tag from the control block (k_set_batch), so the
// sequence depends only on the buffers, K and whether row items may run.
static void launch_rounds(Gen &g, GenArgs a, uint32_t K, bool rows) {
  a.cur = 0;
  // grid-stride loops over the batch's hubs: sized for the largest batch
  const int small_grid = (int)std::min<uint32_t>((g.bmax + 255) / 256, 1024);
  hipLaunchKernelGGL(k_batch_begin, dim3(small_grid), dim3(256), 0, g.st, a);
  for (uint32_t h = 1; h <= K; ++h) {
    a.h = h;
    hipLaunchKernelGGL(k_round_prep, dim3(small_grid), dim3(256), 0, g.st, a);
    if (g.timing) (void)hipEventRecord(g.ev[4 * (h - 1) + 0], g.st);
    hipLaunchKernelGGL(k_expand, dim3(g.grid), dim3(256), 0, g.st, a);
    hipLaunchKernelGGL(k_expand_heavy, dim3(g.grid), dim3(256), 0, g.st, a);
    if (g.timing) (void)hipEventRecord(g.ev[4 * (h - 1) + 1], g.st);
    hipLaunchKernelGGL(k_count, dim3(g.grid), dim3(kGroupThreads), 0, g.st, a);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, g.st, a);
    hipLaunchKernelGGL(k_group, dim3(g.grid), dim3(kGroupThreads), 0, g.st, a);
    if (g.timing) (void)hipEventRecord(g.ev[4 * (h - 1) + 2], g.st);
    {
      // 16- or 32-bit tile values; 32- or 64-bit labels streamed
      auto kp = g.small_vals ? (a.pool32 ? k_prune<uint16_t, uint32_t>
                                         : k_prune<uint16_t, unsigned long long>)
                             : (a.pool32 ? k_prune<uint32_t, uint32_t>
                                         : k_prune<uint32_t, unsigned long long>);
      hipLaunchKernelGGL(kp, dim3(g.prune_grid), dim3(kPruneThreads), g.lds_bytes, g.st, a);
      // row items instead, when k_chunk_scan chose them (one of the two
      // launches returns at once); never for a one-hub batch unless forced,
      // so the plain form pays no second launch per round
      if (rows) {
        auto kr = g.small_vals ? (a.pool32 ? k_prune_rows<uint16_t, uint32_t>
                                           : k_prune_rows<uint16_t, unsigned long long>)
                               : (a.pool32 ? k_prune_rows<uint32_t, uint32_t>
                                           : k_prune_rows<uint32_t, unsigned long long>);
        hipLaunchKernelGGL(kr, dim3(g.prune_grid), dim3(kPruneThreads), g.lds_bytes, g.st, a);
      }
    }
    if (g.timing) (void)hipEventRecord(g.ev[4 * (h - 1) + 3], g.st);
    hipLaunchKernelGGL(k_reserve, dim3(g.grid), dim3(256), 0, g.st, a);
    hipLaunchKernelGGL(k_append, dim3(g.grid), dim3(256), 0, g.st, a);
    hipLaunchKernelGGL(k_finalize, dim3(g.grid), dim3(256), 0, g.st, a);
    a.cur ^= 1;
  }
}

// The buffers a captured batch has baked in (they change when one grows).
static bool same_buffers(const GenArgs &x, const GenArgs &y) {
  return x.pool == y.pool && x.pool_cap == y.pool_cap && x.pool32 == y.pool32 &&
         x.gtile == y.gtile && x.gtile_cap == y.gtile_cap && x.cflat == y.cflat &&
         x.cgroup == y.cgroup && x.ccap == y.ccap && x.F0 == y.F0 && x.F1 == y.F1 &&
         x.frank == y.frank && x.fcap == y.fcap;
}

// Captures the batch sequence (with or without the row-item prune) into a
// HIP graph, once per buffer state; false when capture is not possible here
// (the batch then runs as separate launches).
static bool batch_graph(Gen &g, uint32_t K, bool rows) {
  const int v = rows ? 1 : 0;
  if (g.gexec[v] && same_buffers(g.gkey[v], g.a)) return true;
  if (g.gexec[v]) {
    (void)hipGraphExecDestroy(g.gexec[v]);
    g.gexec[v] = nullptr;
  }
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(g.st, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  launch_rounds(g, g.a, K, rows);
  const hipError_t e = hipStreamEndCapture(g.st, &graph);
  hipGraphExec_t ex = nullptr;
  const bool ok = e == hipSuccess && graph &&
                  hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    (void)hipGetLastError();
    if (g.verbose) fprintf(stderr, "[hsdl] batch graph capture failed; separate launches\n");
    return false;
  }
  g.gexec[v] = ex;
  g.gkey[v] = g.a;
  ++g.graph_captures;
  return true;
}

static int enqueue_batch(Gen &g, uint32_t s0, uint32_t B, uint32_t K, uint32_t nhubs) {
  // k_batch_begin writes the B hubs into the first frontier unchecked
  if (B == 0 || B > g.a.fcap || B > g.bmax || s0 + (uint64_t)B > nhubs)
    return fail(HSDL_EINVAL, "hub batch does not fit the frontier list or the hub sequence");
  hipLaunchKernelGGL(k_set_batch, dim3(1), dim3(1), 0, g.st, g.a.ctrl, s0, B, g.a.tag);
  const bool rows = B > 1 || g.a.row_policy == 2;
  // one graph launch per batch instead of 1 + 11 K launches (the plain form
  // is bound by the host's launch rate); not with per-launch timing events
  if (g.use_graph && !g.timing && batch_graph(g, K, rows)) {
    HSDL_HIP(hipGraphLaunch(g.gexec[rows ? 1 : 0], g.st));
  } else {
    launch_rounds(g, g.a, K, rows);
  }
  HSDL_HIP(hipGetLastError());
  return HSDL_OK;
}

// Re-sorts the lists whose append zones have grown (k_hop_sort), after a
// batch that completed.  The wave slices of the scratch hold the longest
// list so far (grown here); a list that outgrows them before the next growth
// is left unsorted, which only costs reads.
static int sort_zones(Gen &g, uint32_t maxlen, uint32_t resort_div) {
  if (!g.sort_grid) g.sort_grid = std::max(1, g.grid / 4);  // 256 threads: 4 waves each
  const uint64_t waves = (uint64_t)g.sort_grid * 4;
  // tuning knob HSDL_SORT_SLICE caps the slice (longer lists stay unsorted)
  const uint32_t slice_cap = env_u32("HSDL_SORT_SLICE", 1u << 30, 1, 1u << 30);
  if (maxlen > g.sort_per_wave && g.sort_per_wave < slice_cap) {
    uint32_t per = std::max<uint32_t>(1024, g.sort_per_wave);
    while (per < maxlen) per *= 2;
    per = std::min(per, slice_cap);
    const bool p32 = g.a.pool32 != nullptr, p64 = g.a.pool != nullptr;
    const uint64_t bytes = waves * per * ((p64 ? 8ull : 0ull) + (p32 ? 4ull : 0ull));
    if (g.held() - g.sort_tmp.bytes - g.sort_tmp32.bytes + bytes > g.budget) {
      // no room for bigger slices: the longer lists stay unsorted
      if (!g.sort_per_wave) return HSDL_OK;
    } else {
      HSDL_HIP(hipStreamSynchronize(g.st));
      g.sort_tmp.release();
      g.sort_tmp32.release();
      int rc = p64 ? g.take(kScSortTmp, waves * per * 8) : HSDL_OK;
      if (!rc && p32) rc = g.take(kScSortTmp32, waves * per * 4);
      if (rc) return rc;
      g.sort_per_wave = per;
    }
  }
  // tuning knob HSDL_SORT_LDS: labels per wave's LDS slice (12 B each with
  // the 8-byte pool, 4 B with the 32-bit pool alone; four waves per
  // workgroup, at most 64 KB; R-MAT scale 20 builds with both pools: 0 4.252
  // s, 1024 4.249 s, 1360 4.231 s)
  const bool only32 = g.a.pool == nullptr;
  const uint32_t lds_labels = only32 ? env_u32("HSDL_SORT_LDS", 4080, 0, 4095)
                                     : env_u32("HSDL_SORT_LDS", 1360, 0, 1365);
  const size_t lds_bytes = (size_t)lds_labels * 4 * (only32 ? 4 : 12);
  hipLaunchKernelGGL(k_hop_sort, dim3(g.sort_grid), dim3(256), lds_bytes, g.st,
                     g.a, g.sort_tmp.as<unsigned long long>(),
                     g.a.pool32 ? g.sort_tmp32.as<uint32_t>() : nullptr, g.sort_per_wave,
                     resort_div, lds_labels);
  HSDL_HIP(hipGetLastError());
  return HSDL_OK;
}

// Chooses the next hub batch from the last one's measure (the autotuner of
// the batched form; the plain form pins B = 1): its work -- candidates plus
// an eighth of the scanned edges, both fixed by the graph and the schedule,
// so the schedule repeats from run to run -- or, with HSDL_TUNER=time, its
// wall-clock duration.
struct BatchTuner {
  uint32_t B, Bmax;
  double target;
  uint32_t next(double last) {
    double f = last > 0 ? target / last : 2.0;
    f = std::min(2.0, std::max(0.5, f));
    uint64_t nb = (uint64_t)std::llround(B * f);
    nb = std::max<uint64_t>(1, std::min<uint64_t>(nb, Bmax));
    B = (uint32_t)nb;
    return B;
  }
};

int generate(const hsdl_graph *hg, uint8_t K, const hsdl_opts *opts, const uint32_t *group,
             uint64_t ngroup, hsdl_index **out) {
  const uint32_t n = hg->n;
  if (K < 1 || K > kMaxK) return fail(HSDL_EINVAL, "K must be in 1..15");
  if (n == 0) return fail(HSDL_EINVAL, "empty graph");
  if (group) {
    // a group build: hub ids ascending, no repeats, inside the graph
    if (ngroup == 0 || ngroup > n) return fail(HSDL_EINVAL, "a hub group holds 1..n vertices");
    for (uint64_t i = 0; i < ngroup; ++i)
      if (group[i] >= n || (i && group[i] <= group[i - 1]))
        return fail(HSDL_EINVAL, "group hub ids must be ascending vertex ids without repeats");
  }
  const uint32_t nhubs = group ? (uint32_t)ngroup : n;
  if ((uint64_t)K * (uint64_t)std::max(hg->max_w, 1) > kMaxDist)
    return fail(HSDL_ERANGE, "K * max edge weight exceeds the 27-bit label distance");
  int dev = 0;
  int rc = use_device(opts->device, &dev);
  if (rc) return rc;
  const uint64_t t_start = now_ns();
  Gen g;
  g.verbose = getenv("HSDL_VERBOSE") != nullptr;
  // HSDL_PROGRESS=<seconds>: one line on stderr per interval (long builds)
  const double progress_s = (double)env_u32("HSDL_PROGRESS", 0, 0, 86400);
  uint64_t next_progress = 0;
  HSDL_HIP(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
  } sg{g.st};
  g.grid = grid_for_device(dev);
  g.prune_grid = g.grid / 2;

  // device graph (kept with the graph handle)
  if ((rc = upload_graph(const_cast<hsdl_graph *>(hg), dev))) return rc;
  g.timing = opts->kernel_timing != 0;
  if (g.timing) {
    g.ev.resize(4 * K);
    for (auto &e : g.ev) HSDL_HIP(hipEventCreate(&e));
  }
  const uint64_t t_csr = now_ns();  // build time starts with the CSR resident

  size_t free_b = 0, total_b = 0;
  HSDL_HIP(hipMemGetInfo(&free_b, &total_b));
  // dense slots and label pools kept from an earlier build on this graph
  // count as free: this build reuses or replaces them
  hsdl_graph *mg = const_cast<hsdl_graph *>(hg);
  free_b += mg->d_slots_bytes + mg->d_pool_bytes + mg->d_pool32_bytes;
  for (int id = 0; id < kScCount; ++id) free_b += mg->scratch_bytes[id];
  g.mg = mg;
  g.budget = opts->hbm_budget ? opts->hbm_budget : (uint64_t)(free_b * 0.8);
  const uint32_t c0 = 16;  // initial label slots per vertex
  uint64_t pool0 = opts->pool_init_labels
                             ? std::max<uint64_t>(opts->pool_init_labels, (uint64_t)n * c0)
                             : std::max<uint64_t>((uint64_t)n * c0 * 4, 64ull << 20);
  // the last build's arena size on this graph, unless a size is forced
  const uint64_t gtile0 =
      opts->arena_init_words ? std::max<uint64_t>(opts->arena_init_words, g.grid)
                             : std::max<uint64_t>((uint64_t)g.grid * 16384, hg->hint_arena);
  // first candidate / frontier list size; HSDL_LIST_INIT starts them small
  // (the lists grow by rolling the batch back, like the pool)
  uint32_t list0 = env_u32(
      "HSDL_LIST_INIT",
      (uint32_t)std::min<uint64_t>(std::max<uint64_t>((uint64_t)n * 8, 8u << 20), 1u << 30), 64,
      1u << 30);
  // per vertex: segment start, length, capacity, add counter, touched
  // entry, frontier group (16 B), hub slot, hop zone counts
  const uint64_t vert_bytes = (uint64_t)n * (8 + 4 * 4 + 16 + 4 + 4 * (uint64_t)K);
  // dense dedupe slots: at most 1/slot_div of the budget (tuning knob
  // HSDL_SLOT_DIV), <= 65536 hubs
  const bool plain = opts->mode == HSDL_MODE_PLAIN;
  const uint32_t slot_div = env_u32("HSDL_SLOT_DIV", 4, 2, 64);
  uint32_t Bmax =
      (uint32_t)std::min<uint64_t>({g.budget / slot_div / (4ull * n), (uint64_t)n, 1ull << 16});
  if (plain) Bmax = 1;
  if (opts->hub_batch) Bmax = std::min<uint32_t>(opts->hub_batch, Bmax);
  if (Bmax == 0) return fail(HSDL_ENOMEM, "HBM budget below one hub's dedupe slots");
  // slot windows of 2^slot_lw hub slots (tuning knob HSDL_SLOT_WINDOW, log2,
  // default 4: 16 slots, one 64-byte sector), whose slots the batch cap then
  // fills whole: a cap of the budget's is rounded down to whole windows, a
  // caller's hub_batch is kept and the windows narrowed until they divide it
  uint32_t slot_lw = env_u32("HSDL_SLOT_WINDOW", 4, 0, 8);
  if (opts->hub_batch) {
    while (slot_lw && (Bmax & ((1u << slot_lw) - 1u))) --slot_lw;
  } else {
    while (slot_lw && (1u << slot_lw) > Bmax) --slot_lw;
    Bmax &= ~((1u << slot_lw) - 1u);
  }
  g.bmax = Bmax;
  // tuning knob HSDL_GRAPH=0: launch every kernel of a batch separately
  g.use_graph = env_u32("HSDL_GRAPH", 1, 0, 1) != 0;
  list0 = std::max(list0, Bmax);  // a batch's first frontier (its hubs) must fit
  // the lists the last build on this graph grew to (HSDL_LIST_INIT forces a
  // start size instead); each list pair's bytes are checked with the budget
  uint32_t ccap0 = list0, fcap0 = list0;
  if (!getenv("HSDL_LIST_INIT")) {
    ccap0 = std::max(ccap0, hg->hint_ccap);
    fcap0 = std::max(fcap0, hg->hint_fcap);
  }
  // heavy pieces per round (24 MB); when a round has more, the waves keep
  // the items that did not fit (HSDL_HEAVY_CAP lowers the cap)
  const uint32_t heavy_cap = env_u32("HSDL_HEAVY_CAP", 1u << 20, 1, 1u << 20);
  g.fixed = vert_bytes + (uint64_t)Bmax * n * 4 + (Bmax + 1ull) * 16 + (2ull << 20) +
            (uint64_t)heavy_cap * sizeof(HeavyItem);
  // 32-bit copy of the pool for the prune's stream, when hub, hop and
  // distance fields fit in one word (tuning knob HSDL_STREAM32=0 keeps the
  // prune on the 8-byte pool)
  uint32_t p32_hop_bits = 1, p32_dist_bits = 1, p32_hub_bits = 1;
  while ((1ull << p32_hop_bits) <= K) ++p32_hop_bits;
  while ((1ull << p32_dist_bits) <= (uint64_t)K * (uint64_t)std::max(hg->max_w, 1))
    ++p32_dist_bits;
  while ((1ull << p32_hub_bits) <= (uint64_t)(n - 1)) ++p32_hub_bits;
  const bool use32 = env_u32("HSDL_STREAM32", 1, 0, 1) &&
                     p32_hub_bits + p32_hop_bits + p32_dist_bits <= 32;
  // the 32-bit copy, when there is one, is the only pool (4 B per slot);
  // otherwise the 8-byte pool (8 B)
  const uint64_t per_label = use32 ? 4 : 8;
  // the pool kept from the previous build on this graph is taken over when
  // no initial pool size is forced and it fits the budget (no allocation,
  // and no growth if this build needs no more); otherwise it is freed
  {
    const uint64_t kept = use32 ? mg->d_pool32_bytes / 4 : mg->d_pool_bytes / 8;
    const bool adopt = !opts->pool_init_labels && (use32 ? mg->d_pool32 : mg->d_pool) &&
                       kept >= pool0 &&
                       g.fixed + kept * per_label + gtile0 * 4 + (uint64_t)list0 * 36 <= g.budget;
    if (adopt) {
      pool0 = kept;
      if (use32) {
        g.pool32.p = mg->d_pool32;
        g.pool32.bytes = mg->d_pool32_bytes;
        mg->d_pool32 = nullptr;
        mg->d_pool32_bytes = 0;
      } else {
        g.pool.p = mg->d_pool;
        g.pool.bytes = mg->d_pool_bytes;
        mg->d_pool = nullptr;
        mg->d_pool_bytes = 0;
      }
    }
    if (mg->d_pool) (void)hipFree(mg->d_pool);
    if (mg->d_pool32) (void)hipFree(mg->d_pool32);
    mg->d_pool = mg->d_pool32 = nullptr;
    mg->d_pool_bytes = mg->d_pool32_bytes = 0;
  }
  if (g.budget < g.fixed + pool0 * per_label + gtile0 * 4 + (uint64_t)list0 * 36)
    return fail(HSDL_ENOMEM, "HBM budget below the label pool, slots and lists");

  if ((!use32 && !g.pool.p && (rc = g.pool.alloc(pool0 * 8))) ||
      (rc = g.take(kScSegStart, n * 8ull)) ||
      (rc = g.take(kScSegLen, n * 4ull)) || (rc = g.take(kScSegCap, n * 4ull)) ||
      (rc = g.take(kScAddCnt, n * 4ull)) ||
      (rc = g.take(kScTouched, n * 4ull)) || (rc = g.take(kScFgrp, n * 16ull)) ||
      (group && (rc = g.take(kScHubSlot, n * 4ull))) ||
      (rc = g.take(kScHubCcnt, (Bmax + 1ull) * 4)) || (rc = g.take(kScHubOff, (Bmax + 1ull) * 4)) ||
      (rc = g.take(kScHubCur, (Bmax + 1ull) * 4)) || (rc = g.take(kScChunkOff, (Bmax + 1ull) * 4)) ||
      (rc = g.take(kScGtile, gtile0 * 4)) || (rc = g.take(kScCtrl, sizeof(Ctrl))) ||
      (rc = g.take(kScHeavy, (uint64_t)heavy_cap * sizeof(HeavyItem))) ||
      (use32 && !g.pool32.p && (rc = g.pool32.alloc(pool0 * 4))))
    return rc;
  {
    const size_t need = (size_t)Bmax * n * 4;
    if (mg->d_slots_bytes < need) {
      if (mg->d_slots) (void)hipFree(mg->d_slots);
      mg->d_slots = nullptr;
      mg->d_slots_bytes = 0;
      mg->slot_key = 0;
      hipError_t e = hipMalloc(&mg->d_slots, need);
      if (e != hipSuccess) {
        mg->d_slots = nullptr;
        (void)hipGetLastError();
        return fail(HSDL_ENOMEM, "hipMalloc of the dense slots failed");
      }
      mg->d_slots_bytes = need;
    }
  }

  // hop zones (tuning knob HSDL_HOP_ZONES=0 reads every list whole; they
  // need K >= 2 and pay only when batches share lists, so not in the plain
  // form): sorted-zone lengths and per-hop prefix counts
  const bool zones = env_u32("HSDL_HOP_ZONES", 1, 0, 1) && K >= 2 && !plain;
  // a list is sorted again once its append zone passes 1/resort_div of its
  // sorted zone (HSDL_RESORT_DIV; R-MAT scale 18 / 20 builds: 2 0.551 /
  // 5.34 s, 4 0.463 / 4.42 s, 8 0.471 / 4.49 s, 16 0.473 / 4.66 s)
  const uint32_t resort_div = env_u32("HSDL_RESORT_DIV", 4, 1, 1u << 20);
  if (zones && ((rc = g.take(kScSortedLen, n * 4ull)) ||
                (rc = g.take(kScHcnt, (uint64_t)n * (K - 1) * 4))))
    return rc;

  // LDS tile per workgroup: 64 KB; 16-bit distances when they fit
  const uint32_t K1 = K + 1;
  g.small_vals = (uint64_t)K * (uint64_t)std::max(hg->max_w, 1) < 0xFFFFull;
  const uint64_t lds_words = 64 * 1024 / 4;  // + ~4.5 KB of static LDS; two workgroups per CU
  uint32_t lds_slots = 64;
  while ((g.small_vals ? rec_tile_words<uint16_t>(lds_slots + 64, K1)
                       : rec_tile_words<uint32_t>(lds_slots + 64, K1)) <= lds_words)
    lds_slots += 64;
  if (opts->lds_tile_slots) lds_slots = std::min(lds_slots, opts->lds_tile_slots);
  g.lds_bytes = (size_t)lds_words * 4;

  GenArgs &a = g.a;
  a.n = n;
  a.K = K;
  // accepted candidates of a work item are staged in 256 LDS slots
  a.tile = opts->tile ? std::min<uint32_t>(opts->tile, 256u) : 256u;
  // tuning knob HSDL_PRUNE_ROWS: 0 never row items, 1 when dense (default),
  // 2 always
  a.row_policy = env_u32("HSDL_PRUNE_ROWS", 1, 0, 2);
  a.row_density = env_u32("HSDL_ROW_DENSITY", 32, 1, 1u << 20);
  a.row_run = env_u32("HSDL_ROW_RUN", 4, 1, 64);
  // tuning knob HSDL_ROW_BLOCK: vertices per y-block, rounded to 1024
  a.row_block = ((env_u32("HSDL_ROW_BLOCK", 2048, 1024, 1u << 20) + 1023) / 1024) * 1024;
  // tuning knobs HSDL_HEAVY_PAIRS, HSDL_PIECE_PAIRS: (label, edge) pairs of a
  // frontier group above which it is cut into heavy pieces, and per piece
  // (a wave's light groups then hold at most 64 * heavy_pairs pairs).  Swept
  // on R-MAT scale 18 / 20, K=5, expansion seconds: 256 / 2048 0.093 / 0.91,
  // 256 / 8192 0.083 / 0.62, 512 / 16384 0.089 / 0.60, 1024 / 16384 0.099 /
  // 0.66, 128 / 1024 0.13 / 1.44 (every 64 groups of a wave wait for its
  // largest, so big light groups idle lanes; small pieces cost a claim each)
  a.heavy_pairs = env_u32("HSDL_HEAVY_PAIRS", 256, 1, 1u << 16);
  // tuning knob HSDL_CLAIM_MAX: candidates a prune wave claims at once
  // (R-MAT scale 20 builds: 2 4.33 s, 4 4.27 s, 8 4.25 s, 16 4.24 s; scale
  // 18 the same within 1 %)
  a.claim_max = env_u32("HSDL_CLAIM_MAX", 8, 1, 64);
  // indexed tiles need hub << kIdxBits to fit a word (tuning knob
  // HSDL_PRUNE_INDEXED=0: records only; 2: every LDS tile indexed)
  {
    const uint32_t mode = (uint64_t)n <= (1ull << (32 - kIdxBits))
                              ? env_u32("HSDL_PRUNE_INDEXED", 1, 0, 2) : 0u;
    a.idx_tiles = mode == 0 ? kInf32 : mode == 2 ? 0u : kRecSmall;
  }
  a.piece_pairs = env_u32("HSDL_PIECE_PAIRS", 8192, 1, 1u << 16);
  a.lds_slots = lds_slots;
  DevBuf hubs_buf, in_group;
  if (group) {
    std::vector<uint8_t> mark(n, 0);
    for (uint64_t i = 0; i < ngroup; ++i) mark[group[i]] = 1;
    if ((rc = hubs_buf.alloc(ngroup * 4)) || (rc = in_group.alloc(n))) return rc;
    HSDL_HIP(hipMemcpyAsync(hubs_buf.p, group, ngroup * 4, hipMemcpyHostToDevice, g.st));
    HSDL_HIP(hipMemcpyAsync(in_group.p, mark.data(), n, hipMemcpyHostToDevice, g.st));
    HSDL_HIP(hipStreamSynchronize(g.st));
  }
  a.hubs = group ? hubs_buf.as<uint32_t>() : nullptr;
  a.off = static_cast<const unsigned long long *>(hg->d_off);
  a.adj = static_cast<const uint2 *>(hg->d_adj);
  a.pool = use32 ? nullptr : g.pool.as<unsigned long long>();
  a.pool_cap = pool0;
  a.pool32 = use32 ? g.pool32.as<uint32_t>() : nullptr;
  a.p32_hop_bits = p32_hop_bits;
  a.p32_dist_bits = p32_dist_bits;
  a.seg_start = g.seg_start.as<unsigned long long>();
  a.seg_len = g.seg_len.as<uint32_t>();
  a.seg_cap = g.seg_cap.as<uint32_t>();
  a.sorted_len = zones ? g.sorted_len.as<uint32_t>() : nullptr;
  a.hcnt = zones ? g.hcnt.as<uint32_t>() : nullptr;
  a.hcnt_stride = K - 1;
  a.add_cnt = g.add_cnt.as<uint32_t>();
  a.touched = g.touched.as<uint32_t>();
  a.fgrp = g.fgrp.as<FGroup>();
  a.hub_slot = group ? g.hub_slot.as<uint32_t>() : nullptr;
  a.cand_d = static_cast<uint32_t *>(mg->d_slots);
  a.slot_lw = slot_lw;
  {
    // slot layout: candidate distances need bits for K * max weight, the
    // round field bits for K; the rest of the word holds the attempt tag
    const uint64_t dmax = (uint64_t)K * (uint64_t)std::max(hg->max_w, 1);
    uint32_t dist_bits = 1, rbits = 1;
    while ((1ull << dist_bits) <= dmax) ++dist_bits;
    while ((1u << rbits) <= K) ++rbits;
    a.slot_rbits = rbits;
    a.slot_shift = dist_bits + rbits;  // <= 31: dmax < 2^27, K <= 15
  }
  a.hub_ccnt = g.hub_ccnt.as<uint32_t>();
  a.hub_off = g.hub_off.as<uint32_t>();
  a.hub_cur = g.hub_cur.as<uint32_t>();
  a.chunk_off = g.chunk_off.as<uint32_t>();
  a.gtile = g.gtile.as<uint32_t>();
  a.gtile_cap = gtile0;
  a.ctrl = g.ctrl.as<Ctrl>();
  a.heavy = g.heavy.as<HeavyItem>();
  a.heavy_cap = heavy_cap;
  rc = alloc_lists(g, ccap0, fcap0);
  if (rc && (ccap0 != list0 || fcap0 != list0)) rc = alloc_lists(g, list0, list0);
  if (rc) return rc;

  {
    Ctrl c0v{};
    c0v.pool_top = (unsigned long long)n * c0;
    HSDL_HIP(hipMemcpyAsync(g.ctrl.p, &c0v, sizeof(Ctrl), hipMemcpyHostToDevice, g.st));
  }
  hipLaunchKernelGGL(k_init_labels, dim3(1024), dim3(256), 0, g.st, a, c0);
  // The slots hold the values of the last build on this graph.  When it used
  // the same slot encoding, this build continues its tag count-down (those
  // values carry larger tags, so they read as older attempts) instead of
  // clearing Bmax * n words, clearing only the rows above those the last
  // clear covered (Bmax follows the free HBM, which moves a little between
  // builds); otherwise the rows are cleared and the count-down starts again.
  const uint64_t slot_key = (1ull << 63) | ((uint64_t)n << 24) | ((uint64_t)slot_lw << 16) |
                            ((uint64_t)a.slot_rbits << 8) | a.slot_shift;
  const bool slots_kept = mg->slot_key == slot_key;
  mg->slot_key = 0;  // set again when this build ends normally
  if (!slots_kept) {
    hipLaunchKernelGGL(k_fill_u32, dim3(2048), dim3(256), 0, g.st, a.cand_d, (uint64_t)Bmax * n,
                       kInf32);
    mg->slot_rows_valid = Bmax;
  } else if (Bmax > mg->slot_rows_valid) {
    // rows [valid, Bmax) are whole windows: words [valid * n, Bmax * n)
    hipLaunchKernelGGL(k_fill_u32, dim3(2048), dim3(256), 0, g.st,
                       a.cand_d + (uint64_t)mg->slot_rows_valid * n,
                       (uint64_t)(Bmax - mg->slot_rows_valid) * n, kInf32);
    mg->slot_rows_valid = Bmax;
  }
  HSDL_HIP(hipGetLastError());

  // HSDL_BATCH_TARGET_US: the time the tuner aims one batch at (default 50 ms:
  // R-MAT scale 20 K=5 generates in 5.11 s against 5.71 s at 25 ms and 6.74 s
  // at 12.5 ms; scale 18 takes the same at 25 and 50 ms, longer at 100 ms);
  // HSDL_HUB_BATCH_MAX: the largest batch it may choose (default: the slots')
  const uint32_t tune_cap = env_u32("HSDL_HUB_BATCH_MAX", Bmax, 1, Bmax);
  // HSDL_TUNER=time tunes on batch durations (HSDL_BATCH_TARGET_US); the
  // default tunes on work, HSDL_BATCH_TARGET_WORK units (candidates plus an
  // eighth of the scanned edges) per batch, so the schedule repeats from run
  // to run.  Builds of R-MAT scale 18 / 20 K=5: 1.6e9 0.347 / 3.13 s, 8e8
  // 0.347 / 3.15 s, 4e8 0.346 / 3.23 s, 2e8 0.360 / 3.37 s, 1e8 0.381 /
  // 3.60 s (larger targets reach the slots' batch cap sooner); the 1024x1024
  // grid is the same at all of them.
  const char *tm = getenv("HSDL_TUNER");
  const bool tune_time = tm && !strcmp(tm, "time");
  const double tune_target =
      tune_time ? env_u32("HSDL_BATCH_TARGET_US", 50000, 100, 10000000) * 1e-6
                : (double)env_u32("HSDL_BATCH_TARGET_WORK", 1600000000u, 1000, 4000000000u);
  BatchTuner tuner{plain ? 1u : std::min<uint32_t>(tune_cap, 64u), tune_cap, tune_target};
  if (opts->hub_batch) tuner.B = std::min(opts->hub_batch, Bmax);
  uint64_t batches = 0, redos = 0;
  const uint64_t t_loop = now_ns();  // set-up (buffers, lists) ends here
  uint32_t s0 = 0;
  Ctrl c{}, c_ok{};  // c_ok: counters after the last batch that completed
  // attempt tags count down from tag_top; when they run out, the rows used
  // since the last clear are reset to all ones (the oldest possible value)
  const uint32_t tag_top = (uint32_t)((1ull << (32 - g.a.slot_shift)) - 2);
  uint64_t tag_next = slots_kept ? mg->slot_tag_next : tag_top;
  uint32_t rows_used = slots_kept ? mg->slot_rows_used : 0u;
  while (s0 < nhubs) {
    const uint32_t B = std::min<uint32_t>(tuner.B, nhubs - s0);
    const uint64_t tb = now_ns();
    if (tag_next > tag_top) {  // wrapped below zero
      hipLaunchKernelGGL(k_fill_u32, dim3(2048), dim3(256), 0, g.st, g.a.cand_d,
                         (((uint64_t)rows_used + (1u << slot_lw) - 1) >> slot_lw << slot_lw) * n,
                         kInf32);
      HSDL_HIP(hipGetLastError());
      tag_next = tag_top;
      rows_used = 0;
    }
    g.a.tag = (uint32_t)tag_next--;
    rows_used = std::max(rows_used, B);
    if ((rc = enqueue_batch(g, s0, B, K, nhubs))) return rc;
    const uint64_t te = now_ns();
    if ((rc = read_ctrl(g, c))) return rc;
    g.h_enqueue += (te - tb) * 1e-9;
    g.h_wait += (now_ns() - te) * 1e-9;
    if (g.timing) {
      for (uint32_t h = 0; h < K; ++h) {
        float ms1 = 0, ms2 = 0;
        HSDL_HIP(hipEventElapsedTime(&ms1, g.ev[4 * h + 0], g.ev[4 * h + 1]));
        HSDL_HIP(hipEventElapsedTime(&ms2, g.ev[4 * h + 2], g.ev[4 * h + 3]));
        g.t_expand += ms1 * 1e-3;
        g.t_prune += ms2 * 1e-3;
      }
      g.n_prune += K;
    }
    if (c.overflow) {
      // roll the batch back, grow what ran out and run it again
      const uint64_t tg = now_ns();
      GenArgs ra = g.a;
      ra.s0 = s0;
      ra.B = B;
      hipLaunchKernelGGL(k_rollback, dim3(2048), dim3(256), 0, g.st, ra);
      HSDL_HIP(hipGetLastError());
      HSDL_HIP(hipStreamSynchronize(g.st));
      const uint64_t used = std::min<uint64_t>(c.pool_top, g.a.pool_cap);
      if ((c.overflow & kOvfPool) && (rc = grow_pool(g, used, used * 2))) return rc;
      if ((c.overflow & kOvfArena) && (rc = grow_gtile(g))) return rc;
      if (c.overflow & (kOvfCand | kOvfFront)) {
        const uint64_t nc = std::max<uint64_t>((uint64_t)g.a.ccap * 2, (uint64_t)c.ccnt * 5 / 4);
        const uint64_t nf = std::max<uint64_t>((uint64_t)g.a.fcap * 2, nc);
        const uint32_t occ = g.a.ccap, ofc = g.a.fcap;
        const bool grown = nc <= 0xFFFFFFF0ull && nf <= 0xFFFFFFF0ull &&
                           alloc_lists(g, (uint32_t)nc, (uint32_t)nf) == HSDL_OK;
        if (!grown) {
          // lists of twice the size do not fit: the batch runs again at half
          // its hubs with lists of the old sizes, which fitted (a chosen
          // hub_batch or a one-hub batch cannot shrink)
          (void)hipGetLastError();
          if (B == 1 || opts->hub_batch)
            return fail(HSDL_ENOMEM, "a batch's candidate or frontier lists exceed the HBM "
                                     "budget; lower hub_batch");
          g.a.ccap = g.a.fcap = 0;  // the old buffers may be gone
          if ((rc = alloc_lists(g, occ, ofc))) return rc;
          tuner.B = std::max(1u, B / 2);
        }
      }
      // the counters restart from the last completed batch (the attempt's
      // work is redone); the prune's read counters keep it, matching the
      // kernel time that includes it
      Ctrl fix = c;
      fix.overflow = 0;
      fix.pool_top = used;
      fix.cand_total = c_ok.cand_total;
      fix.item_total = c_ok.item_total;
      fix.edge_total = c_ok.edge_total;
      HSDL_HIP(hipMemcpyAsync(g.ctrl.p, &fix, sizeof(Ctrl), hipMemcpyHostToDevice, g.st));
      HSDL_HIP(hipStreamSynchronize(g.st));
      g.h_grow += (now_ns() - tg) * 1e-9;
      if (g.verbose)
        fprintf(stderr, "[hsdl] batch s0=%u B=%u overflow=%u: pool %llu, arena %llu, lists %u/%u; redo\n",
                s0, B, c.overflow, (unsigned long long)g.a.pool_cap,
                (unsigned long long)g.a.gtile_cap, g.a.ccap, g.a.fcap);
      if (++redos > 64) return fail(HSDL_ENOMEM, "buffers kept overflowing");
      continue;
    }
    // keep headroom: grow the pool between batches when it is half full
    if (c.pool_top * 2 > g.a.pool_cap) {
      const uint64_t tg = now_ns();
      if ((rc = grow_pool(g, c.pool_top, 0))) return rc;
      g.h_grow += (now_ns() - tg) * 1e-9;
    }
    const double work = (double)(c.cand_total - c_ok.cand_total) +
                        (double)(c.edge_total - c_ok.edge_total) / 8.0;
    c_ok = c;
    // (no prune follows the last batch: its lists stay as they are)
    if (zones && B > 1 && s0 + B < nhubs && (rc = sort_zones(g, c.maxlen, resort_div)))
      return rc;
    s0 += B;
    ++batches;
    const double dt = (now_ns() - tb) * 1e-9;
    if (g.verbose)
      fprintf(stderr, "[hsdl] batch %llu s0=%u B=%u %.3f ms cand=%llu pool=%llu/%llu\n",
              (unsigned long long)batches, s0 - B, B, dt * 1e3,
              (unsigned long long)c.cand_total, (unsigned long long)c.pool_top,
              (unsigned long long)g.a.pool_cap);
    if (!plain && !opts->hub_batch) tuner.next(tune_time ? dt : work);
    if (progress_s > 0) {
      const uint64_t now = now_ns();
      if (!next_progress) next_progress = t_csr + (uint64_t)(progress_s * 1e9);
      if (now >= next_progress) {
        fprintf(stderr, "[hsdl] progress: %u of %u hubs, %llu batches, %.1f s\n", s0, nhubs,
                (unsigned long long)batches, (now - t_csr) * 1e-9);
        fflush(stderr);
        next_progress = now + (uint64_t)(progress_s * 1e9);
      }
    }
    if (opts->deadline_s > 0 && (now_ns() - t_start) * 1e-9 > opts->deadline_s)
      return fail(HSDL_ETIMEDOUT, "generation passed its deadline");
  }
  if (group) {
    hipLaunchKernelGGL(k_drop_foreign_self, dim3(1024), dim3(256), 0, g.st, a,
                       in_group.as<uint8_t>());
    HSDL_HIP(hipGetLastError());
  }
  mg->slot_key = slot_key;
  mg->slot_tag_next = tag_next;
  mg->slot_rows_used = rows_used;
  const uint64_t t_gen = now_ns();
  // the batch buffers stay for the graph's next build unless the export
  // (labels, and the sort's double buffer and scratch) needs their memory
  {
    uint64_t total_est = 0;
    if (hipMemcpy(&total_est, &g.ctrl.as<Ctrl>()->pool_top, 8, hipMemcpyDeviceToHost) !=
        hipSuccess)
      return fail(HSDL_EDEVICE, "reading the pool top failed");
    size_t fb = 0, tb = 0;
    HSDL_HIP(hipMemGetInfo(&fb, &tb));
    if ((uint64_t)fb < total_est * 24 + (1ull << 30)) {
      g.cflat.release();
      g.cgroup.release();
      g.F0.release();
      g.F1.release();
      g.gtile.release();
    }
  }

  // final CSR sorted per vertex
  hsdl_index *idx = new hsdl_index();
  idx->device = dev;
  idx->n = n;
  idx->K = K;
  auto cleanup = [&](int code) {
    delete idx;
    return code;
  };
  if ((rc = idx->off.alloc((n + 1ull) * 8))) return cleanup(rc);
  if ((rc = scan_counts_u32(a.seg_len, idx->off.as<uint64_t>(), n, g.st))) return cleanup(rc);
  uint64_t total = 0;
  if (hipMemcpy(&total, idx->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost) != hipSuccess)
    return cleanup(fail(HSDL_EDEVICE, "reading label total failed"));
  idx->total = total;
  if ((rc = idx->lab.alloc(std::max<uint64_t>(total, 1) * 8))) return cleanup(rc);
  const uint64_t t_sort0 = now_ns();
  // the 32-bit path when the lists have 32-bit copies (tuning knob
  // HSDL_EXPORT32=0: the 8-byte path); its buffers stay with the graph
  size_t exp_tmp_bytes = 0;
  const uint32_t key_bits = p32_hub_bits + p32_hop_bits + p32_dist_bits;
  bool sorted32 = false;
  if (a.pool32 && total && env_u32("HSDL_EXPORT32", 1, 0, 1) &&
      segmented_sort_u32(nullptr, exp_tmp_bytes, nullptr, nullptr, total, nullptr, n, key_bits,
                         g.st) == HSDL_OK &&
      g.take(kScExpKeys, total * 4) == HSDL_OK && g.take(kScExpSorted, total * 4) == HSDL_OK &&
      g.take(kScExpTmp, std::max<size_t>(exp_tmp_bytes, 16)) == HSDL_OK) {
    hipLaunchKernelGGL(k_gather32, dim3(g.grid), dim3(256), 0, g.st, a,
                       idx->off.as<unsigned long long>(), g.exp_keys.as<uint32_t>());
    HSDL_HIP(hipGetLastError());
    if ((rc = segmented_sort_u32(g.exp_tmp.p, exp_tmp_bytes, g.exp_keys.as<uint32_t>(),
                                 g.exp_sorted.as<uint32_t>(), total, idx->off.as<uint64_t>(), n,
                                 key_bits, g.st)))
      return cleanup(rc);
    hipLaunchKernelGGL(k_widen32, dim3(g.grid), dim3(256), 0, g.st, a,
                       g.exp_sorted.as<uint32_t>(), total, idx->lab.as<unsigned long long>());
    HSDL_HIP(hipGetLastError());
    if (hipStreamSynchronize(g.st) != hipSuccess)
      return cleanup(fail(HSDL_EDEVICE, "export failed"));
    sorted32 = true;
  }
  if (!sorted32) {
    hipLaunchKernelGGL(k_gather, dim3(g.grid), dim3(256), 0, g.st, a,
                       idx->off.as<unsigned long long>(), idx->lab.as<unsigned long long>());
    if (hipStreamSynchronize(g.st) != hipSuccess)
      return cleanup(fail(HSDL_EDEVICE, "gather failed"));
    // the pools stay for the graph's next build; they are released first
    // only when the sort cannot get its buffers otherwise
    rc = segmented_sort_labels(idx->lab.as<uint64_t>(), total, idx->off.as<uint64_t>(), n, g.st);
    if (rc == HSDL_ENOMEM) {
      g.pool.release();
      g.pool32.release();
      rc = segmented_sort_labels(idx->lab.as<uint64_t>(), total, idx->off.as<uint64_t>(), n, g.st);
    }
    if (rc) return cleanup(rc);
  }
  const uint64_t t_sorted = now_ns();
  // parents (optional): derived from the sorted uncleaned labels and the graph
  if (opts->parents && (rc = derive_parents(idx, hg, g.st))) return cleanup(rc);
  const uint64_t t_end = now_ns();
  idx->stats.t_parents_s = (t_end - t_sorted) * 1e-9;
  if (g.verbose)
    fprintf(stderr, "[hsdl] generate: %.3f s set-up, %.3f s batches (%llu, %llu redo; host enqueue "
            "%.3f s, wait %.3f s, grow %.3f s; %llu graph captures), %.3f s export (%.3f s sort), "
            "%llu labels\n",
            (t_loop - t_csr) * 1e-9, (t_gen - t_loop) * 1e-9, (unsigned long long)batches,
            (unsigned long long)redos,
            g.h_enqueue, g.h_wait, g.h_grow, (unsigned long long)g.graph_captures,
            (t_end - t_gen) * 1e-9, (t_end - t_sort0) * 1e-9, (unsigned long long)total);
  mg->hint_ccap = g.a.ccap;
  mg->hint_fcap = g.a.fcap;
  mg->hint_arena = g.a.gtile_cap;
  idx->stats.labels_uncleaned = total;
  idx->dist_max = (uint32_t)std::min<uint64_t>(kMaxDist, (uint64_t)K * std::max(hg->max_w, 1));
  idx->stats.t_generate_s = (t_end - t_csr) * 1e-9;
  idx->stats.t_sort_s = (t_sorted - t_sort0) * 1e-9;
  idx->stats.hub_batch = tuner.B;
  idx->stats.tile = a.tile;
  idx->stats.rounds = K;
  idx->stats.batches = batches;
  idx->stats.candidates = c.cand_total;
  idx->stats.frontier_items = c.item_total;
  idx->stats.scanned_edges = c.edge_total;
  idx->stats.t_expand_s = g.t_expand;
  idx->stats.t_prune_s = g.t_prune;
  idx->stats.prune_launches = g.n_prune;
  idx->stats.prune_label_reads = c.label_reads;
  idx->stats.prune_tile_reads = c.tile_reads;
  idx->stats.prune_row_slots = c.row_slots;
  idx->stats.prune_label_bytes = a.pool32 ? 4 : 8;
  // the pool stays with the graph for its next build (one released for the
  // sort leaves nothing to keep: the next build allocates anew)
  mg->d_pool = g.pool.p;
  mg->d_pool_bytes = g.pool.bytes;
  g.pool.p = nullptr;
  g.pool.bytes = 0;
  mg->d_pool32 = g.pool32.p;
  mg->d_pool32_bytes = g.pool32.bytes;
  g.pool32.p = nullptr;
  g.pool32.bytes = 0;
  *out = idx;
  return HSDL_OK;
}

}  // namespace hsdl

extern "C" int hsdl_graph_release_device(hsdl_graph *g) {
  if (!g) return hsdl::fail(HSDL_EINVAL, "null graph");
  g->drop_device_copy();
  return HSDL_OK;
}

static int graph_upload_body(hsdl_graph *g, int32_t device) {
  if (!g) return hsdl::fail(HSDL_EINVAL, "null graph");
  int dev = 0;
  int rc = hsdl::use_device(device, &dev);
  if (rc) return rc;
  return hsdl::upload_graph(g, dev);
}

extern "C" int hsdl_graph_upload(hsdl_graph *g, int32_t device) {
  return hsdl::guarded([&] { return graph_upload_body(g, device); });
}

static int generate_body(const hsdl_graph *g, uint8_t K, const hsdl_opts *opts,
                             hsdl_index **out) {
  if (!g || !out) return hsdl::fail(HSDL_EINVAL, "null argument");
  hsdl_opts o;
  if (opts)
    o = *opts;
  else
    hsdl_opts_init(&o);
  return hsdl::generate(g, K, &o, nullptr, 0, out);
}

extern "C" int hsdl_generate(const hsdl_graph *g, uint8_t K, const hsdl_opts *opts,
                             hsdl_index **out) {
  return hsdl::guarded([&] { return generate_body(g, K, opts, out); });
}

extern "C" int hsdl_generate_group(const hsdl_graph *g, uint8_t K, const hsdl_opts *opts,
                                   const uint32_t *hubs, uint64_t nhubs, hsdl_index **out) {
  return hsdl::guarded([&] {
    if (!g || !out || !hubs) return hsdl::fail(HSDL_EINVAL, "null argument");
    hsdl_opts o;
    if (opts)
      o = *opts;
    else
      hsdl_opts_init(&o);
    return hsdl::generate(g, K, &o, hubs, nhubs, out);
  });
}
