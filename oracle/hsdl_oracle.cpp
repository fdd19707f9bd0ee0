// hsdl_oracle.cpp -- CPU restatement of the reference's HSDL path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
// entry point and bench.py's cpu_baseline leg are the only callers.  The
// product library (gpu4hophl_code_amd/) never links, loads or calls it.
//
// Every function restates one piece of the reference (rucdatascience/
// GPU4HopHL_code) from its published semantics; each cites the file:line it
// follows.  Nothing here is copied: the data structures (dense per-thread
// tables, an indexed binary heap instead of boost::heap::fibonacci_heap, flat
// CSR inputs) are this project's own.
//
// Pinning (what checks each function against something other than itself):
//   orc_kround      -> equals brute-force canonical labels after cleaning
//                      (orc_canonical_bruteforce) on thousands of tiny seeded
//                      graphs; label invariants (hub<=v, self label, strictly
//                      decreasing dist per hub); queries == Bellman-Ford.
//   orc_dijkstra    -> clean(orc_dijkstra) == clean(orc_kround) == brute force.
//   orc_clean       -> brute-force canonical closed form; both L(u) readings
//                      (already-cleaned vs input) give the same output.
//   orc_query       -> hop-bounded Bellman-Ford (orc_bf) on every (s,t,k) of
//                      tiny graphs and sampled pairs of larger ones.
//   orc_bf          -> Floyd-style min-plus matrix powers on tiny graphs
//                      (tests/test_oracle_pins.py computes those in numpy).
//   orc_load_txt    -> the reference fixture files (|V|=, Edge u v w) and the
//                      reference's last-wins duplicate rule.
//   mutations       -> every function takes a `mutation` code; tests plant a
//                      mistake through it and require a pin to fail.
//
// Build: oracle/Makefile (g++ -O2 -shared -fPIC -pthread).

#include <algorithm>
#include <atomic>
#include <climits>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

constexpr int32_t INF32 = INT32_MAX;

struct Label {
  int32_t hub, hop, dist;
  int32_t parent;  // K-round generation only (0 elsewhere): see kround()
};

struct Graph {
  int32_t n = 0;
  std::vector<int64_t> off;
  std::vector<int32_t> col;
  std::vector<int32_t> w;
  int64_t deg(int32_t v) const { return off[v + 1] - off[v]; }
};

using LabelSet = std::vector<std::vector<Label>>;

// Per-vertex order (hub asc, hop asc, dist asc):
// CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:373-387.
inline bool label_less(const Label &a, const Label &b) {
  if (a.hub != b.hub) return a.hub < b.hub;
  if (a.hop != b.hop) return a.hop < b.hop;
  return a.dist < b.dist;
}

struct Labels {
  LabelSet L;
  // Algorithmic-byte counters filled by the K-round generator (logical sizes:
  // 9 B per label, 8 B per CSR entry, 8 B per frontier item).
  int64_t frontier_items = 0;    // sum over rounds of |F_{h-1}|
  int64_t scanned_edges = 0;     // sum over frontier items of deg(x)
  int64_t candidates = 0;        // deduplicated (s,y) candidates
  int64_t cand_label_reads = 0;  // sum over candidates of |L_{<h}(y)|
  int64_t hub_label_reads = 0;   // sum over active hubs of |L_{<h}(s)|
  int64_t accepted = 0;          // labels added in rounds 1..K
  int64_t rounds_run = 0;
};

void sort_all(LabelSet &L) {
  for (auto &v : L) std::sort(v.begin(), v.end(), label_less);
}

int nthreads_or_default(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class F>
void parallel_for(int64_t n, int threads, F &&fn) {
  threads = nthreads_or_default(threads);
  if (threads <= 1 || n < 2) {
    for (int64_t i = 0; i < n; ++i) fn(0, i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (;;) {
        int64_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(t, i);
      }
    });
  }
  for (auto &th : pool) th.join();
}

// Pruning query shared by both generators and the clean:
//   min { da + db : (a,ha,da) in Lx, (a,hb,db) in T[a], ha + hb <= hop }
// where T is the dense hub-indexed table built from the other side's labels.
// CPU_HopHL/.../hop_constrained_two_hop_labels_generation.h:127-144 (Dijkstra),
// Hybrid_Generation_Clean/.../hop_constrained_two_hop_labels_generation.h:199-212
// (K-round), CPU_HopHL/...generation.h:460-475 (clean).
struct HubTable {
  std::vector<std::vector<std::pair<int32_t, int32_t>>> t;  // hub -> (dist, hop)
  std::vector<int32_t> touched;
  explicit HubTable(int32_t n) : t(n) {}
  void add(int32_t hub, int32_t dist, int32_t hop) {
    if (t[hub].empty()) touched.push_back(hub);
    t[hub].push_back({dist, hop});
  }
  void clear() {
    for (int32_t h : touched) t[h].clear();
    touched.clear();
  }
};

int64_t pair_min(const std::vector<Label> &Lx, const HubTable &T, int32_t hop,
                 int mutation) {
  int64_t q = INF32;
  for (const Label &x : Lx) {
    for (const auto &yy : T.t[x.hub]) {
      bool ok = (mutation == 1) ? (x.hop + yy.second < hop)
                                : (x.hop + yy.second <= hop);
      if (ok) {
        int64_t d = (int64_t)x.dist + yy.first;
        if (d < q) q = d;
      }
    }
  }
  return q;
}

// ---------------------------------------------------------------------------
// K-round frontier generation: the CPU twin of the GPU algorithm,
// HSDL_with_large_queue + its driver,
// Hybrid_Generation_Clean/include/HBPLL/hop_constrained_two_hop_labels_generation.h:136-243,1291-1322
// (round driver), with the GPU round semantics of
// Hybrid_Generation_Clean/src/label_gen.cu:477-593 (expansion, dedupe by min,
// rank pruning) and :87-110 (accept iff candidate < query).
//
// Labels committed in round h carry hop h; such a label can only pair with a
// hop-0 label (a self label) in a round-h query, which the rank order rules
// out, so committing after the round (here) and committing live (reference)
// give the same set.
// ---------------------------------------------------------------------------
// `group` > 0 restates the reference's vertex-group mode
// (Hybrid_Generation_Clean/src/label_gen.cu:647 clears the device labels per
// group, Hybrid_Generation_Clean/src/test.cu:310-328 cuts hubs into groups of
// G_max ids): a hub's pruning query only sees labels whose hub lies in its own
// group, which leaves redundant labels for the clean to remove.
// `gid` (one group id per vertex) generalises the contiguous groups to the
// reference's arbitrary vertex groups (Hybrid_Generation_Clean/src/test.cu:
// 338-341, generate_Group_CDLP): a hub's query sees only labels whose hub has
// its group id.
Labels *kround(const Graph &g, int K, int threads, int mutation, int group = 0,
               const int32_t *gid = nullptr) {
  const int32_t n = g.n;
  Labels *res = new Labels();
  LabelSet &L = res->L;
  L.assign(n, {});
  std::vector<std::vector<std::pair<int32_t, int32_t>>> F(n), Fn(n);  // (x, dx)
  for (int32_t v = 0; v < n; ++v) {
    L[v].push_back({v, 0, 0, v});
    F[v].push_back({v, 0});
  }
  threads = nthreads_or_default(threads);
  struct Scratch {
    std::vector<int32_t> tempD, tempP;
    HubTable T;
    std::vector<int32_t> D;
    int64_t c_items = 0, c_edges = 0, c_cand = 0, c_cand_reads = 0,
            c_hub_reads = 0, c_acc = 0;
    explicit Scratch(int32_t n) : tempD(n, INF32), tempP(n, INF32), T(n) {}
  };
  std::vector<Scratch *> scr;
  for (int t = 0; t < threads; ++t) scr.push_back(new Scratch(n));
  // per-hub accepted labels of the current round: (y, dist, parent)
  struct Acc {
    int32_t y, d, p;
  };
  std::vector<std::vector<Acc>> acc(n);

  for (int h = 1; h <= K; ++h) {
    std::atomic<bool> any{false};
    parallel_for(n, threads, [&](int tid, int64_t si) {
      const int32_t s = (int32_t)si;
      Scratch &S = *scr[tid];
      acc[s].clear();
      if (F[s].empty()) return;
      // snapshot L(s) into the hub table (reference :164-170)
      for (const Label &x : L[s])
        if (gid ? gid[x.hub] == gid[s] : (group <= 0 || x.hub / group == s / group))
          S.T.add(x.hub, x.dist, x.hop);
      S.c_hub_reads += (int64_t)L[s].size();
      // expansion with min-dedupe (reference :172-194)
      for (const auto &fx : F[s]) {
        const int32_t x = fx.first, dx = fx.second;
        S.c_items += 1;
        S.c_edges += g.deg(x);
        for (int64_t e = g.off[x]; e < g.off[x + 1]; ++e) {
          const int32_t y = g.col[e];
          if (s > y) continue;  // rank pruning (reference :178-180)
          const int32_t nd = dx + g.w[e];
          // mutation 2 keeps the first distance seen instead of the minimum
          if (nd < S.tempD[y] && (mutation != 2 || S.tempD[y] == INF32)) {
            if (S.tempD[y] == INF32) S.D.push_back(y);
            S.tempD[y] = nd;
            S.tempP[y] = x;
          } else if (nd == S.tempD[y]) {
            // parent: the smallest frontier vertex at the minimum distance
            // (the reference keeps the first one its scan meets,
            // Hybrid_Generation_Clean/src/label_gen.cu:531-541, which depends
            // on the scan order); mutation 7 keeps the largest instead
            if (mutation == 7 ? x > S.tempP[y] : x < S.tempP[y]) S.tempP[y] = x;
          }
        }
      }
      if (!S.D.empty()) any = true;
      // pruning check (reference :197-225)
      for (int32_t y : S.D) {
        const int32_t dy = S.tempD[y];
        S.c_cand += 1;
        S.c_cand_reads += (int64_t)L[y].size();
        int64_t q = pair_min(L[y], S.T, h, mutation);
        bool accept = (mutation == 3) ? (dy <= q) : (dy < q);
        if (accept) acc[s].push_back({y, dy, S.tempP[y]});
        S.tempD[y] = INF32;
        S.tempP[y] = INF32;
      }
      S.D.clear();
      S.T.clear();
    });
    // commit the round: labels (s, h, d) to L(y); next frontier of s
    int64_t acc_round = 0;
    for (int32_t s = 0; s < n; ++s) {
      Fn[s].clear();
      for (const auto &a : acc[s]) {
        L[a.y].push_back({s, h, a.d, a.p});
        Fn[s].push_back({a.y, a.d});
      }
      acc_round += (int64_t)acc[s].size();
    }
    res->accepted += acc_round;
    std::swap(F, Fn);
    res->rounds_run = h;
    if (!any) break;  // reference :1302 (global_label_generation_tag)
  }
  for (auto *S : scr) {
    res->frontier_items += S->c_items;
    res->scanned_edges += S->c_edges;
    res->candidates += S->c_cand;
    res->cand_label_reads += S->c_cand_reads;
    res->hub_label_reads += S->c_hub_reads;
    delete S;
  }
  sort_all(L);
  return res;
}

// ---------------------------------------------------------------------------
// 2023WWW generation, _2023WWW_thread_function + its driver:
// CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:226-370
// (per hub), :513-604 (hubs 0..N-1).  Run single-threaded in hub order, the
// reference's output with thread_num = 1.  Per hub v_k: a level-synchronous
// search over hops h = 0..K with one distance-so-far per vertex (dist_hop,
// :262-279, improved by the previous level's pushes at the start of a
// level); vertex u of level h is labelled (v_k, h, dist(u)) iff dist(u) is
// below the pair minimum of L(u) and the snapshot of L(v_k) taken before the
// search (:244-256, hop sum <= h, :289-303); a labelled u pushes every
// neighbour it improves (:318-327).  No rank pruning (:284-288 is commented
// out in the reference).  The reference's level lists may hold a vertex
// twice; the second visit makes the same decision and at most repeats a
// label, so the restatement keeps each level as a set.
// ---------------------------------------------------------------------------
Labels *www2023(const Graph &g, int K, int mutation) {
  const int32_t n = g.n;
  Labels *res = new Labels();
  LabelSet &L = res->L;
  L.assign(n, {});
  std::vector<int32_t> dist(n, INF32), stamp(n, -1), changed;
  HubTable T(n);
  std::vector<int32_t> cur, nxt;
  std::vector<std::pair<int32_t, int32_t>> updates;  // (vertex, distance)
  for (int32_t vk = 0; vk < n; ++vk) {
    for (const Label &x : L[vk]) T.add(x.hub, x.dist, x.hop);
    cur.assign(1, vk);
    dist[vk] = 0;
    changed.push_back(vk);
    for (int h = 0; h <= K; ++h) {
      for (const auto &u : updates) {
        if (u.second < dist[u.first]) {
          if (dist[u.first] == INF32) changed.push_back(u.first);
          dist[u.first] = u.second;
        }
      }
      updates.clear();
      nxt.clear();
      for (int32_t u : cur) {
        const int32_t Pu = dist[u];
        const int64_t q = pair_min(L[u], T, h, 0);
        if (Pu < q) {
          L[u].push_back({vk, h, Pu, 0});
          for (int64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
            const int32_t x = g.col[e];
            const int64_t nd = (int64_t)Pu + g.w[e];
            // mutation 8 also pushes a neighbour the walk does not improve
            if (nd < dist[x] || (mutation == 8 && nd == dist[x])) {
              if (stamp[x] != h + 1) {
                stamp[x] = h + 1;
                nxt.push_back(x);
              }
              updates.push_back({x, (int32_t)nd});
            }
          }
        }
      }
      std::swap(cur, nxt);
      for (int32_t x : cur) stamp[x] = -1;
    }
    updates.clear();
    for (int32_t v : changed) dist[v] = INF32;
    changed.clear();
    T.clear();
  }
  sort_all(L);
  res->accepted = 0;
  for (auto &v : L) res->accepted += (int64_t)v.size();
  return res;
}

// ---------------------------------------------------------------------------
// Dijkstra generation of CPU_HopHL (the parity target):
// HSDL_thread_function, CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:45-224,
// heap order (dist asc, hop asc) from CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:42-52,
// driver :513-604 (hubs 0..N-1), sortL :389-419.
// Run single-threaded in hub order, which is the reference's output with
// thread_num = 1 (its multi-threaded uncleaned set depends on the schedule).
// The Fibonacci heap with update() is replaced by an indexed binary heap with
// decrease-key; ties between equal (dist,hop) keys of different vertices do
// not change the label set (each (vertex,hop) state is final when popped).
// ---------------------------------------------------------------------------
struct IndexedHeap {
  // state id = vertex * (K+1) + hop ; key = (dist, hop)
  std::vector<int64_t> heap;      // state ids
  std::vector<int64_t> pos;       // state -> heap index, -1 absent
  std::vector<int32_t> key_dist;  // state -> dist
  int K1;
  IndexedHeap(int32_t n, int K) : pos((size_t)n * (K + 1), -1),
                                  key_dist((size_t)n * (K + 1), INF32), K1(K + 1) {}
  bool less(int64_t a, int64_t b) const {
    if (key_dist[a] != key_dist[b]) return key_dist[a] < key_dist[b];
    int ha = (int)(a % K1), hb = (int)(b % K1);
    if (ha != hb) return ha < hb;
    return a < b;
  }
  void up(int64_t i) {
    while (i > 0) {
      int64_t p = (i - 1) / 2;
      if (!less(heap[i], heap[p])) break;
      std::swap(heap[i], heap[p]);
      pos[heap[i]] = i;
      pos[heap[p]] = p;
      i = p;
    }
  }
  void down(int64_t i) {
    int64_t n = (int64_t)heap.size();
    for (;;) {
      int64_t l = 2 * i + 1, r = l + 1, m = i;
      if (l < n && less(heap[l], heap[m])) m = l;
      if (r < n && less(heap[r], heap[m])) m = r;
      if (m == i) break;
      std::swap(heap[i], heap[m]);
      pos[heap[i]] = i;
      pos[heap[m]] = m;
      i = m;
    }
  }
  void push(int64_t st, int32_t d) {
    key_dist[st] = d;
    heap.push_back(st);
    pos[st] = (int64_t)heap.size() - 1;
    up(pos[st]);
  }
  void decrease(int64_t st, int32_t d) {
    key_dist[st] = d;
    up(pos[st]);
  }
  int64_t pop() {
    int64_t top = heap[0];
    heap[0] = heap.back();
    pos[heap[0]] = 0;
    heap.pop_back();
    pos[top] = -2;  // popped
    if (!heap.empty()) down(0);
    return top;
  }
};

Labels *dijkstra(const Graph &g, int K, int mutation) {
  const int32_t n = g.n;
  Labels *res = new Labels();
  LabelSet &L = res->L;
  L.assign(n, {});
  HubTable T(n);
  IndexedHeap Q(n, K);
  std::vector<int64_t> changed;
  const int K1 = K + 1;
  for (int32_t vk = 0; vk < n; ++vk) {
    L[vk].push_back({vk, 0, 0, 0});  // reference :74-85
    for (const Label &x : L[vk]) T.add(x.hub, x.dist, x.hop);
    int64_t st0 = (int64_t)vk * K1 + 0;
    Q.push(st0, 0);
    changed.push_back(st0);
    while (!Q.heap.empty()) {
      int64_t st = Q.pop();
      const int32_t u = (int32_t)(st / K1);
      const int32_t uh = (int32_t)(st % K1);
      const int32_t P = Q.key_dist[st];
      if (vk > u) continue;  // reference :107-110
      int64_t q = pair_min(L[u], T, uh, mutation);
      bool add = (mutation == 3) ? (P <= q) : (P < q);
      if (add || q == 0) {
        if (add) L[u].push_back({vk, uh, P, 0});
        if (uh + 1 > K) continue;
        for (int64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
          const int32_t y = g.col[e];
          const int32_t nd = P + g.w[e];
          const int64_t sy = (int64_t)y * K1 + (uh + 1);
          if (Q.key_dist[sy] != INF32) {
            // a popped state is final: every later path to it is longer
            if (Q.key_dist[sy] > nd && Q.pos[sy] >= 0) Q.decrease(sy, nd);
          } else {
            Q.push(sy, nd);
            changed.push_back(sy);
          }
        }
      }
    }
    for (int64_t s : changed) {
      Q.key_dist[s] = INF32;
      Q.pos[s] = -1;
    }
    changed.clear();
    T.clear();
  }
  sort_all(L);
  res->accepted = 0;
  for (auto &v : L) res->accepted += (int64_t)v.size();
  return res;
}

// ---------------------------------------------------------------------------
// The threaded CPU_HopHL generation as the reference runs it, for timing:
// hop_constrained_two_hop_labels_generation's thread pool over hubs in id
// order (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:
// 560-575) running HSDL_thread_function (:45-224), with one mutex per vertex
// guarding its label list while a hub queries or extends it (:84-91,
// :128-146, :160-167).  Workers stop taking hubs once `seconds` have passed
// (0: no limit) and finish the hubs they hold.  The uncleaned set depends on
// the schedule (as the reference's does); only the time and the counts are
// used: *hubs_done (hubs finished, a prefix in id order up to the ones still
// in flight at the deadline) and *labels.
// ---------------------------------------------------------------------------
void dijkstra_timed(const Graph &g, int K, int threads, double seconds, int64_t *hubs_done,
                    int64_t *labels) {
  const int32_t n = g.n;
  LabelSet L(n);
  std::vector<std::mutex> mtx(n);
  threads = nthreads_or_default(threads);
  const int K1 = K + 1;
  std::atomic<int64_t> next{0}, done{0}, made{0};
  std::atomic<bool> stop{false};
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    HubTable T(n);
    IndexedHeap Q(n, K);
    std::vector<int64_t> changed;
    for (;;) {
      if (seconds > 0 && std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                 .count() > seconds)
        stop = true;
      if (stop) break;
      const int64_t vki = next.fetch_add(1);
      if (vki >= n) break;
      const int32_t vk = (int32_t)vki;
      {
        std::lock_guard<std::mutex> lk(mtx[vk]);
        L[vk].push_back({vk, 0, 0, 0});
        for (const Label &x : L[vk]) T.add(x.hub, x.dist, x.hop);
      }
      int64_t mine = 1;
      const int64_t st0 = (int64_t)vk * K1;
      Q.push(st0, 0);
      changed.push_back(st0);
      while (!Q.heap.empty()) {
        const int64_t st = Q.pop();
        const int32_t u = (int32_t)(st / K1);
        const int32_t uh = (int32_t)(st % K1);
        const int32_t P = Q.key_dist[st];
        if (vk > u) continue;
        int64_t q;
        bool add;
        {
          std::lock_guard<std::mutex> lk(mtx[u]);
          q = pair_min(L[u], T, uh, 0);
          add = P < q;
          if (add && u != vk) {
            L[u].push_back({vk, uh, P, 0});
            ++mine;
          }
        }
        if ((add || q == 0) && uh + 1 <= K) {
          for (int64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
            const int32_t y = g.col[e];
            const int32_t nd = P + g.w[e];
            const int64_t sy = (int64_t)y * K1 + (uh + 1);
            if (Q.key_dist[sy] != INF32) {
              if (Q.key_dist[sy] > nd && Q.pos[sy] >= 0) Q.decrease(sy, nd);
            } else {
              Q.push(sy, nd);
              changed.push_back(sy);
            }
          }
        }
      }
      for (int64_t x : changed) {
        Q.key_dist[x] = INF32;
        Q.pos[x] = -1;
      }
      changed.clear();
      T.clear();
      done += 1;
      made += mine;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  *hubs_done = done;
  *labels = made;
}

// ---------------------------------------------------------------------------
// The threaded 2023WWW generation as the reference runs it, for timing: the
// same thread pool over hubs in id order
// (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:
// 560-575) running _2023WWW_thread_function (:226-370) -- per hub a
// level-synchronous search without rank pruning, one distance-so-far per
// vertex, a vertex labelled when it beats the pair query against the hub's
// list taken before the search -- with a mutex per vertex around its list.
// Stops taking hubs after `seconds` (0: no limit); returns the hubs finished
// and the labels made (schedule-dependent, as the reference's).
// ---------------------------------------------------------------------------
void www2023_timed(const Graph &g, int K, int threads, double seconds, int64_t *hubs_done,
                   int64_t *labels) {
  const int32_t n = g.n;
  LabelSet L(n);
  std::vector<std::mutex> mtx(n);
  threads = nthreads_or_default(threads);
  std::atomic<int64_t> next{0}, done{0}, made{0};
  std::atomic<bool> stop{false};
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    HubTable T(n);
    std::vector<int32_t> dist(n, INF32), stamp(n, -1), changed, cur, nxt;
    std::vector<std::pair<int32_t, int32_t>> updates;
    for (;;) {
      if (seconds > 0 && std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                 .count() > seconds)
        stop = true;
      if (stop) break;
      const int64_t vki = next.fetch_add(1);
      if (vki >= n) break;
      const int32_t vk = (int32_t)vki;
      {
        std::lock_guard<std::mutex> lk(mtx[vk]);
        for (const Label &x : L[vk]) T.add(x.hub, x.dist, x.hop);
      }
      int64_t mine = 0;
      cur.assign(1, vk);
      dist[vk] = 0;
      changed.push_back(vk);
      for (int h = 0; h <= K; ++h) {
        for (const auto &u : updates)
          if (u.second < dist[u.first]) {
            if (dist[u.first] == INF32) changed.push_back(u.first);
            dist[u.first] = u.second;
          }
        updates.clear();
        nxt.clear();
        for (int32_t u : cur) {
          const int32_t Pu = dist[u];
          bool add;
          {
            std::lock_guard<std::mutex> lk(mtx[u]);
            add = Pu < pair_min(L[u], T, h, 0);
            if (add) {
              L[u].push_back({vk, h, Pu, 0});
              ++mine;
            }
          }
          if (!add) continue;
          for (int64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
            const int32_t x = g.col[e];
            const int64_t nd = (int64_t)Pu + g.w[e];
            if (nd < dist[x]) {
              if (stamp[x] != h + 1) {
                stamp[x] = h + 1;
                nxt.push_back(x);
              }
              updates.push_back({x, (int32_t)nd});
            }
          }
        }
        std::swap(cur, nxt);
        for (int32_t x : cur) stamp[x] = -1;
      }
      updates.clear();
      for (int32_t x : changed) dist[x] = INF32;
      changed.clear();
      T.clear();
      done += 1;
      made += mine;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  *hubs_done = done;
  *labels = made;
}

// ---------------------------------------------------------------------------
// Canonical repair, hop_constrained_clean_L:
// CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:422-511.
// Walk L(v) in (hub,hop,dist) order with kept := {}; keep (u,h,d) iff
//   min { d1 + d2 : (w,h1,d1) in L(u), (w,h2,d2) in kept, h1 + h2 <= h } > d.
// `lu_mode` chooses which L(u) is read for u < v:
//   0: the already-cleaned list (single-threaded reference: vertices are
//      cleaned in id order, so L(u) for u < v has been replaced);
//   1: the input (uncleaned) list, as a GPU clean that reads only its input.
// For u == v the input list is read in both modes (the reference replaces
// L(v) only after the walk).  Tests require the two modes to agree.
// ---------------------------------------------------------------------------
LabelSet clean(const LabelSet &Lin, int lu_mode, int threads, int mutation) {
  const int32_t n = (int32_t)Lin.size();
  LabelSet out(n);
  if (lu_mode == 1) {
    threads = nthreads_or_default(threads);
    std::vector<HubTable *> tabs;
    for (int t = 0; t < threads; ++t) tabs.push_back(new HubTable(n));
    parallel_for(n, threads, [&](int tid, int64_t vi) {
      const int32_t v = (int32_t)vi;
      HubTable &T = *tabs[tid];
      for (const Label &l : Lin[v]) {
        int64_t m = pair_min(Lin[l.hub], T, l.hop, mutation);
        bool keep = (mutation == 4) ? (m >= l.dist) : (m > l.dist);
        if (keep) {
          out[v].push_back(l);
          T.add(l.hub, l.dist, l.hop);
        }
      }
      T.clear();
    });
    for (auto *t : tabs) delete t;
  } else {
    HubTable T(n);
    for (int32_t v = 0; v < n; ++v) {
      for (const Label &l : Lin[v]) {
        const std::vector<Label> &Lu = (l.hub < v) ? out[l.hub] : Lin[l.hub];
        int64_t m = pair_min(Lu, T, l.hop, mutation);
        bool keep = (mutation == 4) ? (m >= l.dist) : (m > l.dist);
        if (keep) {
          out[v].push_back(l);
          T.add(l.hub, l.dist, l.hop);
        }
      }
      T.clear();
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Query, hop_constrained_extract_distance:
// CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:144-222.
// Merge-join of hub-sorted L(s), L(t); INT32_MAX when unreachable, k<0 or
// (k==0 and s!=t); 0 when s==t.  The reference dereferences its end iterator
// before testing it (:173,178); this restatement tests first.
// ---------------------------------------------------------------------------
int32_t query_one(const LabelSet &L, int32_t s, int32_t t, int32_t k, int mutation) {
  if (k < 0) return INF32;
  if (s == t) return 0;
  if (k == 0) return INF32;
  int64_t best = INF32;
  const auto &A = L[s];
  const auto &B = L[t];
  size_t i = 0, j = 0;
  while (i < A.size() && j < B.size()) {
    if (A[i].hub == B[j].hub) {
      size_t ie = i, je = j;
      while (ie < A.size() && A[ie].hub == A[i].hub) ++ie;
      while (je < B.size() && B[je].hub == B[j].hub) ++je;
      for (size_t a = i; a < ie; ++a) {
        for (size_t b = j; b < je; ++b) {
          bool ok = (mutation == 1) ? (A[a].hop + B[b].hop < k)
                                    : (A[a].hop + B[b].hop <= k);
          if (!ok) break;  // L(t) hops ascend within a hub
          int64_t d = (int64_t)A[a].dist + B[b].dist;
          if (d < best) best = d;
        }
      }
      i = ie;
      j = je;
    } else if (A[i].hub > B[j].hub) {
      ++j;
    } else {
      ++i;
    }
  }
  return (int32_t)best;
}

// ---------------------------------------------------------------------------
// Hop-constrained shortest path, hop_constrained_extract_shortest_path:
// CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:224-319.
// One step merges L(s) and L(t) like the distance query and takes the first
// pair, in the merge's order (hub, then position in L(s), then position in
// L(t)), whose dist sum is smaller than every earlier one among pairs with
// hop sum <= k.  It emits (s, parent of the L(s) label) unless that parent is
// s, then (t, parent of the L(t) label) unless that parent is t, moves s and
// t to those parents and lowers k by one per emitted edge; the reference
// recurses, this loops.  It stops when s == t or no pair qualifies (a step
// that moves neither endpoint also stops it: the reference would recurse
// forever, which its own parents never cause).  Mutation 6 takes the last
// pair at the minimum instead of the first.
// ---------------------------------------------------------------------------
void path_one(const LabelSet &L, int32_t s, int32_t t, int32_t k,
              std::vector<std::pair<int32_t, int32_t>> &out, int mutation) {
  for (;;) {
    if (s == t) return;
    int64_t best = INF32;
    int32_t p1 = s, p2 = t;
    const auto &A = L[s];
    const auto &B = L[t];
    size_t i = 0, j = 0;
    while (i < A.size() && j < B.size()) {
      if (A[i].hub == B[j].hub) {
        size_t ie = i, je = j;
        while (ie < A.size() && A[ie].hub == A[i].hub) ++ie;
        while (je < B.size() && B[je].hub == B[j].hub) ++je;
        for (size_t a = i; a < ie; ++a) {
          for (size_t b = j; b < je; ++b) {
            if (A[a].hop + B[b].hop > k) break;
            int64_t d = (int64_t)A[a].dist + B[b].dist;
            if (d < best || (mutation == 6 && d == best)) {
              best = d;
              p1 = A[a].parent;
              p2 = B[b].parent;
            }
          }
        }
        i = ie;
        j = je;
      } else if (A[i].hub > B[j].hub) {
        ++j;
      } else {
        ++i;
      }
    }
    if (best >= INF32) return;
    if (p1 == s && p2 == t) return;
    if (s != p1) {
      out.push_back({s, p1});
      s = p1;
      --k;
    }
    if (t != p2) {
      out.push_back({t, p2});
      t = p2;
      --k;
    }
  }
}

// ---------------------------------------------------------------------------
// Hop-bounded shortest distances from one source:
// CPU_HopHL/include/graph_v_of_v/graph_v_of_v_hop_constrained_shortest_distance.h:7-38
// restated as K rounds of Bellman-Ford over a per-round copy (distance[v] is
// the shortest walk with <= k edges; INT32_MAX when none).
// ---------------------------------------------------------------------------
void bf(const Graph &g, int32_t src, int k, int32_t *dist, int mutation) {
  const int32_t n = g.n;
  std::vector<int32_t> cur(n, INF32), nxt;
  cur[src] = 0;
  int rounds = (mutation == 5) ? k - 1 : k;
  for (int r = 0; r < rounds; ++r) {
    nxt = cur;
    for (int32_t x = 0; x < n; ++x) {
      if (cur[x] == INF32) continue;
      for (int64_t e = g.off[x]; e < g.off[x + 1]; ++e) {
        int32_t y = g.col[e];
        int64_t nd = (int64_t)cur[x] + g.w[e];
        if (nd < nxt[y]) nxt[y] = (int32_t)nd;
      }
    }
    cur.swap(nxt);
  }
  std::memcpy(dist, cur.data(), sizeof(int32_t) * n);
}

// ---------------------------------------------------------------------------
// Brute-force canonical labeling (closed form, tiny graphs only):
//   (u,h,d) in L*(v) <=> u <= v, d = d_h(u,v), (h=0 => u=v,d=0),
//   (h>=1 => d_h(u,v) < d_{h-1}(u,v)),
//   and no w < u, h1+h2 <= h with d_{h1}(u,w) + d_{h2}(w,v) <= d.
// This is the fixed point that rank-pruned hop-constrained 2-hop labeling
// converges to; it uses only hop-bounded distances, never the generators.
// ---------------------------------------------------------------------------
LabelSet canonical_bruteforce(const Graph &g, int K) {
  const int32_t n = g.n;
  // D[h][a][b]
  std::vector<std::vector<int32_t>> D((size_t)(K + 1) * n, std::vector<int32_t>(n));
  for (int h = 0; h <= K; ++h)
    for (int32_t a = 0; a < n; ++a) bf(g, a, h, D[(size_t)h * n + a].data(), 0);
  auto d = [&](int h, int32_t a, int32_t b) { return D[(size_t)h * n + a][b]; };
  LabelSet L(n);
  for (int32_t v = 0; v < n; ++v) {
    for (int32_t u = 0; u <= v; ++u) {
      for (int h = 0; h <= K; ++h) {
        if (h == 0) {
          if (u == v) L[v].push_back({v, 0, 0, 0});
          continue;
        }
        if (u == v) continue;
        int32_t dh = d(h, u, v);
        if (dh == INF32) continue;
        if (!(dh < d(h - 1, u, v))) continue;
        bool dominated = false;
        for (int32_t w = 0; w < u && !dominated; ++w) {
          for (int h1 = 0; h1 <= h && !dominated; ++h1) {
            int32_t a = d(h1, u, w);
            if (a == INF32) continue;
            int32_t b = d(h - h1, w, v);
            if (b == INF32) continue;
            if ((int64_t)a + b <= dh) dominated = true;
          }
        }
        if (!dominated) L[v].push_back({u, h, dh, 0});
      }
    }
  }
  sort_all(L);
  return L;
}

// ---------------------------------------------------------------------------
// Text graph loader: graph_v_of_v<int>::txt_read + add_edge,
// CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:92-108,242-277, with
// sorted_vector_binary_operations_insert (tool_functions/
// sorted_vector_binary_operations.h:108-125): "|V|= n" sizes the graph,
// "Edge u v w" inserts both directions into id-sorted adjacency; an existing
// entry has its weight overwritten (last one wins); a self loop is stored
// once.  Weights go through a double and truncate to int, as std::stod does.
// ---------------------------------------------------------------------------
bool load_txt(const char *path, Graph &g) {
  std::ifstream in(path);
  if (!in.is_open()) return false;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> adj;
  std::string line;
  auto insert = [&](int32_t a, int32_t b, int32_t w) {
    auto &v = adj[a];
    auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(b, INT32_MIN));
    if (it != v.end() && it->first == b)
      it->second = w;
    else
      v.insert(it, {b, w});
  };
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok;
    if (!(ss >> tok)) continue;
    if (tok == "|V|=") {
      long long n;
      ss >> n;
      adj.assign((size_t)n, {});
    } else if (tok == "Edge") {
      long long a, b;
      double wd;
      ss >> a >> b >> wd;
      int32_t w = (int32_t)wd;
      insert((int32_t)a, (int32_t)b, w);
      insert((int32_t)b, (int32_t)a, w);
    }
  }
  g.n = (int32_t)adj.size();
  g.off.assign(g.n + 1, 0);
  for (int32_t v = 0; v < g.n; ++v) g.off[v + 1] = g.off[v] + (int64_t)adj[v].size();
  g.col.resize(g.off[g.n]);
  g.w.resize(g.off[g.n]);
  for (int32_t v = 0; v < g.n; ++v) {
    int64_t p = g.off[v];
    for (auto &e : adj[v]) {
      g.col[p] = e.first;
      g.w[p] = e.second;
      ++p;
    }
  }
  return true;
}

// Degree rank: graph_v_of_v_update_vertexIDs_by_degrees_large_to_small,
// CPU_HopHL/include/graph_v_of_v/graph_v_of_v_update_vertexIDs_by_degrees_large_to_small.h:14-40,
// with the tie order fixed as (degree desc, id asc) since the reference's
// std::sort is unstable.
void reorder(const Graph &g, Graph &out, std::vector<int32_t> &new_to_old) {
  const int32_t n = g.n;
  new_to_old.resize(n);
  for (int32_t i = 0; i < n; ++i) new_to_old[i] = i;
  std::stable_sort(new_to_old.begin(), new_to_old.end(),
                   [&](int32_t a, int32_t b) { return g.deg(a) > g.deg(b); });
  std::vector<int32_t> old_to_new(n);
  for (int32_t i = 0; i < n; ++i) old_to_new[new_to_old[i]] = i;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> adj(n);
  for (int32_t v = 0; v < n; ++v)
    for (int64_t e = g.off[v]; e < g.off[v + 1]; ++e)
      adj[old_to_new[v]].push_back({old_to_new[g.col[e]], g.w[e]});
  out.n = n;
  out.off.assign(n + 1, 0);
  for (int32_t v = 0; v < n; ++v) {
    std::sort(adj[v].begin(), adj[v].end());
    out.off[v + 1] = out.off[v] + (int64_t)adj[v].size();
  }
  out.col.resize(out.off[n]);
  out.w.resize(out.off[n]);
  for (int32_t v = 0; v < n; ++v) {
    int64_t p = out.off[v];
    for (auto &e : adj[v]) {
      out.col[p] = e.first;
      out.w[p] = e.second;
      ++p;
    }
  }
}

Graph make_graph(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w) {
  Graph g;
  g.n = n;
  g.off.assign(off, off + n + 1);
  g.col.assign(col, col + off[n]);
  g.w.assign(w, w + off[n]);
  return g;
}

Labels *wrap(LabelSet &&L) {
  Labels *r = new Labels();
  r->L = std::move(L);
  return r;
}

}  // namespace

// ------------------------------- C interface -------------------------------
// ---------------------------------------------------------------------------
// Checks of a finished label set too large to rebuild on the CPU (flat CSR
// arrays, each vertex's labels sorted by (hub, hop, dist)).
//
// Replay of hub s: the K rounds of the K-round twin
// (Hybrid_Generation_Clean/include/HBPLL/hop_constrained_two_hop_labels_generation.h:172-222,
// the GPU round of Hybrid_Generation_Clean/src/label_gen.cu:477-593,87-110)
// recomputed for s alone, with the queries reading the given set: in s's
// round-h query only labels of hubs a < s (final once every smaller hub has
// run) and s's own labels of hops < h can pair (a hop-h label pairs only with
// a self label, which rank pruning rules out), so the set fixes what s's
// round h must produce.  Returns, per sampled hub, 0 when every round's
// accepted (y, dist) equal the set's (s, h, .) labels, else the first
// differing round.
// ---------------------------------------------------------------------------
struct FlatLabels {
  const int64_t *off;
  const uint32_t *hub;
  const uint8_t *hop;
  const int32_t *dist;
};

void replay_hubs(const Graph &g, int K, const FlatLabels &L, int64_t nh, const int32_t *hubs,
                 int threads, int32_t *bad) {
  const int32_t n = g.n;
  // the set's labels of the sampled hubs: per sample, (hop, y, dist)
  std::vector<int32_t> slot(n, -1);
  for (int64_t i = 0; i < nh; ++i) slot[hubs[i]] = (int32_t)i;
  std::vector<std::vector<std::tuple<int32_t, int32_t, int32_t>>> have(nh);
  for (int32_t y = 0; y < n; ++y)
    for (int64_t j = L.off[y]; j < L.off[y + 1]; ++j) {
      const int32_t sl = slot[L.hub[j]];
      if (sl >= 0 && L.hop[j] > 0) have[sl].emplace_back(L.hop[j], y, L.dist[j]);
    }
  for (auto &v : have) std::sort(v.begin(), v.end());
  threads = nthreads_or_default(threads);
  struct Scratch {
    std::vector<int32_t> cand;
    std::vector<int32_t> touched;
    HubTable T;
    explicit Scratch(int32_t n) : cand(n, INF32), T(n) {}
  };
  std::vector<Scratch *> scr;
  for (int t = 0; t < threads; ++t) scr.push_back(new Scratch(n));
  parallel_for(nh, threads, [&](int tid, int64_t i) {
    Scratch &S = *scr[tid];
    const int32_t s = hubs[i];
    for (int64_t j = L.off[s]; j < L.off[s + 1]; ++j) S.T.add((int32_t)L.hub[j], L.dist[j], L.hop[j]);
    std::vector<std::pair<int32_t, int32_t>> F{{s, 0}}, acc;
    size_t pos = 0;  // position in have[i] (sorted by hop, y)
    bad[i] = 0;
    for (int h = 1; h <= K && !bad[i]; ++h) {
      for (const auto &fx : F)
        for (int64_t e = g.off[fx.first]; e < g.off[fx.first + 1]; ++e) {
          const int32_t y = g.col[e];
          if (y <= s) continue;
          const int32_t nd = fx.second + g.w[e];
          if (nd < S.cand[y]) {
            if (S.cand[y] == INF32) S.touched.push_back(y);
            S.cand[y] = nd;
          }
        }
      std::sort(S.touched.begin(), S.touched.end());
      acc.clear();
      for (int32_t y : S.touched) {
        const int32_t dy = S.cand[y];
        S.cand[y] = INF32;
        int64_t q = INF32;
        for (int64_t j = L.off[y]; j < L.off[y + 1]; ++j) {
          const int32_t a = (int32_t)L.hub[j];
          if (a > s || (a == s && L.hop[j] >= h)) break;
          for (const auto &tb : S.T.t[a])
            if (L.hop[j] + tb.second <= h) q = std::min<int64_t>(q, (int64_t)L.dist[j] + tb.first);
        }
        if (dy < q) acc.emplace_back(y, dy);
      }
      S.touched.clear();
      // compare with the set's hop-h labels of s
      size_t e = pos;
      while (e < have[i].size() && std::get<0>(have[i][e]) == h) ++e;
      bool same = (e - pos) == acc.size();
      for (size_t k = 0; same && k < acc.size(); ++k)
        same = std::get<1>(have[i][pos + k]) == acc[k].first &&
               std::get<2>(have[i][pos + k]) == acc[k].second;
      if (!same) bad[i] = h;
      pos = e;
      F.swap(acc);
    }
    if (!bad[i] && pos != have[i].size()) bad[i] = K + 1;  // labels above hop K
    S.T.clear();
  });
  for (auto *x : scr) delete x;
}

// The normative clean (hop_constrained_clean_L, CPU_HopHL/include/HBPLL/
// hop_constrained_two_hop_labels_generation.h:451-482) of sampled vertices
// of an uncleaned set `in`, reading L(u) from the input (lu_mode 1 of
// clean()), compared with the cleaned set `out`: bad[i] = 1 when vertex
// verts[i]'s kept labels differ from out's list.
void clean_check(int32_t n, const FlatLabels &in, const FlatLabels &out, int64_t nv,
                 const int32_t *verts, int threads, int32_t *bad) {
  threads = nthreads_or_default(threads);
  std::vector<HubTable *> tabs;
  for (int t = 0; t < threads; ++t) tabs.push_back(new HubTable(n));
  parallel_for(nv, threads, [&](int tid, int64_t i) {
    HubTable &T = *tabs[tid];
    const int32_t v = verts[i];
    std::vector<std::tuple<uint32_t, uint8_t, int32_t>> kept;
    for (int64_t j = in.off[v]; j < in.off[v + 1]; ++j) {
      const int32_t u = (int32_t)in.hub[j], h = in.hop[j], d = in.dist[j];
      int64_t m = INF32;
      for (int64_t k = in.off[u]; k < in.off[u + 1]; ++k)
        for (const auto &tb : T.t[in.hub[k]])
          if (in.hop[k] + tb.second <= h) m = std::min<int64_t>(m, (int64_t)in.dist[k] + tb.first);
      if (m > d) {
        kept.emplace_back(in.hub[j], in.hop[j], d);
        T.add(u, d, h);
      }
    }
    T.clear();
    bool same = (int64_t)kept.size() == out.off[v + 1] - out.off[v];
    for (size_t k = 0; same && k < kept.size(); ++k) {
      const int64_t j = out.off[v] + (int64_t)k;
      same = std::get<0>(kept[k]) == out.hub[j] && std::get<1>(kept[k]) == out.hop[j] &&
             std::get<2>(kept[k]) == out.dist[j];
    }
    bad[i] = same ? 0 : 1;
  });
  for (auto *t : tabs) delete t;
}

extern "C" {

typedef struct Labels orc_labels;
typedef struct Graph orc_graph;

orc_labels *orc_kround(int32_t n, const int64_t *off, const int32_t *col,
                       const int32_t *w, int K, int threads, int mutation) {
  Graph g = make_graph(n, off, col, w);
  return kround(g, K, threads, mutation);
}

orc_labels *orc_kround_grouped(int32_t n, const int64_t *off, const int32_t *col,
                               const int32_t *w, int K, int group, int threads) {
  Graph g = make_graph(n, off, col, w);
  return kround(g, K, threads, 0, group);
}

orc_labels *orc_kround_gid(int32_t n, const int64_t *off, const int32_t *col,
                           const int32_t *w, int K, const int32_t *gid, int threads) {
  Graph g = make_graph(n, off, col, w);
  return kround(g, K, threads, 0, 0, gid);
}

orc_labels *orc_2023www(int32_t n, const int64_t *off, const int32_t *col,
                        const int32_t *w, int K, int mutation) {
  Graph g = make_graph(n, off, col, w);
  return www2023(g, K, mutation);
}

orc_labels *orc_dijkstra(int32_t n, const int64_t *off, const int32_t *col,
                         const int32_t *w, int K, int mutation) {
  Graph g = make_graph(n, off, col, w);
  return dijkstra(g, K, mutation);
}

orc_labels *orc_canonical_bruteforce(int32_t n, const int64_t *off, const int32_t *col,
                                     const int32_t *w, int K) {
  Graph g = make_graph(n, off, col, w);
  return wrap(canonical_bruteforce(g, K));
}

orc_labels *orc_clean(const orc_labels *in, int lu_mode, int threads, int mutation) {
  return wrap(clean(in->L, lu_mode, threads, mutation));
}

// Build a label set from flat CSR arrays (any per-vertex order; sorted here);
// `parent` may be null (parents 0).
orc_labels *orc_labels_from_arrays_p(int32_t n, const int64_t *off, const uint32_t *hub,
                                     const uint8_t *hop, const int32_t *dist,
                                     const int32_t *parent) {
  LabelSet L(n);
  for (int32_t v = 0; v < n; ++v)
    for (int64_t i = off[v]; i < off[v + 1]; ++i)
      L[v].push_back({(int32_t)hub[i], (int32_t)hop[i], dist[i], parent ? parent[i] : 0});
  sort_all(L);
  return wrap(std::move(L));
}

orc_labels *orc_labels_from_arrays(int32_t n, const int64_t *off, const uint32_t *hub,
                                   const uint8_t *hop, const int32_t *dist) {
  return orc_labels_from_arrays_p(n, off, hub, hop, dist, nullptr);
}

int64_t orc_labels_total(const orc_labels *l) {
  int64_t t = 0;
  for (auto &v : l->L) t += (int64_t)v.size();
  return t;
}

int32_t orc_labels_n(const orc_labels *l) { return (int32_t)l->L.size(); }

// counters: [frontier_items, scanned_edges, candidates, cand_label_reads,
//            hub_label_reads, accepted, rounds_run]
void orc_labels_counters(const orc_labels *l, int64_t *out7) {
  out7[0] = l->frontier_items;
  out7[1] = l->scanned_edges;
  out7[2] = l->candidates;
  out7[3] = l->cand_label_reads;
  out7[4] = l->hub_label_reads;
  out7[5] = l->accepted;
  out7[6] = l->rounds_run;
}

void orc_labels_export(const orc_labels *l, int64_t *off, uint32_t *hub, uint8_t *hop,
                       int32_t *dist) {
  int64_t p = 0;
  off[0] = 0;
  for (size_t v = 0; v < l->L.size(); ++v) {
    for (const Label &x : l->L[v]) {
      hub[p] = (uint32_t)x.hub;
      hop[p] = (uint8_t)x.hop;
      dist[p] = x.dist;
      ++p;
    }
    off[v + 1] = p;
  }
}

// parent vertex of every label, in the export's order
void orc_labels_parents(const orc_labels *l, int32_t *par) {
  int64_t p = 0;
  for (const auto &v : l->L)
    for (const Label &x : v) par[p++] = x.parent;
}

void orc_labels_free(orc_labels *l) { delete l; }

// Paths for nq (s, t, k): counts[i] edges of query i (at most cap are
// stored, the count is the full length) at edges[2 * cap * i ...] as
// (vertex, parent) pairs in the order the reference emits them.
void orc_paths(const orc_labels *l, int64_t nq, const uint32_t *s, const uint32_t *t,
               const int32_t *k, int32_t cap, int32_t *edges, int32_t *counts, int threads,
               int mutation) {
  parallel_for(nq, threads, [&](int, int64_t i) {
    std::vector<std::pair<int32_t, int32_t>> out;
    path_one(l->L, (int32_t)s[i], (int32_t)t[i], k[i], out, mutation);
    counts[i] = (int32_t)out.size();
    for (int32_t e = 0; e < (int32_t)out.size() && e < cap; ++e) {
      edges[2 * (cap * i + e)] = out[e].first;
      edges[2 * (cap * i + e) + 1] = out[e].second;
    }
  });
}

void orc_query(const orc_labels *l, int64_t nq, const uint32_t *s, const uint32_t *t,
               const int32_t *k, int32_t *out, int threads, int mutation) {
  parallel_for(nq, threads, [&](int, int64_t i) {
    out[i] = query_one(l->L, (int32_t)s[i], (int32_t)t[i], k[i], mutation);
  });
}

void orc_bf(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w,
            int32_t src, int k, int32_t *dist, int mutation) {
  Graph g = make_graph(n, off, col, w);
  bf(g, src, k, dist, mutation);
}

// Text loader + optional degree reorder.  Returns a graph handle.
orc_graph *orc_load_txt(const char *path, int do_reorder) {
  Graph *g = new Graph();
  if (!load_txt(path, *g)) {
    delete g;
    return nullptr;
  }
  if (do_reorder) {
    Graph r;
    std::vector<int32_t> n2o;
    reorder(*g, r, n2o);
    *g = std::move(r);
  }
  return g;
}

orc_graph *orc_reorder(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w,
                       int32_t *new_to_old) {
  Graph g = make_graph(n, off, col, w);
  Graph *r = new Graph();
  std::vector<int32_t> n2o;
  reorder(g, *r, n2o);
  if (new_to_old) std::memcpy(new_to_old, n2o.data(), sizeof(int32_t) * n);
  return r;
}

int32_t orc_graph_n(const orc_graph *g) { return g->n; }
int64_t orc_graph_nnz(const orc_graph *g) { return g->off[g->n]; }
void orc_graph_export(const orc_graph *g, int64_t *off, int32_t *col, int32_t *w) {
  std::memcpy(off, g->off.data(), sizeof(int64_t) * (g->n + 1));
  std::memcpy(col, g->col.data(), sizeof(int32_t) * g->off[g->n]);
  std::memcpy(w, g->w.data(), sizeof(int32_t) * g->off[g->n]);
}
void orc_graph_free(orc_graph *g) { delete g; }

// Per block of `block` vertices: label count and the wrapping sum of
// mix(mix(mix(v << 32 | i) ^ (hub | hop << 32)) + dist) over its labels
// (i = position in L(v), mix = the splitmix64 finaliser); the same digest
// as tests/label_digest.py, compiled for label sets of billions.
static inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void orc_block_digests(int32_t n, const int64_t *off, const uint32_t *hub, const uint8_t *hop,
                       const int32_t *dist, int32_t block, int threads, uint64_t *counts,
                       uint64_t *dig) {
  const int64_t nb = ((int64_t)n + block - 1) / block;
  parallel_for(nb, threads, [&](int, int64_t b) {
    const int64_t v0 = b * block, v1 = std::min<int64_t>(n, v0 + block);
    uint64_t acc = 0;
    for (int64_t v = v0; v < v1; ++v)
      for (int64_t j = off[v]; j < off[v + 1]; ++j) {
        const uint64_t a = mix64(((uint64_t)v << 32) | (uint64_t)(j - off[v]));
        const uint64_t key = (uint64_t)hub[j] | ((uint64_t)hop[j] << 32);
        acc += mix64(mix64(a ^ key) + (uint64_t)(uint32_t)dist[j]);
      }
    counts[b] = (uint64_t)(off[v1] - off[v0]);
    dig[b] = acc;
  });
}

void orc_dijkstra_timed(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w,
                        int K, int threads, double seconds, int64_t *hubs_done, int64_t *labels) {
  Graph g = make_graph(n, off, col, w);
  dijkstra_timed(g, K, threads, seconds, hubs_done, labels);
}

void orc_2023www_timed(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w,
                       int K, int threads, double seconds, int64_t *hubs_done, int64_t *labels) {
  Graph g = make_graph(n, off, col, w);
  www2023_timed(g, K, threads, seconds, hubs_done, labels);
}

void orc_replay_hubs(int32_t n, const int64_t *off, const int32_t *col, const int32_t *w, int K,
                     const int64_t *loff, const uint32_t *lhub, const uint8_t *lhop,
                     const int32_t *ldist, int64_t nh, const int32_t *hubs, int threads,
                     int32_t *bad) {
  Graph g = make_graph(n, off, col, w);
  replay_hubs(g, K, FlatLabels{loff, lhub, lhop, ldist}, nh, hubs, threads, bad);
}

void orc_clean_check(int32_t n, const int64_t *ioff, const uint32_t *ihub, const uint8_t *ihop,
                     const int32_t *idist, const int64_t *ooff, const uint32_t *ohub,
                     const uint8_t *ohop, const int32_t *odist, int64_t nv, const int32_t *verts,
                     int threads, int32_t *bad) {
  clean_check(n, FlatLabels{ioff, ihub, ihop, idist}, FlatLabels{ooff, ohub, ohop, odist}, nv,
              verts, threads, bad);
}

}  // extern "C"
