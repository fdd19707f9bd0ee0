// compat_driver.cpp -- a reference-style driver on include/hsdl_compat.hpp.
//
// Follows the flow of the reference's CPU_HopHL test driver
// (CPU_HopHL/include/HBPLL/test.h:136-227): read a graph in the |V|=/Edge
// text format, rename vertices by degree, generate + canonically repair the
// labels, then check sampled (s,t,k) answers against a hop-bounded
// Bellman-Ford written here (CPU_HopHL/include/graph_v_of_v/
// graph_v_of_v_hop_constrained_shortest_distance.h:7-38 computes the same).
// Prints "check pass" and exits 0 when every answer matches.
//
// usage: compat_driver <graph.txt> <K> <pairs>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "hsdl_compat.hpp"

using namespace hsdl_compat;

static std::vector<int> hop_bounded(graph_v_of_v<int> &g, int src, int k) {
  std::vector<int> cur(g.size(), INT_MAX), nxt;
  cur[src] = 0;
  for (int r = 0; r < k; ++r) {
    nxt = cur;
    for (int x = 0; x < g.size(); ++x) {
      if (cur[x] == INT_MAX) continue;
      for (auto &e : g[x]) {
        long long d = (long long)cur[x] + e.second;
        if (d < nxt[e.first]) nxt[e.first] = (int)d;
      }
    }
    cur.swap(nxt);
  }
  return cur;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s graph.txt K pairs\n", argv[0]);
    return 2;
  }
  const int K = atoi(argv[2]);
  const int pairs = atoi(argv[3]);
  try {
    graph_v_of_v<int> g0;
    g0.txt_read(argv[1]);
    graph_v_of_v<int> g = graph_v_of_v_update_vertexIDs_by_degrees_large_to_small(g0);
    hop_constrained_case_info mm;
    mm.upper_k = K;
    mm.use_canonical_repair = true;
    hop_constrained_two_hop_labels_generation(g, mm);
    printf("labels %.0f -> %.0f, generate %.4f s, clean %.4f s\n",
           mm.label_size_before_canonical_repair, mm.label_size_after_canonical_repair,
           mm.time_generate_labels, mm.time_canonical_repair);
    std::mt19937 rng(12345);
    std::uniform_int_distribution<int> pick(0, g.size() - 1), hop(1, K);
    int bad = 0;
    std::vector<query_info> qs;
    std::vector<int> want;
    for (int i = 0; i < pairs; ++i) {
      const int s = pick(rng), t = pick(rng), k = hop(rng);
      const std::vector<int> d = hop_bounded(g, s, k);
      const int got = hop_constrained_extract_distance(mm, s, t, k);
      if (got != d[t]) {
        printf("mismatch s=%d t=%d k=%d: labels %d, Bellman-Ford %d\n", s, t, k, got, d[t]);
        ++bad;
      }
      qs.push_back({s, t, k});
      want.push_back(d[t]);
    }
    std::vector<int> ans(qs.size());
    gpu_query(mm, (int)qs.size(), qs.data(), ans.data());
    for (size_t i = 0; i < qs.size(); ++i)
      if (ans[i] != want[i]) ++bad;
    // the host-label overload answers the same
    if (!qs.empty() &&
        hop_constrained_extract_distance(mm.L, qs[0].s, qs[0].t, qs[0].h) != want[0])
      ++bad;
    // shortest paths: a walk from s to t of at most k edges whose weight is
    // the distance (the edges extend the s side or the t side until they meet)
    int paths = 0;
    for (size_t i = 0; i < qs.size(); ++i) {
      const auto p = hop_constrained_extract_shortest_path(mm, qs[i].s, qs[i].t, qs[i].h);
      if (want[i] == INT_MAX || qs[i].s == qs[i].t) {
        if (!p.empty()) ++bad;
        continue;
      }
      int cs = qs[i].s, ct = qs[i].t;
      long long w = 0;
      bool ok = (int)p.size() <= qs[i].h;
      for (auto &e : p) {
        bool found = false;
        for (auto &nb : g[e.first])
          if (nb.first == e.second) {
            w += nb.second;
            found = true;
          }
        if (!found) ok = false;
        if (e.first == cs) cs = e.second;
        else if (e.first == ct) ct = e.second;
        else ok = false;
      }
      if (!ok || cs != ct || w != want[i]) {
        printf("path mismatch s=%d t=%d k=%d\n", qs[i].s, qs[i].t, qs[i].h);
        ++bad;
      }
      ++paths;
      // the host-label overload (parents from mm.L) gives the same edges
      if (i < 5 && hop_constrained_extract_shortest_path(mm.L, qs[i].s, qs[i].t, qs[i].h) != p)
        ++bad;
    }
    printf("%d paths checked\n", paths);
    // canonical repair of already-canonical labels changes nothing
    const double before = mm.label_size_after_canonical_repair;
    hop_constrained_clean_L(mm, 1);
    if (mm.label_size_after_canonical_repair != before) ++bad;
    if (bad) {
      printf("check failed: %d mismatches\n", bad);
      return 1;
    }
    printf("check pass\n");
    return 0;
  } catch (const error &e) {
    printf("hsdl error: %s\n", e.what());
    return 3;
  }
}
