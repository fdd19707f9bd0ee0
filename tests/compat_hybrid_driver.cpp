// compat_hybrid_driver.cpp -- the reference's hybrid GPU driver flow on
// include/hsdl_compat.hpp.
//
// Follows Hybrid_Generation_Clean/src/test.cu:219-478 with every core a GPU
// core: read a |V|=/Edge graph, rename vertices by degree, convert it to LDBC
// and CSR (toCSR), cut the vertices into groups of G_max consecutive ids
// (use_cd = 0, :310-320), or take CDLP groups (use_cd = 1, :321-325) or
// shuffled ones, set up hop_constrained_case_info_v2 (init,
// set_nid), hand the groups to the least-loaded of four GPU cores
// (Executive_Core queue, :361-383) that each run label_gen, release the device
// state (destroy_L_cuda, destroy_csr_graph), sort every L(v), then clean per
// group (gpu_clean_init + gpu_clean, :410-447).  Checks, written here:
//   - the cleaned labels answer sampled (s, t, k) like a hop-bounded
//     Bellman-Ford (the reference's GPU_HSDL_checker, :83-149), through the
//     host-label query and through the query module's gpu_query
//     (Query_mdnd/src/gpu_query.cu:84);
//   - Clean_pc's gpu_clean (Clean_pc/src/gpu_clean.cu:140) on the same
//     uncleaned labels gives the same lists as the per-group clean;
//   - the uncleaned-output generation (Clean_pc/include/HBPLL/
//     hop_constrained_two_hop_labels_generation.h:514) gives the same cleaned
//     lists.
// Writes the cleaned labels in the reference's dump layout
// (Clean_pc/main.cu:11-57) to <out> and prints the label totals and
// "check pass".
//
// usage: compat_hybrid_driver <graph.txt> <K> <G_max> <pairs> <out.bin>
//                             [contiguous|cdlp|random]
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <string>
#include <vector>

#include "hsdl_compat.hpp"

using namespace hsdl_compat;

struct Executive_Core {
  int id;
  double time_use;
  int core_type;  // 0: cpu, 1: gpu
  Executive_Core(int x, double y, int z) : id(x), time_use(y), core_type(z) {}
};
inline bool operator<(Executive_Core a, Executive_Core b) {
  if (a.time_use == b.time_use) return a.id > b.id;
  return a.time_use > b.time_use;
}

static bool label_before(const hop_constrained_two_hop_label &i,
                         const hop_constrained_two_hop_label &j) {
  if (i.hub_vertex != j.hub_vertex) return i.hub_vertex < j.hub_vertex;
  if (i.hop != j.hop) return i.hop < j.hop;
  return i.distance < j.distance;
}

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

static bool same_lists(const std::vector<std::vector<hop_constrained_two_hop_label>> &a,
                       const std::vector<std::vector<hop_constrained_two_hop_label>> &b) {
  if (a.size() != b.size()) return false;
  for (size_t v = 0; v < a.size(); ++v) {
    if (a[v].size() != b[v].size()) return false;
    for (size_t i = 0; i < a[v].size(); ++i)
      if (a[v][i].hub_vertex != b[v][i].hub_vertex || a[v][i].hop != b[v][i].hop ||
          a[v][i].distance != b[v][i].distance)
        return false;
  }
  return true;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s graph.txt K G_max pairs out.bin [contiguous|cdlp|random]\n",
            argv[0]);
    return 2;
  }
  const std::string grouping = argc > 6 ? argv[6] : "contiguous";
  const int hop_cst = atoi(argv[2]);
  int G_max = atoi(argv[3]);
  const int pairs = atoi(argv[4]);
  const int thread_num = 1000, thread_num_clean = 5000, GPU_Gen_Num = 4, GPU_Clean_Num = 4;
  try {
    graph_v_of_v<int> g0;
    g0.txt_read(argv[1]);
    graph_v_of_v<int> instance_graph = graph_v_of_v_update_vertexIDs_by_degrees_large_to_small(g0);
    const int V = instance_graph.size();
    int Distributed_Graph_Num = (V + G_max - 1) / G_max;

    LDBC<int> graph(V);
    graph_v_of_v_to_LDBC(graph, instance_graph);
    CSR_graph<int> csr_graph = toCSR(graph);

    // groups as test.cu:309-326 makes them (use_cd 0: contiguous; 1: CDLP),
    // or shuffled (Random_group.cuh)
    Graph_pool<int> graph_pool;
    if (grouping == "cdlp") {
      generate_Group_CDLP(instance_graph, graph_pool.graph_group, G_max);
      Distributed_Graph_Num = (int)graph_pool.graph_group.size();
    } else if (grouping == "random") {
      generate_Group_Random(instance_graph, hop_cst, graph_pool.graph_group, G_max, 7);
      Distributed_Graph_Num = (int)graph_pool.graph_group.size();
    } else {
      graph_pool.graph_group.resize(Distributed_Graph_Num);
      const int Nodes_Per_Graph = (V - 1) / Distributed_Graph_Num + 1;
      for (int i = 0; i < Distributed_Graph_Num; ++i)
        for (int j = Nodes_Per_Graph * i; j < Nodes_Per_Graph * (i + 1) && j < V; ++j)
          graph_pool.graph_group[i].push_back(j);
      G_max = V / Distributed_Graph_Num + 1;
    }
    graph_pool.graph_group_bfs.resize(Distributed_Graph_Num);

    hop_constrained_case_info_v2 *info_gpu = new hop_constrained_case_info_v2();
    info_gpu->init(V, hop_cst, G_max, thread_num, graph_pool.graph_group);
    info_gpu->set_nid(Distributed_Graph_Num, graph_pool.graph_group);

    std::vector<std::vector<hop_constrained_two_hop_label>> L_hybrid(V);
    std::priority_queue<Executive_Core> pq_gen;
    for (int i = 0; i < GPU_Gen_Num; ++i) pq_gen.push(Executive_Core(i, 0, 1));
    for (int i = 0; i < Distributed_Graph_Num; ++i) {
      Executive_Core x = pq_gen.top();
      pq_gen.pop();
      const double t0 = info_gpu->time_generate_labels + info_gpu->time_traverse_labels;
      label_gen(csr_graph, info_gpu, L_hybrid, graph_pool.graph_group[i], i);
      x.time_use += info_gpu->time_generate_labels + info_gpu->time_traverse_labels - t0;
      pq_gen.push(x);
    }
    info_gpu->destroy_L_cuda();
    csr_graph.destroy_csr_graph();
    for (int v = 0; v < V; ++v) std::sort(L_hybrid[v].begin(), L_hybrid[v].end(), label_before);
    long long before = 0;
    for (auto &l : L_hybrid) before += (long long)l.size();
    const auto uncleaned = L_hybrid;

    std::priority_queue<Executive_Core> pq_clean;
    for (int i = 0; i < GPU_Clean_Num; ++i) pq_clean.push(Executive_Core(i, 0, 1));
    gpu_clean_init(instance_graph, L_hybrid, info_gpu, graph_pool, thread_num, hop_cst);
    for (int i = 0; i < Distributed_Graph_Num; ++i) {
      Executive_Core x = pq_clean.top();
      pq_clean.pop();
      gpu_clean(instance_graph, info_gpu, L_hybrid, thread_num_clean, i);
      pq_clean.push(x);
    }
    long long after = 0;
    for (auto &l : L_hybrid) after += (long long)l.size();
    printf("groups %d, labels before clean %lld, after clean %lld\n", Distributed_Graph_Num,
           before, after);

    int bad = 0;
    // Clean_pc's whole-set GPU clean gives the per-group clean's lists
    std::vector<std::vector<hop_constrained_two_hop_label>> res;
    auto uc = uncleaned;
    gpu_clean(instance_graph, uc, res, thread_num, hop_cst);
    if (!same_lists(res, L_hybrid)) {
      printf("Clean_pc gpu_clean differs from the per-group clean\n");
      ++bad;
    }
    // the uncleaned-output generation cleans to the same lists (a set
    // generated without groups is already canonical)
    hop_constrained_case_info mm;
    mm.upper_k = hop_cst;
    mm.compute_parents = false;
    std::vector<std::vector<hop_constrained_two_hop_label>> unc2;
    hop_constrained_two_hop_labels_generation(instance_graph, mm, unc2);
    if (!same_lists(mm.L, L_hybrid)) {
      printf("hop_constrained_two_hop_labels_generation(.., uncleaned_L) cleans differently\n");
      ++bad;
    }
    // the reference checker: sampled (s, t, k) against Bellman-Ford
    std::mt19937 rng(4321);
    std::uniform_int_distribution<int> pick(0, V - 1), hop(1, hop_cst);
    std::vector<query_info> qs;
    std::vector<int> want;
    for (int i = 0; i < pairs; ++i) {
      const int s = pick(rng), t = pick(rng), k = hop(rng);
      const int d = hop_bounded(instance_graph, s, k)[t];
      if (hop_constrained_extract_distance(L_hybrid, s, t, k) != d) ++bad;
      qs.push_back({s, t, k});
      want.push_back(d);
    }
    std::vector<std::vector<label>> Lq(V);
    for (int v = 0; v < V; ++v)
      for (auto &x : L_hybrid[v]) Lq[v].push_back({x.hub_vertex, x.hop, x.distance});
    std::vector<int> ans(qs.size());
    gpu_query(instance_graph, Lq, (int)qs.size(), qs.data(), ans.data(), hop_cst);
    for (size_t i = 0; i < qs.size(); ++i)
      if (ans[i] != want[i]) {
        printf("gpu_query mismatch s=%d t=%d k=%d: %d vs %d\n", qs[i].s, qs[i].t, qs[i].h, ans[i],
               want[i]);
        ++bad;
      }
    // the reference's label dump layout (Clean_pc/main.cu:11-57)
    FILE *f = fopen(argv[5], "wb");
    if (!f) {
      printf("cannot write %s\n", argv[5]);
      return 1;
    }
    size_t n = (size_t)V;
    fwrite(&n, sizeof(size_t), 1, f);
    for (int v = 0; v < V; ++v) {
      size_t c = L_hybrid[v].size();
      fwrite(&c, sizeof(size_t), 1, f);
      for (auto &x : L_hybrid[v]) {
        const int rec[4] = {x.hub_vertex, x.parent_vertex, x.hop, x.distance};
        fwrite(rec, sizeof(int), 4, f);
      }
    }
    fclose(f);
    delete info_gpu;
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
