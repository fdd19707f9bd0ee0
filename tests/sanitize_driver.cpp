// sanitize_driver.cpp -- the library's host-side input paths under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitizers.py, run on the CPU without a GPU).
//
// Exercises: graph construction from random edge lists (duplicates, self
// loops, both flag bits), degree renaming, CSR export, text save/load, the
// .e edge-list loader and the seeded generators; malformed graph files;
// label import from host arrays and label dumps, corrupt in every way the
// host checks (out-of-range hubs, hops above K, bad offsets; truncated
// files, counts larger than the file, negative fields).  Inputs that pass
// every check go on to the device, whose runtime start-up under the
// sanitizers takes most of a minute on a machine without a GPU; the CPU
// tests cover those (HSDL_ENODEV).  Every call must return its documented
// code; the sanitizers abort on any memory or UB error.  Prints
// "sanitize pass".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hsdl.h"

static int failures = 0;
#define EXPECT(cond)                                                     \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::printf("expectation failed at line %d: %s\n", __LINE__, #cond); \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

static void write_file(const std::string &path, const std::string &bytes) {
  FILE *f = std::fopen(path.c_str(), "wb");
  std::fwrite(bytes.data(), 1, bytes.size(), f);
  std::fclose(f);
}

template <class T>
static void put(std::string &b, T v) {
  b.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  std::mt19937_64 rng(2026);
  // graphs from random edge lists
  for (int trial = 0; trial < 60; ++trial) {
    const uint32_t n = 1 + rng() % 300;
    const uint64_t m = rng() % 2000;
    std::vector<uint32_t> u(m), v(m);
    std::vector<int32_t> w(m);
    for (uint64_t i = 0; i < m; ++i) {
      u[i] = (uint32_t)(rng() % n);
      v[i] = (uint32_t)(rng() % n);
      w[i] = 1 + (int32_t)(rng() % 1000);
    }
    hsdl_graph *g = nullptr;
    EXPECT(hsdl_graph_from_edges(n, m, u.data(), v.data(), w.data(), (uint32_t)(trial % 4), &g) ==
           HSDL_OK);
    uint32_t gn = 0;
    uint64_t nnz = 0;
    EXPECT(hsdl_graph_info(g, &gn, &nnz) == HSDL_OK && gn == n);
    std::vector<uint32_t> n2o(n);
    EXPECT(hsdl_graph_reorder_by_degree(g, n2o.data()) == HSDL_OK);
    std::vector<uint64_t> off(n + 1);
    std::vector<uint32_t> col(nnz);
    std::vector<int32_t> ww(nnz);
    EXPECT(hsdl_graph_csr(g, off.data(), col.data(), ww.data()) == HSDL_OK);
    const std::string p = dir + "/g.txt";
    EXPECT(hsdl_graph_save_txt(g, p.c_str()) == HSDL_OK);
    hsdl_graph *g2 = nullptr;
    EXPECT(hsdl_graph_load_txt(p.c_str(), &g2) == HSDL_OK);
    hsdl_free_graph(g2);
    hsdl_free_graph(g);
  }
  // bad edges: a vertex out of range, a weight below 1
  {
    uint32_t u[2] = {0, 5}, v[2] = {1, 1};
    int32_t w[2] = {1, 1};
    hsdl_graph *g = nullptr;
    EXPECT(hsdl_graph_from_edges(3, 2, u, v, w, 0, &g) == HSDL_EINVAL);
    w[1] = 0;
    u[1] = 2;
    EXPECT(hsdl_graph_from_edges(3, 2, u, v, w, 0, &g) == HSDL_EINVAL);
  }
  // generators
  for (uint32_t scale = 1; scale <= 11; scale += 2) {
    hsdl_graph *g = nullptr;
    EXPECT(hsdl_graph_rmat(scale, 8, 0.57, 0.19, 0.19, scale, 1, 100, &g) == HSDL_OK);
    hsdl_free_graph(g);
  }
  {
    hsdl_graph *g = nullptr;
    EXPECT(hsdl_graph_grid(17, 23, 5, 1, 1000, &g) == HSDL_OK);
    hsdl_free_graph(g);
    EXPECT(hsdl_graph_grid(0, 23, 5, 1, 1000, &g) != HSDL_OK);
  }
  // malformed graph files: every one is refused or loaded, never a crash
  const char *texts[] = {"", "|V|= 3\n", "|V|= -4\nEdge 0 1 1\n", "|V|= 3\nEdge 0 9 1\n",
                         "|V|= 3\nEdge 0 1\n", "garbage\n\n\n", "|V|= 99999999999999\n",
                         "|V|= 3\nEdge 0 1 -5\n", "|V|= 2\nEdge 1 0 2.9\nEdge 0 1 7\n",
                         "|V|= 3\nEdge 0 1 1e30\n", "x |V|= 3\n0 1 1e30\n", "x |V|= 3\n0 1 nan\n",
                         "x |V|= -4\n0 1 1\n", "x |V|= 3\n0 1 2\n2 1 9\n1 0 -3\n"};
  for (const char *t : texts) {
    const std::string p = dir + "/bad.txt";
    write_file(p, t);
    hsdl_graph *g = nullptr;
    const int rc = hsdl_graph_load_txt(p.c_str(), &g);
    EXPECT(rc == HSDL_OK || rc == HSDL_EIO || rc == HSDL_EINVAL || rc == HSDL_ENOMEM);
    if (rc == HSDL_OK) hsdl_free_graph(g);
    hsdl_graph *g2 = nullptr;
    const int rc2 = hsdl_graph_load_edgelist(p.c_str(), 0, &g2);
    EXPECT(rc2 == HSDL_OK || rc2 == HSDL_EIO || rc2 == HSDL_EINVAL || rc2 == HSDL_ENOMEM);
    if (rc2 == HSDL_OK) hsdl_free_graph(g2);
  }
  // label import from host arrays: invalid ones are refused before the
  // device
  {
    uint64_t off[3] = {0, 1, 3};
    uint32_t hub[3] = {0, 7, 1};
    uint8_t hop[3] = {0, 1, 0};
    int32_t dist[3] = {0, 4, 0};
    hsdl_index *ix = nullptr;
    EXPECT(hsdl_index_from_labels(2, 3, off, hub, hop, dist, -1, &ix) == HSDL_EINVAL);
    hub[1] = 0;
    hop[1] = 9;
    EXPECT(hsdl_index_from_labels(2, 3, off, hub, hop, dist, -1, &ix) == HSDL_ERANGE);
    hop[1] = 1;
    uint64_t boff[3] = {0, 3, 1};
    EXPECT(hsdl_index_from_labels(2, 3, boff, hub, hop, dist, -1, &ix) == HSDL_EINVAL);
  }
  // label dumps (Clean_pc/main.cu:11-57 layout)
  auto dump = [&](const std::vector<std::vector<int32_t>> &rows) {
    std::string b;
    put<uint64_t>(b, rows.size());
    for (auto &r : rows) {
      put<uint64_t>(b, r.size() / 4);
      for (int32_t x : r) put<int32_t>(b, x);
    }
    return b;
  };
  const std::string good = dump({{0, 0, 0, 0}, {0, 0, 1, 7, 1, 0, 0, 0}});
  struct Case {
    std::string bytes;
    int want;
  };
  std::vector<Case> cases;
  cases.push_back({dump({{0, 0, 0, 0}, {0, 0, 33, 7, 1, 0, 0, 0}}), HSDL_ERANGE});
  cases.push_back({dump({{0, 0, 0, 0}, {-1, 0, 1, 7, 1, 0, 0, 0}}), HSDL_EIO});
  cases.push_back({dump({{0, 0, 0, 0}, {0, 0, 1, -7, 1, 0, 0, 0}}), HSDL_EIO});
  cases.push_back({good.substr(0, good.size() - 3), HSDL_EIO});
  cases.push_back({std::string(), HSDL_EIO});
  {
    std::string b = good;
    const uint64_t huge = 1ull << 39;
    std::memcpy(&b[8], &huge, 8);
    cases.push_back({b, HSDL_EIO});
    const uint64_t many = 1ull << 31;
    std::string c = good;
    std::memcpy(&c[0], &many, 8);
    cases.push_back({c, HSDL_EIO});
  }
  for (size_t i = 0; i < cases.size(); ++i) {
    const std::string p = dir + "/dump" + std::to_string(i) + ".bin";
    write_file(p, cases[i].bytes);
    hsdl_index *ix = nullptr;
    const int rc = hsdl_load_labels(p.c_str(), 5, -1, &ix);
    if (rc != cases[i].want) std::printf("dump case %zu: got %d want %d\n", i, rc, cases[i].want);
    EXPECT(rc == cases[i].want);
  }
  if (failures) {
    std::printf("sanitize failed: %d\n", failures);
    return 1;
  }
  std::printf("sanitize pass\n");
  return 0;
}
