// graph.cpp -- host-side graph construction for the HSDL engine: edge list ->
// symmetric CSR with the reference's add_edge rules, degree-rank renaming,
// text / edge-list loaders and seeded synthetic generators.
//
// Reference semantics followed:
//   add_edge / txt_read / txt_save:
//     CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:92-108,215-277
//   degree order:
//     CPU_HopHL/include/graph_v_of_v/graph_v_of_v_update_vertexIDs_by_degrees_large_to_small.h:14-40
//   ".e" edge lists: data/ca-CondMat.txt/makedata.cpp:1-36 (writer),
//     GPU4HSDL_EXP/CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:242-295 (reader)
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <numeric>
#include <string>
#include <thread>

#include "common.hpp"

namespace hsdl {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
const char *last_error_cstr() { return g_last_error.c_str(); }

uint64_t now_ns() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

// Rows hold [neighbours < v] [v itself if a self loop] [neighbours > v]; the
// canonical edge keys arrive sorted by (a,b), so each part fills in id order.
void csr_from_unique_edges(uint32_t n, const std::vector<uint64_t> &keys,
                           const std::vector<int32_t> &w, hsdl_graph &g) {
  g.n = n;
  std::vector<uint64_t> deg(n, 0), lower(n, 0);
  std::vector<uint8_t> self(n, 0);
  for (uint64_t k : keys) {
    uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    if (a == b) {
      deg[a] += 1;
      self[a] = 1;
    } else {
      deg[a] += 1;
      deg[b] += 1;
      lower[b] += 1;
    }
  }
  g.off.assign((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) g.off[v + 1] = g.off[v] + deg[v];
  g.col.assign(g.off[n], 0);
  g.w.assign(g.off[n], 0);
  std::vector<uint64_t> plo(n), phi(n);
  for (uint32_t v = 0; v < n; ++v) {
    plo[v] = g.off[v];
    phi[v] = g.off[v] + lower[v] + self[v];
  }
  int32_t mw = 0;
  for (size_t i = 0; i < keys.size(); ++i) {
    uint32_t a = (uint32_t)(keys[i] >> 32), b = (uint32_t)keys[i];
    int32_t ww = w[i];
    mw = std::max(mw, ww);
    if (a == b) {
      uint64_t p = g.off[a] + lower[a];
      g.col[p] = a;
      g.w[p] = ww;
    } else {
      g.col[phi[a]] = b;
      g.w[phi[a]++] = ww;
      g.col[plo[b]] = a;
      g.w[plo[b]++] = ww;
    }
  }
  g.max_w = mw;
}

static void sort_u64_pairs(std::vector<uint64_t> &keys, std::vector<int32_t> &vals) {
  std::vector<uint32_t> idx(keys.size());
  std::iota(idx.begin(), idx.end(), 0u);
  std::sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
  std::vector<uint64_t> k2(keys.size());
  std::vector<int32_t> v2(keys.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    k2[i] = keys[idx[i]];
    v2[i] = vals[idx[i]];
  }
  keys.swap(k2);
  vals.swap(v2);
}

void reorder_by_degree(hsdl_graph &g, std::vector<uint32_t> &new_to_old) {
  const uint32_t n = g.n;
  // stable counting sort by degree, descending
  uint64_t maxdeg = 0;
  for (uint32_t v = 0; v < n; ++v) maxdeg = std::max(maxdeg, g.off[v + 1] - g.off[v]);
  std::vector<uint64_t> cnt(maxdeg + 2, 0);
  for (uint32_t v = 0; v < n; ++v) cnt[maxdeg - (g.off[v + 1] - g.off[v]) + 1]++;
  for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
  new_to_old.assign(n, 0);
  for (uint32_t v = 0; v < n; ++v) new_to_old[cnt[maxdeg - (g.off[v + 1] - g.off[v])]++] = v;
  std::vector<uint32_t> old_to_new(n);
  for (uint32_t i = 0; i < n; ++i) old_to_new[new_to_old[i]] = i;
  std::vector<uint64_t> keys;
  std::vector<int32_t> ws;
  keys.reserve(g.off[n] / 2 + 16);
  ws.reserve(g.off[n] / 2 + 16);
  for (uint32_t v = 0; v < n; ++v) {
    for (uint64_t e = g.off[v]; e < g.off[v + 1]; ++e) {
      uint32_t u = g.col[e];
      if (u < v) continue;  // each undirected edge once (self loops too)
      uint32_t a = old_to_new[v], b = old_to_new[u];
      if (a > b) std::swap(a, b);
      keys.push_back(((uint64_t)a << 32) | b);
      ws.push_back(g.w[e]);
    }
  }
  sort_u64_pairs(keys, ws);
  csr_from_unique_edges(n, keys, ws, g);
}

// Edge list -> unique canonical edges with add_edge's update rule.
static int build_from_edge_list(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                                const int32_t *w, uint32_t flags, hsdl_graph &g) {
  std::vector<uint64_t> keys;
  std::vector<int32_t> ws;
  keys.reserve(m);
  ws.reserve(m);
  std::vector<uint32_t> order;
  order.reserve(m);
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = u[i], b = v[i];
    if (a >= n || b >= n) return fail(HSDL_EINVAL, "edge endpoint out of range");
    if (w[i] < 1) return fail(HSDL_EINVAL, "edge weights must be >= 1");
    if (a == b && (flags & HSDL_DROP_SELF_LOOPS)) continue;
    if (a > b) std::swap(a, b);
    keys.push_back(((uint64_t)a << 32) | b);
    ws.push_back(w[i]);
  }
  // stable sort by key keeps input order among duplicates
  std::vector<uint32_t> idx(keys.size());
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
  std::vector<uint64_t> uk;
  std::vector<int32_t> uw;
  uk.reserve(idx.size());
  uw.reserve(idx.size());
  for (size_t i = 0; i < idx.size();) {
    size_t j = i;
    int32_t best = ws[idx[i]];
    while (j < idx.size() && keys[idx[j]] == keys[idx[i]]) {
      if (flags & HSDL_DEDUP_MIN)
        best = std::min(best, ws[idx[j]]);
      else
        best = ws[idx[j]];  // last one wins (add_edge overwrites)
      ++j;
    }
    uk.push_back(keys[idx[i]]);
    uw.push_back(best);
    i = j;
  }
  csr_from_unique_edges(n, uk, uw, g);
  return HSDL_OK;
}

}  // namespace hsdl

using namespace hsdl;

extern "C" {

const char *hsdl_last_error(void) { return hsdl::last_error_cstr(); }
uint64_t hsdl_now_ns(void) { return hsdl::now_ns(); }

const char *hsdl_strerror(int code) {
  switch (code) {
    case HSDL_OK: return "ok";
    case HSDL_EINVAL: return "invalid argument";
    case HSDL_ENOMEM: return "out of memory or over the memory budget";
    case HSDL_EIO: return "i/o error";
    case HSDL_EDEVICE: return "HIP runtime error";
    case HSDL_ETIMEDOUT: return "deadline exceeded";
    case HSDL_EBUSY: return "queue full";
    case HSDL_ERANGE: return "value outside the packed label range";
    case HSDL_ENODEV: return "no usable GPU";
    case HSDL_ESTATE: return "invalid state";
    default: return "unknown error";
  }
}

static int graph_from_edges_body(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                          const int32_t *w, uint32_t flags, hsdl_graph **out) {
  if (!out || (m && (!u || !v || !w))) return fail(HSDL_EINVAL, "null argument");
  hsdl_graph *g = new hsdl_graph();
  int rc = build_from_edge_list(n, m, u, v, w, flags, *g);
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return HSDL_OK;
}

int hsdl_graph_from_edges(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                          const int32_t *w, uint32_t flags, hsdl_graph **out) {
  return hsdl::guarded([&] { return graph_from_edges_body(n, m, u, v, w, flags, out); });
}

static bool parse_int(const char *&p, long long &x) {
  while (*p == ' ' || *p == '\t') ++p;
  char *end;
  errno = 0;
  x = strtoll(p, &end, 10);
  if (end == p || errno) return false;
  p = end;
  return true;
}
static bool parse_double(const char *&p, double &x) {
  while (*p == ' ' || *p == '\t') ++p;
  char *end;
  x = strtod(p, &end);
  if (end == p) return false;
  p = end;
  return true;
}

static int graph_load_txt_body(const char *path, hsdl_graph **out) {
  if (!path || !out) return fail(HSDL_EINVAL, "null argument");
  std::ifstream in(path);
  if (!in.is_open()) return fail(HSDL_EIO, std::string("cannot open ") + path);
  std::string line;
  long long n = -1;
  std::vector<uint32_t> us, vs;
  std::vector<int32_t> ws;
  while (std::getline(in, line)) {
    const char *p = line.c_str();
    while (*p == ' ') ++p;
    if (!strncmp(p, "|V|=", 4)) {
      p += 4;
      if (!parse_int(p, n) || n < 0 || n > 0xFFFFFFFFll) return fail(HSDL_EIO, "bad |V|= line");
    } else if (!strncmp(p, "Edge", 4)) {
      p += 4;
      long long a, b;
      double wd;
      if (!parse_int(p, a) || !parse_int(p, b) || !parse_double(p, wd))
        return fail(HSDL_EIO, "bad Edge line: " + line);
      if (n < 0 || a < 0 || b < 0 || a >= n || b >= n)
        return fail(HSDL_EIO, "Edge endpoint outside |V|: " + line);
      // std::stod then an int conversion truncates; outside int32 the
      // conversion is undefined, so such weights are refused
      if (!(wd > -2147483649.0 && wd < 2147483648.0)) return fail(HSDL_EIO, "weight outside int32: " + line);
      us.push_back((uint32_t)a);
      vs.push_back((uint32_t)b);
      ws.push_back((int32_t)wd);
    }
  }
  if (n < 0) return fail(HSDL_EIO, "missing |V|= line");
  return hsdl_graph_from_edges((uint32_t)n, us.size(), us.data(), vs.data(), ws.data(), 0, out);
}

int hsdl_graph_load_txt(const char *path, hsdl_graph **out) {
  return hsdl::guarded([&] { return graph_load_txt_body(path, out); });
}

static int graph_save_txt_body(const hsdl_graph *g, const char *path) {
  if (!g || !path) return fail(HSDL_EINVAL, "null argument");
  FILE *f = fopen(path, "w");
  if (!f) return fail(HSDL_EIO, std::string("cannot open ") + path);
  uint64_t m = 0;
  for (uint32_t v = 0; v < g->n; ++v)
    for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e)
      if (v <= g->col[e]) ++m;
  fprintf(f, "|V|= %u\n|E|= %llu\n\n", g->n, (unsigned long long)m);
  for (uint32_t v = 0; v < g->n; ++v)
    for (uint64_t e = g->off[v]; e < g->off[v + 1]; ++e)
      if (v <= g->col[e]) fprintf(f, "Edge %u %u %d\n", v, g->col[e], g->w[e]);
  fprintf(f, "\nEOF\n");
  return fclose(f) == 0 ? HSDL_OK : fail(HSDL_EIO, "write failed");
}

int hsdl_graph_save_txt(const hsdl_graph *g, const char *path) {
  return hsdl::guarded([&] { return graph_save_txt_body(g, path); });
}

static int graph_load_edgelist_body(const char *path, uint32_t flags, hsdl_graph **out) {
  if (!path || !out) return fail(HSDL_EINVAL, "null argument");
  std::ifstream in(path);
  if (!in.is_open()) return fail(HSDL_EIO, std::string("cannot open ") + path);
  std::string line;
  long long n = -1;
  if (!std::getline(in, line)) return fail(HSDL_EIO, "empty file");
  // header: "... |V|= N; |E|= M)" (makedata.cpp); "|V| =" also accepted
  size_t pos = line.find("|V|=");
  size_t skip = 4;
  if (pos == std::string::npos) {
    pos = line.find("|V| =");
    skip = 5;
  }
  // like the reference's reader (GPU4HSDL_EXP/CPU_HopHL/include/graph_v_of_v/
  // graph_v_of_v.h:243-271), a header without a readable vertex count is
  // refused; so is one outside 0..2^32-1
  if (pos == std::string::npos) return fail(HSDL_EIO, "header has no |V|= count");
  {
    const char *p = line.c_str() + pos + skip;
    if (!parse_int(p, n) || n < 0 || n > 0xFFFFFFFFll) return fail(HSDL_EIO, "bad |V|= count");
  }
  std::vector<uint32_t> us, vs;
  std::vector<int32_t> ws;
  long long maxid = -1;
  while (std::getline(in, line)) {
    const char *p = line.c_str();
    long long a, b;
    double wd;
    if (!parse_int(p, a)) continue;
    if (!parse_int(p, b) || !parse_double(p, wd)) return fail(HSDL_EIO, "bad edge line: " + line);
    if (a < 0 || b < 0 || a > 0xFFFFFFFEll || b > 0xFFFFFFFEll)
      return fail(HSDL_EIO, "vertex id out of range: " + line);
    if (!(wd > -2147483649.0 && wd < 2147483648.0)) return fail(HSDL_EIO, "weight outside int32: " + line);
    maxid = std::max(maxid, std::max(a, b));
    us.push_back((uint32_t)a);
    vs.push_back((uint32_t)b);
    ws.push_back((int32_t)wd);
  }
  if (maxid >= n) return fail(HSDL_EIO, "edge endpoint outside |V|");
  return hsdl_graph_from_edges((uint32_t)n, us.size(), us.data(), vs.data(), ws.data(), flags, out);
}

int hsdl_graph_load_edgelist(const char *path, uint32_t flags, hsdl_graph **out) {
  return hsdl::guarded([&] { return graph_load_edgelist_body(path, flags, out); });
}

static int graph_reorder_by_degree_body(hsdl_graph *g, uint32_t *new_to_old) {
  if (!g) return fail(HSDL_EINVAL, "null graph");
  g->drop_device_copy();
  std::vector<uint32_t> n2o;
  reorder_by_degree(*g, n2o);
  if (new_to_old) std::copy(n2o.begin(), n2o.end(), new_to_old);
  return HSDL_OK;
}

int hsdl_graph_reorder_by_degree(hsdl_graph *g, uint32_t *new_to_old) {
  return hsdl::guarded([&] { return graph_reorder_by_degree_body(g, new_to_old); });
}

static int graph_rmat_body(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int32_t wmin, int32_t wmax, hsdl_graph **out) {
  if (!out || scale < 1 || scale > 31 || edge_factor < 1 || wmin < 1 || wmax < wmin ||
      a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return fail(HSDL_EINVAL, "bad R-MAT parameters");
  const uint32_t n = 1u << scale;
  const uint64_t samples = (uint64_t)edge_factor << scale;
  // 53-bit integer thresholds of the cumulative quadrant probabilities
  const double two53 = 9007199254740992.0;
  const uint64_t tA = (uint64_t)std::llround(a * two53);
  const uint64_t tB = (uint64_t)std::llround((a + b) * two53);
  const uint64_t tC = (uint64_t)std::llround((a + b + c) * two53);
  // Samples are drawn by T independent streams (stream i seeded seed+i*2^32
  // handles samples i, i+T, ...) so the output does not depend on threads.
  const unsigned T = 64;
  std::vector<std::vector<uint64_t>> part(T);
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t0 = 0; t0 < hw; ++t0) {
    th.emplace_back([&, t0] {
      for (unsigned s = t0; s < T; s += hw) {
        Xoshiro256ss rng(seed + ((uint64_t)s << 32));
        auto &vec = part[s];
        vec.reserve(samples / T + 1);
        for (uint64_t i = s; i < samples; i += T) {
          uint32_t r = 0, cc = 0;
          for (uint32_t l = 0; l < scale; ++l) {
            uint64_t x = rng.next() >> 11;
            uint32_t q = x < tA ? 0 : x < tB ? 1 : x < tC ? 2 : 3;
            r = (r << 1) | (q >> 1);
            cc = (cc << 1) | (q & 1);
          }
          if (r == cc) continue;  // self loop dropped
          uint32_t lo = std::min(r, cc), hi = std::max(r, cc);
          vec.push_back(((uint64_t)lo << 32) | hi);
        }
      }
    });
  }
  for (auto &x : th) x.join();
  std::vector<uint64_t> keys;
  size_t tot = 0;
  for (auto &p : part) tot += p.size();
  keys.reserve(tot);
  for (auto &p : part) {
    keys.insert(keys.end(), p.begin(), p.end());
    std::vector<uint64_t>().swap(p);
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  Xoshiro256ss wr(seed + 0x100);
  std::vector<int32_t> ws(keys.size());
  const uint32_t range = (uint32_t)(wmax - wmin + 1);
  for (size_t i = 0; i < keys.size(); ++i) ws[i] = wmin + (int32_t)wr.bounded(range);
  hsdl_graph *g = new hsdl_graph();
  csr_from_unique_edges(n, keys, ws, *g);
  std::vector<uint32_t> n2o;
  reorder_by_degree(*g, n2o);
  *out = g;
  return HSDL_OK;
}

int hsdl_graph_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int32_t wmin, int32_t wmax, hsdl_graph **out) {
  return hsdl::guarded([&] { return graph_rmat_body(scale, edge_factor, a, b, c, seed, wmin, wmax, out); });
}

static int graph_grid_body(uint32_t rows, uint32_t cols, uint64_t seed, int32_t wmin, int32_t wmax,
                    hsdl_graph **out) {
  if (!out || rows < 1 || cols < 1 || (uint64_t)rows * cols > 0xFFFFFFFFull || wmin < 1 ||
      wmax < wmin)
    return fail(HSDL_EINVAL, "bad grid parameters");
  const uint32_t n = rows * cols;
  std::vector<uint64_t> keys;
  keys.reserve(2ull * n);
  for (uint32_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c) {
      uint32_t v = r * cols + c;
      if (c + 1 < cols) keys.push_back(((uint64_t)v << 32) | (v + 1));
      if (r + 1 < rows) keys.push_back(((uint64_t)v << 32) | (v + cols));
    }
  std::sort(keys.begin(), keys.end());
  Xoshiro256ss wr(seed + 0x100);
  std::vector<int32_t> ws(keys.size());
  const uint32_t range = (uint32_t)(wmax - wmin + 1);
  for (size_t i = 0; i < keys.size(); ++i) ws[i] = wmin + (int32_t)wr.bounded(range);
  hsdl_graph *g = new hsdl_graph();
  csr_from_unique_edges(n, keys, ws, *g);
  std::vector<uint32_t> n2o;
  reorder_by_degree(*g, n2o);
  *out = g;
  return HSDL_OK;
}

int hsdl_graph_grid(uint32_t rows, uint32_t cols, uint64_t seed, int32_t wmin, int32_t wmax,
                    hsdl_graph **out) {
  return hsdl::guarded([&] { return graph_grid_body(rows, cols, seed, wmin, wmax, out); });
}

int hsdl_graph_info(const hsdl_graph *g, uint32_t *n, uint64_t *nnz) {
  if (!g) return fail(HSDL_EINVAL, "null graph");
  if (n) *n = g->n;
  if (nnz) *nnz = g->off[g->n];
  return HSDL_OK;
}

int hsdl_graph_csr(const hsdl_graph *g, uint64_t *off, uint32_t *col, int32_t *w) {
  if (!g) return fail(HSDL_EINVAL, "null graph");
  if (off) std::copy(g->off.begin(), g->off.end(), off);
  if (col) std::copy(g->col.begin(), g->col.end(), col);
  if (w) std::copy(g->w.begin(), g->w.end(), w);
  return HSDL_OK;
}

void hsdl_free_graph(hsdl_graph *g) { delete g; }

void hsdl_opts_init(hsdl_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = HSDL_MODE_BATCHED;
  o->device = -1;
}

void hsdl_server_opts_init(hsdl_server_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_batch = 65536;
  o->max_wait_us = 200;
  o->inflight = 4;
  o->queue_capacity = 1ull << 22;
}

}  // extern "C"
