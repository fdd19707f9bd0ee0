/*
 * hsdl.h -- C ABI of the MI355X hop-constrained shortest-distance labeling
 * (HSDL) engine.
 *
 * The reference (rucdatascience/GPU4HopHL_code) exposes this path as C++
 * functions in headers; each entry point below names the reference function it
 * replaces (file:line).  All functions return 0 (HSDL_OK) or a negative
 * HSDL_E* code; hsdl_strerror() gives text and hsdl_last_error() the detail of
 * the last failure on the calling thread.  Handles are opaque.  Calls on
 * different handles are thread-safe; calls on one handle must be serialised by
 * the caller, except hsdl_query / hsdl_query_device, which may run
 * concurrently on one index.
 *
 * Conventions shared with the reference:
 *   - vertex ids are 0..n-1; after hsdl_graph_reorder_by_degree a smaller id
 *     is a higher rank, and L(v) only holds hubs u <= v;
 *   - a label is (hub u32, hop u8, dist i32); per vertex, labels are sorted by
 *     (hub asc, hop asc, dist asc);
 *   - unreachable answers are INT32_MAX.
 */
#ifndef HSDL_H
#define HSDL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HSDL_OK = 0,
  HSDL_EINVAL = -1,    /* bad argument */
  HSDL_ENOMEM = -2,    /* host or device allocation failed / budget exceeded */
  HSDL_EIO = -3,       /* file could not be opened or parsed */
  HSDL_EDEVICE = -4,   /* HIP runtime error */
  HSDL_ETIMEDOUT = -5, /* deadline in hsdl_opts passed */
  HSDL_EBUSY = -6,     /* server queue full */
  HSDL_ERANGE = -7,    /* value outside the packed label range */
  HSDL_ENODEV = -8,    /* no usable GPU */
  HSDL_ESTATE = -9     /* call not valid in the object's current state */
};

typedef struct hsdl_graph hsdl_graph;
typedef struct hsdl_index hsdl_index;
typedef struct hsdl_server hsdl_server;

/* Flags for hsdl_graph_from_edges. */
#define HSDL_DROP_SELF_LOOPS 0x1u /* drop (v,v) edges (the reference keeps one) */
#define HSDL_DEDUP_MIN 0x2u       /* duplicate edges keep the minimum weight
                                     instead of the reference's last-wins rule */

/* ---------------------------------------------------------------- graphs */

/* Undirected graph from an edge list.  Replaces graph_v_of_v<int>::add_edge
 * (CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:92-108): both directions are
 * stored, adjacency is sorted by neighbour id, a duplicate edge overwrites the
 * earlier weight (last wins) unless HSDL_DEDUP_MIN, a self loop is stored
 * once unless HSDL_DROP_SELF_LOOPS.  Weights must be >= 1. */
int hsdl_graph_from_edges(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                          const int32_t *w, uint32_t flags, hsdl_graph **out);

/* Text format "|V|= n" / "Edge u v w".  Replaces graph_v_of_v<int>::txt_read
 * (CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:242-277). */
int hsdl_graph_load_txt(const char *path, hsdl_graph **out);

/* Writes the same text format.  Replaces graph_v_of_v<int>::txt_save
 * (CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:215-240). */
int hsdl_graph_save_txt(const hsdl_graph *g, const char *path);

/* SNAP-style edge list "u v w" per line after one header line (the ".e"
 * files under the reference's data/, written by data/<x>/makedata.cpp:1-36).
 * Edges go through the same add_edge rule as hsdl_graph_from_edges(flags). */
int hsdl_graph_load_edgelist(const char *path, uint32_t flags, hsdl_graph **out);

/* Renames vertices by (degree desc, id asc).  Replaces
 * graph_v_of_v_update_vertexIDs_by_degrees_large_to_small
 * (CPU_HopHL/include/graph_v_of_v/graph_v_of_v_update_vertexIDs_by_degrees_large_to_small.h:14-40),
 * whose std::sort leaves ties unordered; here ties go by old id.  new_to_old
 * (n entries) may be NULL. */
int hsdl_graph_reorder_by_degree(hsdl_graph *g, uint32_t *new_to_old);

/* Seeded synthetic graphs (xoshiro256** seeded through splitmix64; bounded
 * integers by Lemire's method), already in (degree desc, id asc) order.
 * R-MAT: edge_factor * 2^scale samples with quadrant probabilities
 * (a,b,c,1-a-b-c), canonicalised to (min,max), self loops and duplicates
 * dropped, weights U[wmin,wmax] drawn in sorted edge order.
 * Grid: rows x cols, 4-neighbour, weights U[wmin,wmax] in sorted edge order. */
int hsdl_graph_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                    uint64_t seed, int32_t wmin, int32_t wmax, hsdl_graph **out);
int hsdl_graph_grid(uint32_t rows, uint32_t cols, uint64_t seed, int32_t wmin,
                    int32_t wmax, hsdl_graph **out);

/* Copies the CSR to HBM on `device` (-1 = current) and keeps it with the
 * graph, so hsdl_generate starts from a device-resident graph.  Optional:
 * hsdl_generate uploads it itself otherwise.  Freed with the graph. */
int hsdl_graph_upload(hsdl_graph *g, int32_t device);
/* Frees the graph's device state: the CSR copy and the dense dedupe slots
 * and label pool that hsdl_generate keeps with it for the next build (hub
 * batch x n x 4 B, up to a quarter of the build's HBM budget; 8-12 B per
 * label slot).  hsdl_free_graph does the same.  (No counterpart in the
 * reference.) */
int hsdl_graph_release_device(hsdl_graph *g);

/* n and stored adjacency entries (2*undirected edges, self loops once). */
int hsdl_graph_info(const hsdl_graph *g, uint32_t *n, uint64_t *nnz);
/* Symmetric CSR copy: off[n+1], col[nnz], w[nnz]. */
int hsdl_graph_csr(const hsdl_graph *g, uint64_t *off, uint32_t *col, int32_t *w);
void hsdl_free_graph(hsdl_graph *g);

/* ------------------------------------------------------------- generation */

enum { HSDL_MODE_BATCHED = 0, HSDL_MODE_PLAIN = 1 };

typedef struct hsdl_opts {
  uint32_t mode;          /* HSDL_MODE_BATCHED (default) or HSDL_MODE_PLAIN  */
  uint32_t hub_batch;     /* hubs per batch; 0 = autotune                    */
  uint32_t tile;          /* candidates per prune work item, at most 256;
                             0 = 256                                       */
  uint64_t hbm_budget;    /* bytes the build may hold on the device; 0 = 80%
                             of free HBM at the call                          */
  double deadline_s;      /* wall-clock limit, 0 = none (HSDL_ETIMEDOUT)     */
  int32_t device;         /* HIP device ordinal, -1 = current                */
  uint64_t pool_init_labels;  /* first label-pool size, 0 = automatic; the
                                 pool grows (rolling back and re-running the
                                 batch that ran out) as labels accumulate   */
  uint64_t arena_init_words;  /* first size of the global tile arena for hubs
                                 whose label tile exceeds LDS, 0 = automatic */
  uint32_t kernel_timing;     /* 1: time the expand and prune launches with
                                 HIP events on the build stream (stats)     */
  uint32_t lds_tile_slots;    /* cap on the hub tile slots kept in LDS, 0 =
                                 as many as 60 KB hold; larger tiles use the
                                 global arena                               */
  uint32_t parents;           /* 1: give every label its parent vertex (the
                                 smallest frontier vertex at the label's
                                 distance), needed by hsdl_extract_paths   */
  uint32_t reserved[1];
} hsdl_opts;

/* Fills opts with the defaults above. */
void hsdl_opts_init(hsdl_opts *opts);

/* K-round hop-bounded label generation (uncleaned) on the GPU.  Replaces
 * label_gen (Hybrid_Generation_Clean/src/label_gen.cu:603-837) and, as a set,
 * hop_constrained_two_hop_labels_generation before its canonical repair
 * (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:513-604).
 * The index stays resident in HBM.  K must be 1..15.  The dense dedupe slots
 * and the label pool stay with the graph's device copy for the next build
 * on this graph (hsdl_graph_release_device frees them); calls on one graph
 * must not overlap. */
int hsdl_generate(const hsdl_graph *g, uint8_t K, const hsdl_opts *opts, hsdl_index **out);

/* Generation for one group of hubs, as the reference's hybrid driver runs
 * it per vertex group: label_gen(csr, info, L, nid_vec, nid_vec_id)
 * (Hybrid_Generation_Clean/src/label_gen.cu:603-837) clears the device
 * labels (:647, clear_L) and seeds only the group's vertices (:183-190,
 * init_T), so each hub's pruning sees only labels of hubs in its own group
 * and the union over the groups holds redundant labels for the clean
 * (Hybrid_Generation_Clean/src/test.cu:356-410).  hubs[0..nhubs) are the
 * group's vertex ids, ascending without repeats; the index holds exactly the
 * labels whose hub is in the group (a vertex outside it has no self label),
 * sorted per vertex.  Same options and resident index as hsdl_generate. */
int hsdl_generate_group(const hsdl_graph *g, uint8_t K, const hsdl_opts *opts,
                        const uint32_t *hubs, uint64_t nhubs, hsdl_index **out);

/* Vertex groups for group builds, each of at most max_size vertices:
 * group_of[v] (n entries) receives vertex v's group, numbered in the order a
 * driver takes them, and *ngroups their count.  Replaces the reference's
 * vgroup generators (Hybrid_Generation_Clean/include/vgroup/):
 *   HSDL_GROUP_CONTIGUOUS  ids in order, max_size per group;
 *   HSDL_GROUP_RANDOM      a shuffle seeded by `seed`, cut every max_size
 *                          (Random_group.cuh:10-33, whose seed is random);
 *   HSDL_GROUP_CDLP        `rounds` rounds of label propagation on the device
 *                          with community sizes capped at max_size, then
 *                          communities cut to max_size, merged greedily and
 *                          ordered by total degree, largest first
 *                          (CDLP_group.cuh:20-107; CDLP/GPU_Community_
 *                          Detection.cuh:104-210).  Deterministic: ties go
 *                          to the smallest label and the size cap reads the
 *                          previous round's sizes, where the reference's
 *                          threads race on them.
 * `device` is used by CDLP only (-1: the current device). */
#define HSDL_GROUP_CONTIGUOUS 0u
#define HSDL_GROUP_RANDOM 1u
#define HSDL_GROUP_CDLP 2u
int hsdl_graph_groups(const hsdl_graph *g, uint32_t method, uint32_t max_size, uint32_t rounds,
                      uint64_t seed, int32_t device, uint32_t *group_of, uint32_t *ngroups);

/* Canonical repair in place, on the GPU.  Replaces hop_constrained_clean_L
 * (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:422-511)
 * and gpu_clean (Hybrid_Generation_Clean/src/label_clean.cu:244-329;
 * Clean_pc/src/gpu_clean.cu:140). */
int hsdl_clean(hsdl_index *idx);

/* Frees the working buffers hsdl_clean keeps on a device (-1: every device)
 * for the next clean: its flags, counts, 32-bit label copy and tile arena,
 * and the index arrays a clean replaced.  No counterpart in the reference
 * (memory management of this library); the graph's build buffers go with
 * hsdl_graph_release_device. */
int hsdl_release_scratch(int32_t device);

/* Build statistics of the last generate/clean on this index. */
typedef struct hsdl_stats {
  uint64_t labels_uncleaned;
  uint64_t labels_cleaned; /* 0 until hsdl_clean */
  double t_generate_s;     /* device-resident CSR -> uncleaned index in HBM */
  double t_sort_s;         /* included in t_generate_s                      */
  double t_clean_s;
  uint32_t hub_batch, tile, rounds, clean_passes;
  uint64_t batches;
  uint64_t candidates;     /* deduplicated (hub,vertex) candidates checked  */
  uint64_t frontier_items;
  uint64_t scanned_edges;
  /* filled when hsdl_opts.kernel_timing is set */
  double t_expand_s, t_prune_s;     /* summed HIP-event durations          */
  uint64_t prune_launches;
  uint64_t prune_label_reads;       /* labels of L(y) streamed by the prune */
  uint64_t prune_tile_reads;        /* labels of L(s) staged into tiles     */
  uint64_t clean_seq_vertices;      /* vertices the clean walked in order   */
  uint64_t clean_label_reads;       /* labels of the heads' lists L(u) read
                                       by the clean's first pass            */
  double t_parents_s;               /* parent derivation (opts.parents),
                                       included in t_generate_s            */
  uint32_t prune_label_bytes;       /* bytes per label of L(y) the prune
                                       streamed: 4 when hub, hop and dist
                                       fit one word (a 32-bit copy of the
                                       label pool), else 8                  */
  uint32_t clean_label_bytes;       /* same for the heads' lists the clean's
                                       first pass streamed (0: no clean)   */
  uint64_t prune_row_slots;         /* dense (hub, vertex) slots the prune's
                                       row items scanned for candidates     */
} hsdl_stats;
int hsdl_index_stats(const hsdl_index *idx, hsdl_stats *st);

/* --------------------------------------------------------------- queries */

/* Batched (s,t,k) distance queries from host arrays; blocks until done.
 * Replaces hop_constrained_extract_distance
 * (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:144-222) and
 * gpu_query (Query_mdnd/src/gpu_query.cu:84-177).  k < 0 -> INT32_MAX;
 * s == t -> 0; k == 0 -> INT32_MAX.  The first call on an index makes a
 * 32-bit copy of its labels (4 B per label, kept with the index) when hub,
 * hop and distance fit one word; hsdl_query_device does the same (waiting
 * for the copy once), except on a stream under graph capture. */
int hsdl_query(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
               const int8_t *k, int32_t *out);

/* Same on device pointers, enqueued on `hip_stream` (hipStream_t, NULL = the
 * null stream); returns after enqueueing. */
int hsdl_query_device(const hsdl_index *idx, size_t n, const uint32_t *d_s,
                      const uint32_t *d_t, const int8_t *d_k, int32_t *d_out,
                      void *hip_stream);

/* hsdl_query without any allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * answers and blocking behaviour as hsdl_query, at most
 * HSDL_NOALLOC_MAX_QUERIES queries per call (HSDL_EINVAL beyond).  The
 * requests and answers are staged in a fixed device array of the library
 * (13 B per query, 13.6 MB per GPU, held until the process ends) instead of
 * buffers allocated per call; calls that overlap on one GPU share that array
 * and race, so the caller serialises them. */
#define HSDL_NOALLOC_MAX_QUERIES ((size_t)1 << 20)
int hsdl_query_noalloc(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, int32_t *out);

/* Batched hop-constrained shortest paths from host arrays; blocks until
 * done.  Replaces hop_constrained_extract_shortest_path
 * (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:224-319) and
 * returns the same edges: query i's path is counts[i] (vertex, parent) pairs
 * at edges[2*cap*i ...] (the first cap are stored, the rest of the query's
 * 2*cap words are 0), in the reference's order -- per step the source
 * side's edge, then the terminal side's.  Empty when s == t or t is out of
 * reach within k hops.  A path has at most max(k, 0) edges.  Needs parents. */
int hsdl_extract_paths(const hsdl_index *idx, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, uint32_t cap, uint32_t *edges, uint32_t *counts);

/* ---------------------------------------------------- labels in and out */

int hsdl_label_count(const hsdl_index *idx, uint64_t *total);
/* off[n+1]; hub/hop/dist[total]; hubs are vertex ids of the graph.  Any of
 * hub/hop/dist may be NULL (all three NULL: only the offsets are copied). */
int hsdl_export_labels(const hsdl_index *idx, uint64_t *off, uint32_t *hub, uint8_t *hop,
                       int32_t *dist);
/* Index from host label arrays (any per-vertex order; sorted on import).  Lets
 * a caller clean or serve labels made elsewhere (e.g. by the CPU path). */
int hsdl_index_from_labels(uint32_t n, uint8_t K, const uint64_t *off, const uint32_t *hub,
                           const uint8_t *hop, const int32_t *dist, int32_t device,
                           hsdl_index **out);
int hsdl_index_info(const hsdl_index *idx, uint32_t *n, uint8_t *K);

/* Parent vertex of every label, in hsdl_export_labels order: the vertex
 * before the label's vertex on the hub's path (the reference's
 * parent_vertex, hop_constrained_two_hop_labels.h:6-21), fixed here as the
 * smallest such vertex at the label's distance.  HSDL_EINVAL when the index
 * has none (built without opts.parents). */
int hsdl_export_parents(const hsdl_index *idx, uint32_t *parent);
/* Attaches parents (host array, one per label in hsdl_export_labels order)
 * to an index made from imported labels. */
int hsdl_index_set_parents(hsdl_index *idx, const uint32_t *parent);

/* Label checkpoint in the reference's dump layout (Clean_pc/main.cu:11-57:
 * size_t n, then per vertex size_t count + count x {int hub, parent, hop,
 * dist}).  parent is written as 0 when the index has no parents; a file
 * whose parents are all 0 loads without parents (a real parent set gives
 * vertex v >= 1 the parent v on its self label). */
int hsdl_save_labels(const hsdl_index *idx, const char *path);
int hsdl_load_labels(const char *path, uint8_t K, int32_t device, hsdl_index **out);

void hsdl_free_index(hsdl_index *idx);

/* --------------------------------------------------------------- serving */

/* Continuous-batching query server over a resident index (no counterpart in
 * the reference, which runs one query at a time: Query_mdnd/src/gpu_query.cu:139-155).
 * Requests submitted while earlier batches run on the GPU are packed into the
 * next launch as in-flight slots and device memory free up; a batch closes
 * when it holds max_batch queries or its oldest query has waited max_wait_us. */
typedef struct hsdl_server_opts {
  uint32_t max_batch;      /* queries per launch (default 65536)            */
  uint32_t max_wait_us;    /* batching window (default 200)                 */
  uint32_t inflight;       /* concurrent batches on the device (default 4)  */
  uint64_t queue_capacity; /* submitted, not yet polled (default 1<<22)     */
  uint64_t hbm_budget;     /* staging bytes for batch slots (13 B per
                              request per slot); 0 = a quarter of the free
                              HBM at creation.  Admission by memory is a
                              static cap: the slot count is fixed at
                              creation as min(inflight, budget / slot
                              size), and a closed batch waits for a free
                              slot.  Mapped staging holds the slots in
                              pinned host memory, copy staging in HBM.     */
  uint32_t staging;        /* HSDL_STAGE_MAPPED (default) or HSDL_STAGE_COPY */
  uint32_t reserved[7];
} hsdl_server_opts;
/* Where a batch's requests and answers live: MAPPED = pinned, device-mapped
 * host memory read and written by the query kernel itself (one launch per
 * batch, no DMA); COPY = device buffers filled and drained by async copies. */
enum { HSDL_STAGE_MAPPED = 0, HSDL_STAGE_COPY = 1 };

void hsdl_server_opts_init(hsdl_server_opts *o);
int hsdl_server_create(const hsdl_index *idx, const hsdl_server_opts *o, hsdl_server **out);
/* Enqueue n requests; ids are assigned consecutively from *first_id.
 * arrival_ns (may be NULL = now, CLOCK_MONOTONIC ns) stamps latency start.
 * Returns HSDL_EBUSY if the queue lacks room (nothing enqueued). */
int hsdl_server_submit(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, const uint64_t *arrival_ns, uint64_t *first_id);
/* Copies up to cap finished results (ids, answers, latencies in ns) in
 * completion order; *n_out receives the count. */
int hsdl_server_poll(hsdl_server *srv, size_t cap, uint64_t *ids, int32_t *answers,
                     uint64_t *latency_ns, size_t *n_out);
/* Blocks until every submitted request has finished. */
int hsdl_server_drain(hsdl_server *srv);
typedef struct hsdl_server_stats {
  uint64_t submitted, completed, batches;
  double mean_batch;
} hsdl_server_stats;
int hsdl_server_stats_get(const hsdl_server *srv, hsdl_server_stats *st);
void hsdl_server_destroy(hsdl_server *srv);

/* Open-loop trace replay: submits request i at start + arrival_offset_ns[i]
 * (start = the call), stamped with that arrival time, from the calling thread
 * while the server's own thread batches and launches; returns once all have
 * finished, with answers[i] and latency_ns[i] (finish - arrival) per request.
 * arrival offsets must be non-decreasing. */
int hsdl_server_replay(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, const uint64_t *arrival_offset_ns, int32_t *answers,
                       uint64_t *latency_ns);

/* -------------------------------------------------------------- misc */

/* Device copy benchmark (no counterpart in the reference): copies `bytes`
 * between two device buffers `iters` times with 16 B vector accesses and
 * reports read+write GB/s -- the achievable HBM bandwidth the kernels are
 * compared with, and the known-traffic kernel PMC byte counters are
 * calibrated against. */
int hsdl_device_copy_bench(uint64_t bytes, uint32_t iters, double *gbps);
/* Returning 32-bit atomicMin on uniformly random words of a `bytes` array:
 * *gops = atomics per second / 1e9 (the dense dedupe slots' operation; no
 * counterpart in the reference, used for the expansion roofline). */
int hsdl_device_atomic_bench(uint64_t bytes, uint64_t ops, double *gops);

/* Random 4-byte gathers (no counterpart in the reference): `reads` loads at
 * uniformly random words of a `bytes` array; *gbps = 4 B per read / time.
 * The known-traffic kernel for calibrating PMC byte counters on random
 * access, where each load moves a whole memory sector. */
int hsdl_device_gather_bench(uint64_t bytes, uint64_t reads, double *gbps);

/* SHA-256 (hex) of the library's sources as they were built: which
 * sources a loaded libhsdl.so comes from (no counterpart in the reference). */
const char *hsdl_build_id(void);

const char *hsdl_strerror(int code);
const char *hsdl_last_error(void);
uint64_t hsdl_now_ns(void);

#ifdef __cplusplus
}
#endif
#endif /* HSDL_H */
