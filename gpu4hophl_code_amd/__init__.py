"""MI355X-native hop-constrained shortest-distance labeling (HSDL).

Python host binding over the C ABI of ``libhsdl.so`` (``include/hsdl.h``).
The names mirror the reference's entry points
(rucdatascience/GPU4HopHL_code):

* ``Graph`` ~ ``graph_v_of_v<int>`` (``add_edge``/``txt_read``/``txt_save``,
  CPU_HopHL/include/graph_v_of_v/graph_v_of_v.h:92-108,215-277) plus the
  degree renaming of
  graph_v_of_v_update_vertexIDs_by_degrees_large_to_small.h:14-40;
* ``hop_constrained_two_hop_labels_generation`` ~ the generation of
  CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels_generation.h:513-604
  and label_gen of Hybrid_Generation_Clean/src/label_gen.cu:603;
* ``hop_constrained_clean_L`` ~ the canonical repair at :422-511;
* ``hop_constrained_extract_distance`` ~
  CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:144-222, batched
  (``gpu_query`` of Query_mdnd/src/gpu_query.cu:84 is the same call).

All compute runs in the HIP library; there is no CPU fallback.  Importing this
package loads ``libhsdl.so`` and raises if it has not been built.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "HSDLError", "Graph", "Index", "Server", "lib", "hop_constrained_two_hop_labels_generation",
    "hop_constrained_clean_L", "hop_constrained_extract_distance", "gpu_query",
    "MODE_BATCHED", "MODE_PLAIN", "DROP_SELF_LOOPS", "DEDUP_MIN", "INT32_MAX",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhsdl.so")

MODE_BATCHED = 0
MODE_PLAIN = 1
DROP_SELF_LOOPS = 0x1
DEDUP_MIN = 0x2
NOALLOC_MAX_QUERIES = 1 << 20
INT32_MAX = 2147483647

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


class hsdl_opts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("hub_batch", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("deadline_s", ctypes.c_double), ("device", ctypes.c_int32),
                ("pool_init_labels", ctypes.c_uint64), ("arena_init_words", ctypes.c_uint64),
                ("kernel_timing", ctypes.c_uint32), ("lds_tile_slots", ctypes.c_uint32),
                ("parents", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 1)]


class hsdl_stats(ctypes.Structure):
    _fields_ = [("labels_uncleaned", ctypes.c_uint64), ("labels_cleaned", ctypes.c_uint64),
                ("t_generate_s", ctypes.c_double), ("t_sort_s", ctypes.c_double),
                ("t_clean_s", ctypes.c_double), ("hub_batch", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("rounds", ctypes.c_uint32),
                ("clean_passes", ctypes.c_uint32), ("batches", ctypes.c_uint64),
                ("candidates", ctypes.c_uint64), ("frontier_items", ctypes.c_uint64),
                ("scanned_edges", ctypes.c_uint64), ("t_expand_s", ctypes.c_double),
                ("t_prune_s", ctypes.c_double), ("prune_launches", ctypes.c_uint64),
                ("prune_label_reads", ctypes.c_uint64), ("prune_tile_reads", ctypes.c_uint64),
                ("clean_seq_vertices", ctypes.c_uint64), ("clean_label_reads", ctypes.c_uint64),
                ("t_parents_s", ctypes.c_double), ("prune_label_bytes", ctypes.c_uint32),
                ("clean_label_bytes", ctypes.c_uint32), ("prune_row_slots", ctypes.c_uint64)]


class hsdl_server_opts(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("max_wait_us", ctypes.c_uint32),
                ("inflight", ctypes.c_uint32), ("queue_capacity", ctypes.c_uint64),
                ("hbm_budget", ctypes.c_uint64), ("staging", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 7)]


class hsdl_server_stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("mean_batch", ctypes.c_double)]


# (name, restype, argtypes) of every entry point declared in include/hsdl.h
_SIGNATURES = [
    ("hsdl_graph_from_edges", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p, _i32p, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("hsdl_graph_load_txt", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    ("hsdl_graph_save_txt", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("hsdl_graph_load_edgelist", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("hsdl_graph_reorder_by_degree", ctypes.c_int, [_vp, _vp]),
    ("hsdl_graph_rmat", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("hsdl_graph_grid", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("hsdl_graph_upload", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("hsdl_graph_release_device", ctypes.c_int, [_vp]),
    ("hsdl_graph_info", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]),
    ("hsdl_graph_csr", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("hsdl_free_graph", None, [_vp]),
    ("hsdl_opts_init", None, [ctypes.POINTER(hsdl_opts)]),
    ("hsdl_generate", ctypes.c_int, [_vp, ctypes.c_uint8, ctypes.POINTER(hsdl_opts), ctypes.POINTER(_vp)]),
    ("hsdl_generate_group", ctypes.c_int, [_vp, ctypes.c_uint8, ctypes.POINTER(hsdl_opts), _u32p,
                                           ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("hsdl_graph_groups", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint64, ctypes.c_int32, _u32p,
                                          ctypes.POINTER(ctypes.c_uint32)]),
    ("hsdl_clean", ctypes.c_int, [_vp]),
    ("hsdl_release_scratch", ctypes.c_int, [ctypes.c_int32]),
    ("hsdl_index_stats", ctypes.c_int, [_vp, ctypes.POINTER(hsdl_stats)]),
    ("hsdl_query", ctypes.c_int, [_vp, ctypes.c_size_t, _u32p, _u32p, _i8p, _i32p]),
    ("hsdl_query_device", ctypes.c_int, [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp]),
    ("hsdl_query_noalloc", ctypes.c_int, [_vp, ctypes.c_size_t, _u32p, _u32p, _i8p, _i32p]),
    ("hsdl_extract_paths", ctypes.c_int, [_vp, ctypes.c_size_t, _u32p, _u32p, _i8p, ctypes.c_uint32, _vp, _u32p]),
    ("hsdl_export_parents", ctypes.c_int, [_vp, _u32p]),
    ("hsdl_index_set_parents", ctypes.c_int, [_vp, _u32p]),
    ("hsdl_label_count", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    ("hsdl_export_labels", ctypes.c_int, [_vp, _u64p, _vp, _vp, _vp]),
    ("hsdl_index_from_labels", ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint8, _u64p, _u32p, _u8p, _i32p, ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("hsdl_index_info", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)]),
    ("hsdl_save_labels", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("hsdl_load_labels", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint8, ctypes.c_int32, ctypes.POINTER(_vp)]),
    ("hsdl_free_index", None, [_vp]),
    ("hsdl_server_opts_init", None, [ctypes.POINTER(hsdl_server_opts)]),
    ("hsdl_server_create", ctypes.c_int, [_vp, ctypes.POINTER(hsdl_server_opts), ctypes.POINTER(_vp)]),
    ("hsdl_server_submit", ctypes.c_int, [_vp, ctypes.c_size_t, _u32p, _u32p, _i8p, _vp, ctypes.POINTER(ctypes.c_uint64)]),
    ("hsdl_server_poll", ctypes.c_int, [_vp, ctypes.c_size_t, _u64p, _i32p, _u64p, ctypes.POINTER(ctypes.c_size_t)]),
    ("hsdl_server_drain", ctypes.c_int, [_vp]),
    ("hsdl_server_stats_get", ctypes.c_int, [_vp, ctypes.POINTER(hsdl_server_stats)]),
    ("hsdl_server_destroy", None, [_vp]),
    ("hsdl_server_replay", ctypes.c_int, [_vp, ctypes.c_size_t, _u32p, _u32p, _i8p, _u64p, _i32p, _u64p]),
    ("hsdl_device_copy_bench", ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    ("hsdl_device_gather_bench", ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    ("hsdl_device_atomic_bench", ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    ("hsdl_build_id", ctypes.c_char_p, []),
    ("hsdl_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("hsdl_last_error", ctypes.c_char_p, []),
    ("hsdl_now_ns", ctypes.c_uint64, []),
]

EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


class HSDLError(RuntimeError):
    def __init__(self, code, detail=""):
        self.code = code
        msg = _lib.hsdl_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"hsdl error {code} ({msg}): {detail}")


def lib():
    """Loads libhsdl.so (built by __graft_entry__.build / csrc/Makefile)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C gpu4hophl_code_amd/csrc` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise HSDLError(rc, lib().hsdl_last_error().decode())


def now_ns():
    return lib().hsdl_now_ns()


# the sources libhsdl.so is built from, in the order csrc/Makefile hashes them
HASHED_SOURCES = ["index.hip", "gen.hip", "clean.hip", "query.hip", "server.hip", "util.hip",
                  "parents.hip", "groups.hip", "graph.cpp", "../../include/hsdl.h", "common.hpp", "device.hpp"]


def source_hash():
    """SHA-256 of the library's sources in this tree (hex)."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    for name in HASHED_SOURCES:
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def build_id():
    """SHA-256 of the sources the loaded libhsdl.so was built from."""
    return lib().hsdl_build_id().decode()


class Graph:
    """Undirected weighted graph held by the library (host CSR)."""

    def __init__(self, handle):
        self.h = _vp(handle) if not isinstance(handle, _vp) else handle

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.hsdl_free_graph(self.h)
            self.h = None

    @staticmethod
    def from_edges(n, u, v, w, flags=0):
        h = _vp()
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        w = np.ascontiguousarray(w, np.int32)
        _check(lib().hsdl_graph_from_edges(n, len(u), u, v, w, flags, ctypes.byref(h)))
        return Graph(h)

    @staticmethod
    def load_txt(path):
        h = _vp()
        _check(lib().hsdl_graph_load_txt(os.fsencode(path), ctypes.byref(h)))
        return Graph(h)

    @staticmethod
    def load_edgelist(path, flags=0):
        h = _vp()
        _check(lib().hsdl_graph_load_edgelist(os.fsencode(path), flags, ctypes.byref(h)))
        return Graph(h)

    @staticmethod
    def rmat(scale, edge_factor, seed, wmin=1, wmax=100, a=0.57, b=0.19, c=0.19):
        h = _vp()
        _check(lib().hsdl_graph_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax, ctypes.byref(h)))
        return Graph(h)

    @staticmethod
    def grid(rows, cols, seed, wmin=1, wmax=1000):
        h = _vp()
        _check(lib().hsdl_graph_grid(rows, cols, seed, wmin, wmax, ctypes.byref(h)))
        return Graph(h)

    def upload(self, device=-1):
        """Keeps a device copy of the CSR for later builds."""
        _check(lib().hsdl_graph_upload(self.h, device))

    def release_device(self):
        """Frees the device CSR copy and the dedupe slots kept between builds."""
        _check(lib().hsdl_graph_release_device(self.h))

    def save_txt(self, path):
        _check(lib().hsdl_graph_save_txt(self.h, os.fsencode(path)))

    def reorder_by_degree(self):
        """Renames vertices by (degree desc, id asc); returns new_to_old."""
        n, _ = self.info()
        n2o = np.zeros(n, np.uint32)
        _check(lib().hsdl_graph_reorder_by_degree(self.h, n2o.ctypes.data_as(_vp)))
        return n2o

    def info(self):
        n = ctypes.c_uint32()
        nnz = ctypes.c_uint64()
        _check(lib().hsdl_graph_info(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        return n.value, nnz.value

    @property
    def n(self):
        return self.info()[0]

    GROUP_METHODS = {"contiguous": 0, "random": 1, "cdlp": 2}

    def groups(self, max_size, method="cdlp", rounds=20, seed=0, device=-1):
        """Vertex groups of at most max_size vertices (hsdl_graph_groups):
        the group of every vertex, numbered in build order, and their count."""
        out = np.zeros(self.n, np.uint32)
        cnt = ctypes.c_uint32(0)
        _check(lib().hsdl_graph_groups(self.h, self.GROUP_METHODS[method], max_size, rounds, seed,
                                       device, out, ctypes.byref(cnt)))
        return out, cnt.value

    def csr(self):
        n, nnz = self.info()
        off = np.zeros(n + 1, np.uint64)
        col = np.zeros(nnz, np.uint32)
        w = np.zeros(nnz, np.int32)
        _check(lib().hsdl_graph_csr(self.h, off.ctypes.data_as(_vp), col.ctypes.data_as(_vp),
                                    w.ctypes.data_as(_vp)))
        return off, col, w


class Index:
    """Label index resident in HBM."""

    def __init__(self, handle):
        self.h = handle if isinstance(handle, _vp) else _vp(handle)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.hsdl_free_index(self.h)
            self.h = None

    def free(self):
        if self.h is not None and self.h.value:
            lib().hsdl_free_index(self.h)
        self.h = None

    @staticmethod
    def from_labels(n, K, off, hub, hop, dist, device=-1):
        h = _vp()
        _check(lib().hsdl_index_from_labels(
            n, K, np.ascontiguousarray(off, np.uint64), np.ascontiguousarray(hub, np.uint32),
            np.ascontiguousarray(hop, np.uint8), np.ascontiguousarray(dist, np.int32), device,
            ctypes.byref(h)))
        return Index(h)

    @staticmethod
    def load(path, K, device=-1):
        h = _vp()
        _check(lib().hsdl_load_labels(os.fsencode(path), K, device, ctypes.byref(h)))
        return Index(h)

    def save(self, path):
        _check(lib().hsdl_save_labels(self.h, os.fsencode(path)))

    def info(self):
        n = ctypes.c_uint32()
        K = ctypes.c_uint8()
        _check(lib().hsdl_index_info(self.h, ctypes.byref(n), ctypes.byref(K)))
        return n.value, K.value

    def label_count(self):
        t = ctypes.c_uint64()
        _check(lib().hsdl_label_count(self.h, ctypes.byref(t)))
        return t.value

    def stats(self):
        s = hsdl_stats()
        _check(lib().hsdl_index_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in hsdl_stats._fields_}

    def labels(self):
        """(off[n+1], hub, hop, dist) with per-vertex (hub, hop, dist) order."""
        n, _ = self.info()
        tot = self.label_count()
        off = np.zeros(n + 1, np.uint64)
        hub = np.zeros(tot, np.uint32)
        hop = np.zeros(tot, np.uint8)
        dist = np.zeros(tot, np.int32)
        _check(lib().hsdl_export_labels(self.h, off, hub.ctypes.data_as(_vp),
                                        hop.ctypes.data_as(_vp), dist.ctypes.data_as(_vp)))
        return off, hub, hop, dist

    def parents(self):
        """Parent vertex of every label, in the order of labels() (needs an
        index built with parents=True)."""
        out = np.zeros(self.label_count(), np.uint32)
        _check(lib().hsdl_export_parents(self.h, out))
        return out

    def set_parents(self, parent):
        """Attaches parents (one per label, labels() order) to an index."""
        _check(lib().hsdl_index_set_parents(self.h, np.ascontiguousarray(parent, np.uint32)))

    def paths(self, s, t, k, cap=None):
        """Hop-constrained shortest paths: (edges uint32[nq, cap, 2], counts
        uint32[nq]); edges are (vertex, parent) pairs in the reference's
        order.  cap defaults to the largest k (a path has at most k edges)."""
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        k = np.ascontiguousarray(k, np.int8)
        if not (len(s) == len(t) == len(k)):
            raise ValueError("s, t, k must have equal length")
        if cap is None:
            cap = max(1, int(k.max(initial=0)))
        edges = np.zeros((len(s), cap, 2), np.uint32)
        counts = np.zeros(len(s), np.uint32)
        _check(lib().hsdl_extract_paths(self.h, len(s), s, t, k, cap, edges.ctypes.data_as(_vp),
                                        counts))
        return edges, counts

    def hubs(self):
        """(off[n+1], hub) only."""
        n, _ = self.info()
        off = np.zeros(n + 1, np.uint64)
        hub = np.zeros(self.label_count(), np.uint32)
        _check(lib().hsdl_export_labels(self.h, off, hub.ctypes.data_as(_vp), None, None))
        return off, hub

    def offsets(self):
        """off[n+1] only (label list of v = [off[v], off[v+1]))."""
        n, _ = self.info()
        off = np.zeros(n + 1, np.uint64)
        _check(lib().hsdl_export_labels(self.h, off, None, None, None))
        return off

    def clean(self):
        _check(lib().hsdl_clean(self.h))
        return self

    def query(self, s, t, k):
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        k = np.ascontiguousarray(k, np.int8)
        if not (len(s) == len(t) == len(k)):
            raise ValueError("s, t, k must have equal length")
        out = np.zeros(len(s), np.int32)
        _check(lib().hsdl_query(self.h, len(s), s, t, k, out))
        return out

    def query_noalloc(self, s, t, k):
        """hsdl_query without device allocation (at most NOALLOC_MAX_QUERIES
        per call; the library's fixed staging array, shared per GPU)."""
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        k = np.ascontiguousarray(k, np.int8)
        if not (len(s) == len(t) == len(k)):
            raise ValueError("s, t, k must have equal length")
        out = np.zeros(len(s), np.int32)
        _check(lib().hsdl_query_noalloc(self.h, len(s), s, t, k, out))
        return out

    def query_device(self, d_s, d_t, d_k, d_out, n, stream=0):
        """Device-pointer form (e.g. torch tensors' data_ptr()); enqueue only."""
        _check(lib().hsdl_query_device(self.h, n, _vp(d_s), _vp(d_t), _vp(d_k), _vp(d_out),
                                       _vp(stream)))


def _opts(mode=MODE_BATCHED, hub_batch=0, tile=0, hbm_budget=0, deadline_s=0.0, device=-1,
          pool_init_labels=0, arena_init_words=0, kernel_timing=False, lds_tile_slots=0,
          parents=False):
    o = hsdl_opts()
    lib().hsdl_opts_init(ctypes.byref(o))
    o.parents = 1 if parents else 0
    o.kernel_timing = 1 if kernel_timing else 0
    o.lds_tile_slots = lds_tile_slots
    o.pool_init_labels = pool_init_labels
    o.arena_init_words = arena_init_words
    o.mode = mode
    o.hub_batch = hub_batch
    o.tile = tile
    o.hbm_budget = hbm_budget
    o.deadline_s = deadline_s
    o.device = device
    return o


def hop_constrained_two_hop_labels_generation(graph, K, mode=MODE_BATCHED, hub_batch=0, tile=0,
                                              hbm_budget=0, deadline_s=0.0, device=-1,
                                              clean=False, pool_init_labels=0,
                                              arena_init_words=0, kernel_timing=False,
                                              lds_tile_slots=0, parents=False, hubs=None):
    """Generates the K-hop label index of `graph` on the GPU (uncleaned unless
    clean=True, which also applies hop_constrained_clean_L); parents=True
    also gives every label its parent vertex (for path extraction).  `hubs`
    (ascending vertex ids) restricts the build to one hub group whose
    pruning sees only the group's labels (hsdl_generate_group)."""
    o = _opts(mode, hub_batch, tile, hbm_budget, deadline_s, device, pool_init_labels,
              arena_init_words, kernel_timing, lds_tile_slots, parents)
    h = _vp()
    if hubs is None:
        _check(lib().hsdl_generate(graph.h, K, ctypes.byref(o), ctypes.byref(h)))
    else:
        hubs = np.ascontiguousarray(hubs, np.uint32)
        _check(lib().hsdl_generate_group(graph.h, K, ctypes.byref(o), hubs, len(hubs),
                                         ctypes.byref(h)))
    idx = Index(h)
    if clean:
        idx.clean()
    return idx


def release_scratch(device=-1):
    """Frees the working buffers the clean keeps on a device for the next
    clean (-1: every device)."""
    _check(lib().hsdl_release_scratch(device))


def hop_constrained_clean_L(index):
    """Canonical repair of `index` in place (on the GPU)."""
    return index.clean()


def hop_constrained_extract_distance(index, s, t, k):
    """Distances of (s,t,k) queries; scalars or arrays.  INT32_MAX = none."""
    scalar = np.isscalar(s)
    out = index.query(np.atleast_1d(s), np.atleast_1d(t), np.atleast_1d(k))
    return int(out[0]) if scalar else out


gpu_query = hop_constrained_extract_distance


def hop_constrained_extract_shortest_path(index, s, t, k):
    """Edges [(vertex, parent), ...] of the hop-constrained shortest path from
    s to t, in the order hop_constrained_extract_shortest_path
    (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:224-319)
    returns them; empty when s == t or t is out of reach."""
    edges, counts = index.paths([s], [t], [k], cap=max(1, int(k)))
    return [tuple(int(x) for x in e) for e in edges[0, :counts[0]]]


class Server:
    """Continuous-batching query server over a resident index."""

    STAGING = {"mapped": 0, "copy": 1}

    def __init__(self, index, max_batch=65536, max_wait_us=200, inflight=4,
                 queue_capacity=1 << 22, hbm_budget=0, staging="mapped"):
        o = hsdl_server_opts()
        lib().hsdl_server_opts_init(ctypes.byref(o))
        o.max_batch = max_batch
        o.max_wait_us = max_wait_us
        o.inflight = inflight
        o.queue_capacity = queue_capacity
        o.hbm_budget = hbm_budget
        o.staging = self.STAGING[staging]
        self.index = index  # keeps the index alive
        self.h = _vp()
        _check(lib().hsdl_server_create(index.h, ctypes.byref(o), ctypes.byref(self.h)))

    def submit(self, s, t, k, arrival_ns=None):
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        k = np.ascontiguousarray(k, np.int8)
        first = ctypes.c_uint64()
        arr = None
        if arrival_ns is not None:
            arr = np.ascontiguousarray(arrival_ns, np.uint64)
        rc = lib().hsdl_server_submit(self.h, len(s), s, t, k,
                                      arr.ctypes.data_as(_vp) if arr is not None else None,
                                      ctypes.byref(first))
        if rc == -6:  # HSDL_EBUSY
            return None
        _check(rc)
        return first.value

    def poll(self, cap=1 << 20):
        ids = np.zeros(cap, np.uint64)
        ans = np.zeros(cap, np.int32)
        lat = np.zeros(cap, np.uint64)
        got = ctypes.c_size_t()
        _check(lib().hsdl_server_poll(self.h, cap, ids, ans, lat, ctypes.byref(got)))
        g = got.value
        return ids[:g], ans[:g], lat[:g]

    def drain(self):
        _check(lib().hsdl_server_drain(self.h))

    def stats(self):
        s = hsdl_server_stats()
        _check(lib().hsdl_server_stats_get(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in hsdl_server_stats._fields_}

    def replay(self, s, t, k, arrival_offset_ns):
        """Open-loop replay; returns (answers, latency_ns) per request."""
        s = np.ascontiguousarray(s, np.uint32)
        t = np.ascontiguousarray(t, np.uint32)
        k = np.ascontiguousarray(k, np.int8)
        arr = np.ascontiguousarray(arrival_offset_ns, np.uint64)
        ans = np.zeros(len(s), np.int32)
        lat = np.zeros(len(s), np.uint64)
        _check(lib().hsdl_server_replay(self.h, len(s), s, t, k, arr, ans, lat))
        return ans, lat

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().hsdl_server_destroy(self.h)
        self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()
