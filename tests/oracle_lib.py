"""ctypes binding of the CPU oracle (oracle/libhsdl_oracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libhsdl_oracle.so")

_lib = None

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    L = ctypes.CDLL(ORACLE_SO)
    vp = ctypes.c_void_p
    L.orc_kround.restype = vp
    L.orc_kround.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_kround_grouped.restype = vp
    L.orc_kround_grouped.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_kround_gid.restype = vp
    L.orc_kround_gid.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, _i32p, ctypes.c_int]
    L.orc_block_digests.restype = None
    L.orc_block_digests.argtypes = [ctypes.c_int32, _i64p, _u32p, _u8p, _i32p, ctypes.c_int32,
                                    ctypes.c_int, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"),
                                    np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
    L.orc_dijkstra_timed.restype = None
    L.orc_dijkstra_timed.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
    L.orc_2023www_timed.restype = None
    L.orc_2023www_timed.argtypes = L.orc_dijkstra_timed.argtypes
    L.orc_replay_hubs.restype = None
    L.orc_replay_hubs.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, _i64p, _u32p,
                                  _u8p, _i32p, ctypes.c_int64, _i32p, ctypes.c_int, _i32p]
    L.orc_clean_check.restype = None
    L.orc_clean_check.argtypes = [ctypes.c_int32, _i64p, _u32p, _u8p, _i32p, _i64p, _u32p, _u8p,
                                  _i32p, ctypes.c_int64, _i32p, ctypes.c_int, _i32p]
    L.orc_2023www.restype = vp
    L.orc_2023www.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, ctypes.c_int]
    L.orc_dijkstra.restype = vp
    L.orc_dijkstra.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int, ctypes.c_int]
    L.orc_canonical_bruteforce.restype = vp
    L.orc_canonical_bruteforce.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int]
    L.orc_clean.restype = vp
    L.orc_clean.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_labels_from_arrays.restype = vp
    L.orc_labels_from_arrays.argtypes = [ctypes.c_int32, _i64p, _u32p, _u8p, _i32p]
    L.orc_labels_total.restype = ctypes.c_int64
    L.orc_labels_total.argtypes = [vp]
    L.orc_labels_n.restype = ctypes.c_int32
    L.orc_labels_n.argtypes = [vp]
    L.orc_labels_counters.restype = None
    L.orc_labels_counters.argtypes = [vp, _i64p]
    L.orc_labels_export.restype = None
    L.orc_labels_export.argtypes = [vp, _i64p, _u32p, _u8p, _i32p]
    L.orc_labels_from_arrays_p.restype = vp
    L.orc_labels_from_arrays_p.argtypes = [ctypes.c_int32, _i64p, _u32p, _u8p, _i32p, _i32p]
    L.orc_labels_parents.restype = None
    L.orc_labels_parents.argtypes = [vp, _i32p]
    L.orc_paths.restype = None
    L.orc_paths.argtypes = [vp, ctypes.c_int64, _u32p, _u32p, _i32p, ctypes.c_int32, _i32p,
                            _i32p, ctypes.c_int, ctypes.c_int]
    L.orc_labels_free.restype = None
    L.orc_labels_free.argtypes = [vp]
    L.orc_query.restype = None
    L.orc_query.argtypes = [vp, ctypes.c_int64, _u32p, _u32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int]
    L.orc_bf.restype = None
    L.orc_bf.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, ctypes.c_int32, ctypes.c_int, _i32p, ctypes.c_int]
    L.orc_load_txt.restype = vp
    L.orc_load_txt.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.orc_reorder.restype = vp
    L.orc_reorder.argtypes = [ctypes.c_int32, _i64p, _i32p, _i32p, _i32p]
    L.orc_graph_n.restype = ctypes.c_int32
    L.orc_graph_n.argtypes = [vp]
    L.orc_graph_nnz.restype = ctypes.c_int64
    L.orc_graph_nnz.argtypes = [vp]
    L.orc_graph_export.restype = None
    L.orc_graph_export.argtypes = [vp, _i64p, _i32p, _i32p]
    L.orc_graph_free.restype = None
    L.orc_graph_free.argtypes = [vp]
    _lib = L
    return L


class CSR:
    """Undirected graph as symmetric CSR with id-sorted adjacency."""

    def __init__(self, n, off, col, w):
        self.n = int(n)
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.w = np.ascontiguousarray(w, dtype=np.int32)

    @staticmethod
    def from_edges(n, u, v, w):
        """Symmetric CSR from an undirected edge list without duplicates."""
        u = np.asarray(u, dtype=np.int64)
        v = np.asarray(v, dtype=np.int64)
        w = np.asarray(w, dtype=np.int64)
        src = np.concatenate([u, v])
        dst = np.concatenate([v, u])
        ww = np.concatenate([w, w])
        order = np.lexsort((dst, src))
        src, dst, ww = src[order], dst[order], ww[order]
        off = np.zeros(n + 1, dtype=np.int64)
        np.add.at(off, src + 1, 1)
        off = np.cumsum(off)
        return CSR(n, off, dst, ww)


class OracleLabels:
    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_labels_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_labels_n(self.h)

    def total(self):
        return lib().orc_labels_total(self.h)

    def counters(self):
        c = np.zeros(7, dtype=np.int64)
        lib().orc_labels_counters(self.h, c)
        keys = ["frontier_items", "scanned_edges", "candidates", "cand_label_reads",
                "hub_label_reads", "accepted", "rounds_run"]
        return dict(zip(keys, c.tolist()))

    def arrays(self):
        n, tot = self.n, self.total()
        off = np.zeros(n + 1, dtype=np.int64)
        hub = np.zeros(tot, dtype=np.uint32)
        hop = np.zeros(tot, dtype=np.uint8)
        dist = np.zeros(tot, dtype=np.int32)
        lib().orc_labels_export(self.h, off, hub, hop, dist)
        return off, hub, hop, dist

    def parents(self):
        """Parent vertex of every label, in the order of arrays()."""
        par = np.zeros(self.total(), dtype=np.int32)
        lib().orc_labels_parents(self.h, par)
        return par

    def as_lists(self):
        off, hub, hop, dist = self.arrays()
        return [list(zip(hub[off[v]:off[v + 1]].tolist(), hop[off[v]:off[v + 1]].tolist(),
                         dist[off[v]:off[v + 1]].tolist())) for v in range(self.n)]


def kround(g, K, threads=0, mutation=0):
    return OracleLabels(lib().orc_kround(g.n, g.off, g.col, g.w, K, threads, mutation))


def kround_grouped(g, K, group, threads=0):
    """K-round labels with pruning restricted to hubs of the same id group."""
    return OracleLabels(lib().orc_kround_grouped(g.n, g.off, g.col, g.w, K, group, threads))


def kround_gid(g, K, gid, threads=0):
    """K-round labels with pruning restricted to hubs of the same group id
    (gid: one id per vertex)."""
    gid = np.ascontiguousarray(gid, np.int32)
    return OracleLabels(lib().orc_kround_gid(g.n, g.off, g.col, g.w, K, gid, threads))


def www2023(g, K, mutation=0):
    """The 2023WWW per-hub level search, single-threaded in hub order."""
    return OracleLabels(lib().orc_2023www(g.n, g.off, g.col, g.w, K, mutation))


def dijkstra(g, K, mutation=0):
    return OracleLabels(lib().orc_dijkstra(g.n, g.off, g.col, g.w, K, mutation))


def canonical(g, K):
    return OracleLabels(lib().orc_canonical_bruteforce(g.n, g.off, g.col, g.w, K))


def clean(labels, lu_mode=1, threads=0, mutation=0):
    return OracleLabels(lib().orc_clean(labels.h, lu_mode, threads, mutation))


def labels_from_arrays(n, off, hub, hop, dist):
    return OracleLabels(lib().orc_labels_from_arrays(
        n, np.ascontiguousarray(off, np.int64), np.ascontiguousarray(hub, np.uint32),
        np.ascontiguousarray(hop, np.uint8), np.ascontiguousarray(dist, np.int32)))


def labels_from_arrays_p(n, off, hub, hop, dist, parent):
    return OracleLabels(lib().orc_labels_from_arrays_p(
        n, np.ascontiguousarray(off, np.int64), np.ascontiguousarray(hub, np.uint32),
        np.ascontiguousarray(hop, np.uint8), np.ascontiguousarray(dist, np.int32),
        np.ascontiguousarray(parent, np.int32)))


def paths(labels, s, t, k, cap, threads=0, mutation=0):
    """Hop-constrained shortest paths: (edges int32[nq, cap, 2], counts int32[nq]);
    edges are (vertex, parent) pairs in the order the reference emits them."""
    s = np.ascontiguousarray(s, np.uint32)
    t = np.ascontiguousarray(t, np.uint32)
    k = np.ascontiguousarray(k, np.int32)
    edges = np.zeros((len(s), cap, 2), dtype=np.int32)
    counts = np.zeros(len(s), dtype=np.int32)
    lib().orc_paths(labels.h, len(s), s, t, k, cap, edges.reshape(-1), counts, threads, mutation)
    return edges, counts


def query(labels, s, t, k, threads=0, mutation=0):
    s = np.ascontiguousarray(s, np.uint32)
    t = np.ascontiguousarray(t, np.uint32)
    k = np.ascontiguousarray(k, np.int32)
    out = np.zeros(len(s), dtype=np.int32)
    lib().orc_query(labels.h, len(s), s, t, k, out, threads, mutation)
    return out


def bf(g, src, k, mutation=0):
    out = np.zeros(g.n, dtype=np.int32)
    lib().orc_bf(g.n, g.off, g.col, g.w, int(src), int(k), out, mutation)
    return out


def _graph_from_handle(h):
    L = lib()
    n = L.orc_graph_n(h)
    nnz = L.orc_graph_nnz(h)
    off = np.zeros(n + 1, np.int64)
    col = np.zeros(nnz, np.int32)
    w = np.zeros(nnz, np.int32)
    L.orc_graph_export(h, off, col, w)
    L.orc_graph_free(h)
    return CSR(n, off, col, w)


def load_txt(path, reorder=False):
    h = lib().orc_load_txt(path.encode(), 1 if reorder else 0)
    if not h:
        raise FileNotFoundError(path)
    return _graph_from_handle(h)


def reorder(g):
    n2o = np.zeros(g.n, np.int32)
    h = lib().orc_reorder(g.n, g.off, g.col, g.w, n2o)
    return _graph_from_handle(h), n2o


def random_graph(rng, n, p, wmax):
    """Tiny Erdos-Renyi graph (numpy Generator), degree-ordered."""
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(len(iu)) < p
    u, v = iu[keep], ju[keep]
    w = rng.integers(1, wmax + 1, size=len(u))
    g = CSR.from_edges(n, u, v, w)
    g2, _ = reorder(g)
    return g2


def _flat(labels):
    off, hub, hop, dist = labels
    return (np.ascontiguousarray(off, np.int64), np.ascontiguousarray(hub, np.uint32),
            np.ascontiguousarray(hop, np.uint8), np.ascontiguousarray(dist, np.int32))


def replay_hubs(g, K, labels, hubs, threads=0):
    """Per sampled hub: 0 when its K rounds, recomputed from the finished
    label set `labels` (off, hub, hop, dist arrays), give exactly the set's
    labels of that hub, else the first round that differs."""
    off, hub, hop, dist = _flat(labels)
    hubs = np.ascontiguousarray(hubs, np.int32)
    bad = np.zeros(len(hubs), np.int32)
    lib().orc_replay_hubs(g.n, g.off, g.col, g.w, K, off, hub, hop, dist, len(hubs), hubs,
                          threads, bad)
    return bad


def clean_check(uncleaned, cleaned, verts, threads=0):
    """Per sampled vertex: 1 when the normative clean of its list in
    `uncleaned` differs from its list in `cleaned` (both label arrays)."""
    a = _flat(uncleaned)
    b = _flat(cleaned)
    verts = np.ascontiguousarray(verts, np.int32)
    bad = np.zeros(len(verts), np.int32)
    lib().orc_clean_check(len(a[0]) - 1, *a, *b, len(verts), verts, threads, bad)
    return bad


def block_digests(labels, block=1024, threads=0):
    """tests/label_digest.block_digests computed in C (large label sets)."""
    off, hub, hop, dist = _flat(labels)
    n = len(off) - 1
    nb = (n + block - 1) // block
    counts = np.zeros(nb, np.uint64)
    dig = np.zeros(nb, np.uint64)
    lib().orc_block_digests(n, off, hub, hop, dist, block, threads, counts, dig)
    return counts, dig


def dijkstra_timed(g, K, threads=0, seconds=0.0):
    """The reference's threaded CPU_HopHL generation, stopped after `seconds`
    (0: to the end): (hubs finished, labels made, wall seconds)."""
    import time
    hubs = ctypes.c_int64()
    labels = ctypes.c_int64()
    t0 = time.perf_counter()
    lib().orc_dijkstra_timed(g.n, g.off, g.col, g.w, K, threads, seconds, ctypes.byref(hubs),
                             ctypes.byref(labels))
    return hubs.value, labels.value, time.perf_counter() - t0


def www2023_timed(g, K, threads=0, seconds=0.0):
    """The reference's threaded 2023WWW generation, stopped after `seconds`
    (0: to the end): (hubs finished, labels made, wall seconds)."""
    import time
    hubs = ctypes.c_int64()
    labels = ctypes.c_int64()
    t0 = time.perf_counter()
    lib().orc_2023www_timed(g.n, g.off, g.col, g.w, K, threads, seconds, ctypes.byref(hubs),
                            ctypes.byref(labels))
    return hubs.value, labels.value, time.perf_counter() - t0
