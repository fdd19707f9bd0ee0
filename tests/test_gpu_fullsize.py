"""Full-size parity: the HIP build of the benchmark workloads against the
oracle, at sizes the oracle cannot finish inside a test.

* R-MAT scale 18, K=5 and the 1024 x 1024 grid, K=8: the whole label set,
  before and after the clean, against committed per-block digests of the
  oracle's sets (tests/golden/make_golden.py --full), plus the oracle's
  frontier-item count.
* R-MAT scale 18: the reference checker's density
  (Hybrid_Generation_Clean/src/test.cu:83-149): 1000 sources with a hop bound
  drawn in [1, K], a hop-bounded Bellman-Ford from each, 1000 terminals per
  source, and every path re-weighed on the graph.
* R-MAT scale 20, K=5 (the 32-bit label packing, 20 + 3 + 9 bits): sampled
  hubs replayed round by round from the GPU's own labels, the normative CPU
  clean of sampled vertices of the GPU's uncleaned set, and Bellman-Ford on
  sampled pairs.
Every comparison is exact."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gpu4hophl_code_amd as H
import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)  # the box's CPU share for one GPU


def log(msg):
    print(f"[fullsize {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def csr_of(g):
    off, col, w = g.csr()
    return O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)


def bf_many(c, sources, ks):
    """Hop-bounded Bellman-Ford distances from every source (rows)."""
    def one(i):
        return O.bf(c, int(sources[i]), int(ks[i]))
    with ThreadPoolExecutor(THREADS) as ex:
        return np.stack(list(ex.map(one, range(len(sources)))))


def check_digests(idx, z, cleaned):
    arr = idx.labels()
    counts, dig = O.block_digests(arr, threads=THREADS)
    sfx = "_cleaned" if cleaned else ""
    assert int(arr[0][-1]) == int(z["total" + sfx])
    bad = np.flatnonzero((counts != z["counts" + sfx]) | (dig != z["dig" + sfx]))
    assert len(bad) == 0, f"{len(bad)} vertex blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("name,make,K", [
    ("full_c2_rmat18_k5.npz", lambda: H.Graph.rmat(18, 16, 0x5EED0002, 1, 100), 5),
    ("full_c3_grid1024_k8.npz", lambda: H.Graph.grid(1024, 1024, 0x5EED0003, 1, 1000), 8),
])
def test_full_label_sets_match_oracle_digests(golden_dir, name, make, K):
    z = np.load(os.path.join(golden_dir, name))
    g = make()
    n, nnz = g.info()
    assert n == int(z["n"]) and nnz == int(z["nnz"])
    log(f"{name}: batched build")
    idx = H.hop_constrained_two_hop_labels_generation(g, K)
    assert idx.stats()["frontier_items"] == int(z["frontier_items"])
    check_digests(idx, z, cleaned=False)
    log(f"{name}: clean")
    idx.clean()
    check_digests(idx, z, cleaned=True)
    log(f"{name}: ok, {int(z['total'])} labels")


def test_c2_reference_checker_density():
    """1000 sources x 1000 terminals with a hop bound per source in [1, 5]
    (the reference's GPU_HSDL_checker), answers equal to Bellman-Ford, and
    every extracted path a walk whose re-weighed length is the answer."""
    g = H.Graph.rmat(18, 16, 0x5EED0002, 1, 100)
    c = csr_of(g)
    K = 5
    log("c2 checker: build with parents, clean")
    idx = H.hop_constrained_two_hop_labels_generation(g, K, clean=True, parents=True)
    rng = np.random.default_rng(0x5EED0083)
    src = rng.integers(0, c.n, 1000)
    ks = rng.integers(1, K + 1, 1000)
    log("c2 checker: Bellman-Ford from 1000 sources")
    D = bf_many(c, src, ks)
    ter = rng.integers(0, c.n, (1000, 1000))
    s = np.repeat(src, 1000).astype(np.uint32)
    t = ter.reshape(-1).astype(np.uint32)
    k = np.repeat(ks, 1000).astype(np.int8)
    want = D[np.repeat(np.arange(1000), 1000), ter.reshape(-1)]
    got = idx.query(s, t, k)
    assert np.array_equal(got, want)
    log("c2 checker: paths")
    edges, counts = idx.paths(s, t, k, cap=K)
    # re-weigh every path on the CSR: weight of (a, b) by binary search in a's row
    e = edges.reshape(len(s), K, 2).astype(np.int64)
    mask = np.arange(K)[None, :] < counts[:, None]
    a, b = e[..., 0][mask], e[..., 1][mask]
    # CSR entries keyed row * n + col are globally sorted
    rows = np.repeat(np.arange(c.n, dtype=np.int64), np.diff(c.off))
    keys = rows * c.n + c.col
    pos = np.minimum(np.searchsorted(keys, a * c.n + b), len(keys) - 1)
    assert (keys[pos] == a * c.n + b).all(), "a path step is not an edge"
    wsum = np.zeros(len(s), np.int64)
    np.add.at(wsum, np.repeat(np.arange(len(s)), counts.astype(np.int64)), c.w[pos].astype(np.int64))
    reach = (got != H.INT32_MAX) & (s != t)
    assert np.array_equal(wsum[reach], got[reach].astype(np.int64))
    assert (counts[~reach] == 0).all()
    log(f"c2 checker: ok, {int(reach.sum())} reachable pairs re-weighed")


def test_c4_scale20_replay_clean_and_bellman_ford():
    """R-MAT scale 20, K=5: hubs replayed from the GPU's labels (the 8
    highest-ranked, and 200 drawn uniformly), the CPU clean of 2000 sampled
    vertices of the GPU's uncleaned set equal to the GPU clean's lists, and
    100 sources x 1000 terminals against Bellman-Ford."""
    K = 5
    log("c4: graph")
    g = H.Graph.rmat(20, 16, 0x5EED0004, 1, 100)
    c = csr_of(g)
    log("c4: batched build")
    idx = H.hop_constrained_two_hop_labels_generation(g, K)
    st = idx.stats()
    assert st["prune_label_bytes"] == 4  # 20 + 3 + 9 = 32 bits
    unc = idx.labels()
    log(f"c4: {len(unc[1])} labels exported; replay")
    rng = np.random.default_rng(0x5EED0084)
    hubs = np.unique(np.concatenate([np.arange(8), rng.integers(8, c.n, 200)])).astype(np.int32)
    bad = O.replay_hubs(c, K, unc, hubs, threads=THREADS)
    assert not bad.any(), f"hubs {hubs[bad != 0][:8]} differ at rounds {bad[bad != 0][:8]}"
    log("c4: clean")
    idx.clean()
    assert idx.stats()["clean_label_bytes"] == 4
    cln = idx.labels()
    verts = np.unique(rng.integers(0, c.n, 2000)).astype(np.int32)
    bad = O.clean_check(unc, cln, verts, threads=THREADS)
    assert not bad.any(), f"vertices {verts[bad != 0][:8]} clean differently"
    del unc
    log("c4: Bellman-Ford")
    src = rng.integers(0, c.n, 100)
    ks = rng.integers(1, K + 1, 100)
    D = bf_many(c, src, ks)
    ter = rng.integers(0, c.n, (100, 1000))
    got = idx.query(np.repeat(src, 1000).astype(np.uint32), ter.reshape(-1).astype(np.uint32),
                    np.repeat(ks, 1000).astype(np.int8))
    assert np.array_equal(got, D[np.repeat(np.arange(100), 1000), ter.reshape(-1)])
    log(f"c4: ok, {st['labels_uncleaned']} labels")
