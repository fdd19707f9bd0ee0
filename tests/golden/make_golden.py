"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

Run from the repo root after building (python tests/golden/make_golden.py).
Inputs are the library's seeded generators (their CSR is stored too, so a
change in a generator shows up as a fixture mismatch); expected outputs come
from the oracle (oracle/hsdl_oracle.cpp), which tests/test_oracle_pins.py pins
against brute-force canonical labels and hop-bounded distances.

Fixtures:
  c1_rmat12_k4.npz   R-MAT scale 12, edge factor 8, (0.57,0.19,0.19,0.05),
                     weights U[1,100], degree order, K=4: graph CSR, labels
                     (uncleaned == cleaned), 10,000 uniform (s,t) at k=4 and
                     their answers.
  fixtures_k.npz     the reference's fixture graphs (simple_iterative_tests,
                     simple_iterative_tests_2023www, euroroad2) in degree
                     order: cleaned labels for several K.
  dijkstra_k.npz     the CPU_HopHL path (Dijkstra generation, then the
                     canonical clean) on the R-MAT scale-12 graph above (K=4)
                     and on the three fixture graphs, K in {1, 2, 3, 4, 5, 8}:
                     cleaned labels.
  full_c2_rmat18_k5.npz, full_c3_grid1024_k8.npz (--full; minutes of CPU)
                     the K-round label sets of R-MAT scale 18, edge factor
                     16, weights U[1,100], K=5 and of the 1024 x 1024 grid,
                     weights U[1,1000], K=8: per block of 1024 vertices the
                     label count and a 64-bit digest (tests/label_digest.py),
                     the same for the set after the canonical clean, the
                     label totals and the oracle's algorithmic-byte
                     counters (frontier items, scanned edges, candidates,
                     labels of L_{<h}(y) and L_{<h}(s) read, accepted).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import gpu4hophl_code_amd as H  # noqa: E402
import label_digest as D  # noqa: E402
import oracle_lib as O  # noqa: E402

DIJKSTRA_KS = (1, 2, 3, 4, 5, 8)
FULL = {
    "c2": ("full_c2_rmat18_k5.npz", lambda: H.Graph.rmat(18, 16, 0x5EED0002, 1, 100), 5),
    "c3": ("full_c3_grid1024_k8.npz", lambda: H.Graph.grid(1024, 1024, 0x5EED0003, 1, 1000), 8),
}

C1_SEED = 0x5EED0001
C1_QSEED = 0x5EED0001 + 0x200


def csr_of(g):
    off, col, w = g.csr()
    return O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)


def c1_queries(n, count=10000, seed=C1_QSEED):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, count, dtype=np.uint32)
    t = rng.integers(0, n, count, dtype=np.uint32)
    k = np.full(count, 4, np.int8)
    return s, t, k


def fixture_graphs():
    names = ["simple_iterative_tests_2023www.txt", "simple_iterative_tests.txt", "euroroad2.txt"]
    for nm in names:
        g = H.Graph.load_txt(os.path.join(HERE, nm))
        g.reorder_by_degree()
        yield nm, g


def write_full(name, threads):
    fname, make, K = FULL[name]
    c = csr_of(make())
    lab = O.kround(c, K, threads=threads)
    off, hub, hop, dist = lab.arrays()
    counts, dig = D.block_digests(off, hub, hop, dist)
    del off, hub, hop, dist
    ctr = {k: np.int64(v) for k, v in lab.counters().items()}
    # the canonical clean of the set (the normative CPU predicate)
    cl = O.clean(lab, 1, threads=threads)
    ccounts, cdig = D.block_digests(*cl.arrays())
    np.savez(os.path.join(HERE, fname), counts=counts, dig=dig, total=np.int64(lab.total()),
             counts_cleaned=ccounts, dig_cleaned=cdig, total_cleaned=np.int64(cl.total()),
             n=np.int64(c.n), nnz=np.int64(len(c.col)), **ctr)
    print(fname, lab.total(), cl.total(), ctr)


def write_dijkstra():
    out = {}
    # the R-MAT scale-12 workload at its K (4); the fixture graphs at every K
    graphs = [("rmat12", H.Graph.rmat(12, 8, C1_SEED, 1, 100), (4,))] + \
        [(nm, g, DIJKSTRA_KS) for nm, g in fixture_graphs()]
    for nm, g, ks in graphs:
        c = csr_of(g)
        for K in ks:
            lab = O.clean(O.dijkstra(c, K))
            o2, h2, p2, d2 = lab.arrays()
            key = f"{nm}|{K}"
            out[key + "|off"] = o2
            out[key + "|hub"] = h2
            out[key + "|hop"] = p2
            out[key + "|dist"] = d2
    np.savez_compressed(os.path.join(HERE, "dijkstra_k.npz"), **out)


def main():
    if "--full" in sys.argv:
        names = [a for a in sys.argv[1:] if a in FULL] or list(FULL)
        for nm in names:
            write_full(nm, threads=min(16, os.cpu_count() or 1))
        return
    write_dijkstra()
    g = H.Graph.rmat(12, 8, C1_SEED, 1, 100)
    c = csr_of(g)
    kr = O.kround(c, 4)
    off, hub, hop, dist = kr.arrays()
    cl = O.clean(kr)
    assert cl.total() == kr.total()
    s, t, k = c1_queries(c.n)
    ans = O.query(cl, s, t, k)
    np.savez_compressed(os.path.join(HERE, "c1_rmat12_k4.npz"), off=c.off, col=c.col, w=c.w,
                        lab_off=off, hub=hub, hop=hop, dist=dist, qs=s, qt=t, qk=k, ans=ans)
    out = {}
    for nm, g2 in fixture_graphs():
        c2 = csr_of(g2)
        for K in (1, 2, 3, 4, 5, 8):
            lab = O.clean(O.kround(c2, K))
            o2, h2, p2, d2 = lab.arrays()
            key = f"{nm}|{K}"
            out[key + "|off"] = o2
            out[key + "|hub"] = h2
            out[key + "|hop"] = p2
            out[key + "|dist"] = d2
    np.savez_compressed(os.path.join(HERE, "fixtures_k.npz"), **out)
    for f in ("c1_rmat12_k4.npz", "fixtures_k.npz", "dijkstra_k.npz"):
        with open(os.path.join(HERE, f), "rb") as fh:
            print(f, hashlib.sha256(fh.read()).hexdigest()[:16], os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
