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
import oracle_lib as O  # noqa: E402

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


def main():
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
    for f in ("c1_rmat12_k4.npz", "fixtures_k.npz"):
        with open(os.path.join(HERE, f), "rb") as fh:
            print(f, hashlib.sha256(fh.read()).hexdigest()[:16], os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
