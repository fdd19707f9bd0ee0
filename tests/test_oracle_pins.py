"""Pins the CPU oracle (oracle/hsdl_oracle.cpp) against checks that do not
reuse its own formulas:

* brute-force canonical labels from the closed form (hop-bounded distances
  only) on hundreds of tiny seeded graphs, including small weight ranges that
  force ties;
* the reference's two generators (Dijkstra of CPU_HopHL and the K-round twin)
  agreeing after cleaning, under both readings of L(u) in the clean;
* label queries equal hop-bounded distances for every (s,t,k) of tiny graphs;
* hop-bounded Bellman-Ford equal to min-plus matrix powers computed in numpy;
* the reference's own fixture graphs (|V|=/Edge format) loaded with its
  last-wins duplicate rule;
* planted mistakes (the oracle's `mutation` codes) each fail one of the above.
"""
import itertools
import os

import numpy as np
import pytest

import oracle_lib as O

INF = 2147483647


def minplus_hop_distances(g, K):
    """D[h][a][b] = shortest walk with <= h edges, by min-plus products."""
    n = g.n
    W = np.full((n, n), np.inf)
    for v in range(n):
        for e in range(g.off[v], g.off[v + 1]):
            W[v, g.col[e]] = min(W[v, g.col[e]], g.w[e])
    D = [np.where(np.eye(n, dtype=bool), 0.0, np.inf)]
    for _ in range(K):
        prev = D[-1]
        nxt = np.minimum(prev, np.min(prev[:, :, None] + W[None, :, :], axis=1))
        D.append(nxt)
    return D


def tiny_graphs(seed, count, nmax=24):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        n = int(rng.integers(2, nmax))
        p = float(rng.uniform(0.08, 0.6))
        wmax = int(rng.choice([1, 2, 3, 10, 100]))
        K = int(rng.integers(1, 6))
        yield O.random_graph(rng, n, p, wmax), K


def test_generators_clean_to_bruteforce_canonical():
    for i, (g, K) in enumerate(tiny_graphs(11, 300)):
        want = O.canonical(g, K).as_lists()
        kr = O.kround(g, K)
        dj = O.dijkstra(g, K)
        # single-threaded rank-ordered generation is already canonical
        assert kr.as_lists() == want and dj.as_lists() == want
        # group-restricted pruning (the reference's vertex groups) leaves
        # redundant labels; the clean must remove exactly those
        gr = O.kround_grouped(g, K, 1 + i % 5)
        for lab in (kr, dj, gr):
            for mode in (0, 1):
                assert O.clean(lab, mode).as_lists() == want


def test_grouped_generation_is_redundant_superset():
    extra = 0
    for i, (g, K) in enumerate(tiny_graphs(15, 100)):
        want = set((v, l) for v, lst in enumerate(O.canonical(g, K).as_lists()) for l in lst)
        got = set((v, l) for v, lst in enumerate(O.kround_grouped(g, K, 2).as_lists()) for l in lst)
        assert want <= got
        extra += len(got - want)
    assert extra > 0


def test_kround_invariants():
    for g, K in tiny_graphs(12, 100, nmax=30):
        L = O.kround(g, K).as_lists()
        for v, lst in enumerate(L):
            assert (v, 0, 0) in lst
            assert all(h <= v for h, _, _ in lst)
            assert all(hop <= K for _, hop, _ in lst)
            by_hub = {}
            for hub, hop, d in lst:
                by_hub.setdefault(hub, []).append((hop, d))
            for hub, hd in by_hub.items():
                hd.sort()
                dists = [d for _, d in hd]
                assert all(a > b for a, b in zip(dists, dists[1:])), (v, hub, hd)


def test_queries_equal_hop_bounded_distances():
    for g, K in tiny_graphs(13, 60, nmax=18):
        L = O.clean(O.kround(g, K))
        D = minplus_hop_distances(g, K)
        n = g.n
        s, t, k = [], [], []
        for a, b in itertools.product(range(n), range(n)):
            for kk in range(-1, K + 1):
                s.append(a), t.append(b), k.append(kk)
        got = O.query(L, s, t, k)
        for i in range(len(s)):
            kk = k[i]
            if kk < 0:
                want = INF
            elif s[i] == t[i]:
                want = 0
            elif kk == 0:
                want = INF
            else:
                d = D[kk][s[i], t[i]]
                want = INF if np.isinf(d) else int(d)
            assert got[i] == want, (s[i], t[i], kk)


def test_bf_equals_minplus():
    for g, K in tiny_graphs(14, 40, nmax=20):
        D = minplus_hop_distances(g, K)
        for src in range(g.n):
            for k in range(K + 1):
                got = O.bf(g, src, k)
                want = np.where(np.isinf(D[k][src]), INF, D[k][src]).astype(np.int64)
                assert np.array_equal(got.astype(np.int64), want)


def test_2023www_fixture_by_hand(golden_dir):
    # V=10, 13 edges: small enough to check some distances by hand
    g = O.load_txt(os.path.join(golden_dir, "simple_iterative_tests_2023www.txt"))
    assert g.n == 10 and len(g.col) == 26
    # neighbours of 0 in id order with their weights (file lines 4-8)
    nb = list(zip(g.col[g.off[0]:g.off[1]].tolist(), g.w[g.off[0]:g.off[1]].tolist()))
    assert nb == [(1, 2), (3, 1), (5, 3), (7, 4), (9, 1)]
    L = O.clean(O.kround(g, 3))
    # 0 -> 6: 0-3-4-6 = 1+1+1 = 3 (3 hops); 0-1-2-6 = 2+1+1 = 4 (3 hops);
    # with 2 hops: 0-3-2-6? 3-2 is 2 -> 0-3-2 = 3 then -6 needs a 3rd hop; none.
    assert O.query(L, [0], [6], [3])[0] == 3
    assert O.query(L, [0], [6], [2])[0] == INF
    # 7 -> 2: 7-1-2 = 1+1 = 2
    assert O.query(L, [7], [2], [2])[0] == 2


def test_fixture_last_wins_and_self_loop(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("|V|= 4\n|E|= 4\n\nEdge 0 1 5\nEdge 1 0 7\nEdge 2 2 3\nEdge 2 3 4.9\n\nEOF\n")
    g = O.load_txt(str(p))
    adj = {v: list(zip(g.col[g.off[v]:g.off[v + 1]].tolist(), g.w[g.off[v]:g.off[v + 1]].tolist()))
           for v in range(4)}
    assert adj[0] == [(1, 7)] and adj[1] == [(0, 7)]
    assert adj[2] == [(2, 3), (3, 4)]  # self loop once, 4.9 truncates to 4
    assert adj[3] == [(2, 4)]


def test_simple_iterative_fixture_queries_vs_bf(golden_dir):
    g = O.load_txt(os.path.join(golden_dir, "simple_iterative_tests.txt"), reorder=True)
    assert g.n == 200
    for K in (2, 4):
        L = O.clean(O.kround(g, K))
        assert O.clean(O.dijkstra(g, K)).as_lists() == L.as_lists()
        rng = np.random.default_rng(K)
        for src in rng.integers(0, g.n, 12):
            for k in range(1, K + 1):
                want = O.bf(g, int(src), k)
                got = O.query(L, np.full(g.n, src), np.arange(g.n), np.full(g.n, k))
                assert np.array_equal(got, want)


MUTATIONS = {
    1: "hop budget < instead of <= (pruning, clean and query)",
    2: "dedupe keeps the first distance instead of the minimum",
    3: "accept on <= instead of <",
    4: "clean keeps on >= instead of >",
    5: "Bellman-Ford one round short",
}


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_planted_mistakes_are_caught(mutation):
    """Each planted mistake must break some pin on a handful of tiny graphs."""
    caught = False
    for g, K in tiny_graphs(100 + mutation, 80, nmax=16):
        want = O.canonical(g, K).as_lists()
        if mutation == 5:
            D = minplus_hop_distances(g, K)
            for src in range(g.n):
                got = O.bf(g, src, K, mutation=5)
                ref = np.where(np.isinf(D[K][src]), INF, D[K][src]).astype(np.int64)
                if not np.array_equal(got.astype(np.int64), ref):
                    caught = True
            if caught:
                break
            continue
        if mutation == 3:
            # the clean repairs equal-distance duplicates, so this one shows
            # in the uncleaned invariant (dist strictly decreasing per hub)
            for lst in O.kround(g, K, mutation=3).as_lists():
                seen = {}
                for hub, hop, d in lst:
                    if hub in seen and seen[hub] <= d:
                        caught = True
                    seen[hub] = d
            if caught:
                break
            continue
        if mutation == 4:
            got = O.clean(O.kround_grouped(g, K, 2), mutation=4).as_lists()
            if got != want:
                caught = True
                break
            continue
        m_gen = mutation if mutation in (1, 2) else 0
        m_clean = mutation if mutation == 1 else 0
        got = O.clean(O.kround(g, K, mutation=m_gen), mutation=m_clean).as_lists()
        if got != want:
            caught = True
            break
        if mutation == 1:
            L = O.clean(O.kround(g, K))
            q = O.query(L, [0] * g.n, list(range(g.n)), [K] * g.n, mutation=1)
            if not np.array_equal(q, O.bf(g, 0, K)):
                caught = True
                break
    assert caught, MUTATIONS[mutation]
