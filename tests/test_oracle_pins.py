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


# ---------------------------------------------------------------- parents
def labels_with_parents(L):
    """Per-vertex lists of (hub, hop, dist, parent) from an oracle label set."""
    off, hub, hop, dist = L.arrays()
    par = L.parents()
    return [list(zip(hub[off[v]:off[v + 1]].tolist(), hop[off[v]:off[v + 1]].tolist(),
                     dist[off[v]:off[v + 1]].tolist(), par[off[v]:off[v + 1]].tolist()))
            for v in range(L.n)]


def parent_rule(g, lists):
    """The parent each K-round label must carry, derived from the finished
    label set alone: the self label's is the vertex; for (hub, h, d) at y with
    h >= 1 it is the smallest neighbour x holding (hub, h-1, d - w(x, y)) --
    the frontier vertices of round h that reach y at distance d."""
    have = [set((hub, hop, d) for hub, hop, d, _ in lst) for lst in lists]
    want = []
    for y, lst in enumerate(lists):
        row = []
        for hub, hop, d, _ in lst:
            if hop == 0:
                row.append(y)
                continue
            xs = [int(g.col[e]) for e in range(g.off[y], g.off[y + 1])
                  if (hub, hop - 1, d - int(g.w[e])) in have[int(g.col[e])]]
            row.append(min(xs) if xs else None)
        want.append(row)
    return want


def test_kround_parents_follow_the_min_id_rule():
    """The oracle tracks parents while it expands (the smallest frontier
    vertex at the minimum); the rule above recovers them from the labels.
    Weight ranges of 1-3 force equal-distance ties.  The clean keeps each
    kept label's parent."""
    ties = 0
    for g, K in tiny_graphs(21, 200, nmax=28):
        L = O.kround(g, K)
        lists = labels_with_parents(L)
        want = parent_rule(g, lists)
        for y, lst in enumerate(lists):
            got = [p for _, _, _, p in lst]
            assert got == want[y], (y, lst, want[y])
            for hub, hop, d, _ in lst:
                if hop:
                    n_at_min = sum(1 for e in range(g.off[y], g.off[y + 1])
                                   if (hub, hop - 1, d - int(g.w[e])) in
                                   set((a, b, c) for a, b, c, _ in lists[int(g.col[e])]))
                    ties += n_at_min > 1
        by_key = {(v, hub, hop): p for v, lst in enumerate(lists) for hub, hop, _, p in lst}
        for v, lst in enumerate(labels_with_parents(O.clean(L))):
            for hub, hop, _, p in lst:
                assert by_key[(v, hub, hop)] == p
    assert ties > 0


# ------------------------------------------------------------------ paths
def path_by_argmin(lists, s, t, k):
    """hop_constrained_extract_shortest_path
    (CPU_HopHL/include/HBPLL/hop_constrained_two_hop_labels.h:224-319) with
    each step's pair chosen as the lexicographic minimum of (dist sum, hub,
    position in L(s), position in L(t)) over all pairs with hop sum <= k --
    the pair the reference's merge meets first among those at the minimum."""
    out = []
    while s != t:
        best = None
        for i, (ha, pa, da, fa) in enumerate(lists[s]):
            for j, (hb, pb, db, fb) in enumerate(lists[t]):
                if ha == hb and pa + pb <= k:
                    key = (da + db, ha, i, j)
                    if best is None or key < best[0]:
                        best = (key, fa, fb)
        if best is None:
            break
        _, p1, p2 = best
        if p1 == s and p2 == t:
            break
        if s != p1:
            out.append((s, p1))
            s, k = p1, k - 1
        if t != p2:
            out.append((t, p2))
            t, k = p2, k - 1
    return out


def walk_weight(g, s, t, edges):
    """Total weight of the path's edges if they form a walk from s to t (the
    s side's edges extend s, the t side's extend t, and the two ends meet),
    else None."""
    cs, ct, total = s, t, 0
    for a, b in edges:
        nb = g.col[g.off[a]:g.off[a + 1]]
        j = np.searchsorted(nb, b)
        if j >= len(nb) or nb[j] != b:
            return None
        total += int(g.w[g.off[a] + j])
        if a == cs:
            cs = b
        elif a == ct:
            ct = b
        else:
            return None
    return total if cs == ct else None


def test_paths_are_shortest_walks_and_follow_the_merge_order():
    """On every (s, t, k) of tiny graphs (cleaned K-round labels): the path
    is a walk from s to t whose weight is the hop-bounded distance and whose
    length is at most k (empty when s == t or t is out of reach), and it is
    the one the declarative argmin above picks."""
    checked = 0
    for g, K in tiny_graphs(23, 80, nmax=16):
        L = O.clean(O.kround(g, K))
        lists = labels_with_parents(L)
        D = minplus_hop_distances(g, K)
        n = g.n
        s, t, k = zip(*[(a, b, kk) for a in range(n) for b in range(n) for kk in range(0, K + 1)])
        edges, counts = O.paths(L, s, t, k, 2 * K + 2)
        for i in range(len(s)):
            got = [tuple(e) for e in edges[i, :counts[i]].tolist()]
            assert got == path_by_argmin(lists, s[i], t[i], k[i]), (s[i], t[i], k[i])
            d = D[k[i]][s[i], t[i]]
            if s[i] == t[i] or np.isinf(d):
                assert got == []
                continue
            assert len(got) <= k[i]
            assert walk_weight(g, s[i], t[i], got) == int(d), (s[i], t[i], k[i], got)
            checked += 1
    assert checked > 1000


def test_planted_parent_and_path_mistakes_are_caught():
    """Largest parent id instead of the smallest (mutation 7) breaks the
    parent rule; the last pair at the minimum instead of the first
    (mutation 6) breaks the merge-order check."""
    caught7 = caught6 = False
    for g, K in tiny_graphs(31, 60, nmax=20):
        L7 = O.kround(g, K, mutation=7)
        lists = labels_with_parents(L7)
        want = parent_rule(g, lists)
        if any([p for _, _, _, p in lst] != want[y] for y, lst in enumerate(lists)):
            caught7 = True
        L = O.clean(O.kround(g, K))
        lists = labels_with_parents(L)
        n = g.n
        s, t = zip(*[(a, b) for a in range(n) for b in range(n)])
        k = [K] * len(s)
        edges, counts = O.paths(L, s, t, k, 2 * K + 2, mutation=6)
        for i in range(len(s)):
            got = [tuple(e) for e in edges[i, :counts[i]].tolist()]
            if got != path_by_argmin(lists, s[i], t[i], k[i]):
                caught6 = True
                break
        if caught6 and caught7:
            break
    assert caught7 and caught6


# ---------------------------------------------------------------- 2023WWW
def test_2023www_labels_are_first_hop_exact_distances_and_clean_to_canonical():
    """The 2023WWW search (no rank pruning, one distance-so-far per vertex
    across hops): every label (hub, h, d) of u has d equal to the <= h-edge
    distance from hub to u and strictly below the <= h-1-edge one (h is the
    first hop count reaching d); the clean of its labels is the brute-force
    canonical labeling.  Without rank pruning the queries still keep every
    hub at or above its vertex (hub <= u): for u < hub the pair is already
    covered by hub u's own labels.  Run in hub order it yields exactly the
    K-round set.  Mutation 8 (a neighbour pushed without improvement) breaks
    the first property."""
    for g, K in tiny_graphs(41, 200, nmax=22):
        W = O.www2023(g, K)
        D = minplus_hop_distances(g, K)
        for u, lst in enumerate(W.as_lists()):
            for hub, h, d in lst:
                assert d == D[h][hub, u], (u, hub, h, d)
                assert h == 0 or d < D[h - 1][hub, u], (u, hub, h, d)
                assert hub <= u
        want = O.canonical(g, K).as_lists()
        assert O.clean(W, 0).as_lists() == want and O.clean(W, 1).as_lists() == want
        assert W.as_lists() == O.kround(g, K).as_lists()
    caught = False
    for g, K in tiny_graphs(43, 80, nmax=18):
        D = minplus_hop_distances(g, K)
        for u, lst in enumerate(O.www2023(g, K, mutation=8).as_lists()):
            if any(h > 0 and d >= D[h - 1][hub, u] for hub, h, d in lst):
                caught = True
        if caught:
            break
    assert caught


def test_group_id_generation_pins():
    """The oracle's vertex-group generation with arbitrary group ids
    (orc_kround_gid; the reference's per-group label_gen, whose pruning sees
    only labels of the group's hubs):

    * one group for all vertices is the plain K-round set;
    * groups v // G are the contiguous groups of orc_kround_grouped;
    * singleton groups leave each hub only its own labels as witnesses, so
      the set is the closed form (s, h, D_h(y)) for every y >= s with
      D_h(y) < D_{h-1}(y), D_h the shortest <= h-edge walk from s through
      vertices >= s (numpy min-plus on that subgraph);
    * any grouping cleans to the brute-force canonical labels."""
    rng = np.random.default_rng(77)
    for i, (g, K) in enumerate(tiny_graphs(31, 80)):
        n = g.n
        plain = O.kround(g, K).as_lists()
        assert O.kround_gid(g, K, np.zeros(n, np.int32)).as_lists() == plain
        G = 1 + i % 4
        assert (O.kround_gid(g, K, np.arange(n, dtype=np.int32) // G).as_lists()
                == O.kround_grouped(g, K, G).as_lists())
        single = O.kround_gid(g, K, np.arange(n, dtype=np.int32)).as_lists()
        W = np.full((n, n), np.inf)
        for v in range(n):
            for e in range(g.off[v], g.off[v + 1]):
                W[v, g.col[e]] = min(W[v, g.col[e]], g.w[e])
        want = [[] for _ in range(n)]
        for s in range(n):
            sub = W[s:, s:]
            D = np.full(n - s, np.inf)
            D[0] = 0.0
            want[s].append((s, 0, 0))
            for h in range(1, K + 1):
                nd = np.minimum(D, np.min(D[:, None] + sub, axis=0))
                for j in np.flatnonzero(nd < D):
                    if j > 0:
                        want[s + j].append((s, h, int(nd[j])))
                D = nd
        want = [sorted(x) for x in want]
        assert single == want, i
        gid = rng.integers(0, 3, n).astype(np.int32)
        canon = O.canonical(g, K).as_lists()
        assert O.clean(O.kround_gid(g, K, gid)).as_lists() == canon


def test_replay_and_clean_checkers_have_teeth():
    """The full-size checkers (orc_replay_hubs: a hub's K rounds recomputed
    from a finished label set; orc_clean_check: the normative clean of
    sampled vertices) accept the oracle's own sets and catch planted
    mistakes: a changed distance, a dropped label, an extra label, and an
    uncleaned list passed off as cleaned."""
    rng = np.random.default_rng(99)
    caught = {"dist": 0, "drop": 0, "extra": 0, "clean": 0}
    for g, K in tiny_graphs(41, 60, nmax=40):
        lab = O.kround(g, K)
        off, hub, hop, dist = lab.arrays()
        hubs = np.arange(g.n, dtype=np.int32)
        assert not O.replay_hubs(g, K, (off, hub, hop, dist), hubs).any()
        nonself = np.flatnonzero(hop > 0)
        if len(nonself):
            j = int(rng.choice(nonself))
            d2 = dist.copy()
            d2[j] += 1
            bad = O.replay_hubs(g, K, (off, hub, hop, d2), hubs)
            caught["dist"] += int(bad[hub[j]] != 0)
            keep = np.ones(len(hub), bool)
            keep[j] = False
            v = np.searchsorted(off, j, side="right") - 1
            off2 = off.copy()
            off2[v + 1:] -= 1
            bad = O.replay_hubs(g, K, (off2, hub[keep], hop[keep], dist[keep]), hubs)
            caught["drop"] += int(bad[hub[j]] != 0)
        # an extra label (hub 0, hop K) appended at the end of the last vertex
        # when it has no hub-0 hop-K label
        v = g.n - 1
        if v > 0 and not ((hub[off[v]:off[v + 1]] == 0) & (hop[off[v]:off[v + 1]] == K)).any():
            ins = off[v + 1]
            pos = off[v] + np.searchsorted(hub[off[v]:off[v + 1]], 1)  # after hub 0's labels
            hub3 = np.insert(hub, pos, 0)
            hop3 = np.insert(hop, pos, K)
            dist3 = np.insert(dist, pos, 1)
            off3 = off.copy()
            off3[v + 1:] += 1
            bad = O.replay_hubs(g, K, (off3, hub3, hop3, dist3), hubs)
            caught["extra"] += int(bad[0] != 0)
            assert ins >= pos
        red = O.kround_grouped(g, K, 2)
        cl = O.clean(red)
        verts = np.arange(g.n, dtype=np.int32)
        assert not O.clean_check(red.arrays(), cl.arrays(), verts).any()
        if red.total() > cl.total():
            caught["clean"] += int(O.clean_check(red.arrays(), red.arrays(), verts).any())
    assert all(c > 10 for c in caught.values()), caught


def test_label_digest_python_and_c_agree():
    """The label-set digest of the full-size fixtures has two
    implementations (numpy in tests/label_digest.py, C in the oracle library
    for billions of labels); they agree, and a changed, moved or missing
    label changes the digest of its block only."""
    import label_digest as D
    rng = np.random.default_rng(5)
    for g, K in tiny_graphs(51, 20, nmax=60):
        off, hub, hop, dist = O.kround(g, K).arrays()
        for block in (1, 7, 1024):
            c1, d1 = D.block_digests(off, hub, hop, dist, block=block, chunk=97)
            c2, d2 = O.block_digests((off, hub, hop, dist), block=block)
            assert np.array_equal(c1, c2) and np.array_equal(d1, d2)
        j = int(rng.integers(0, len(hub)))
        v = np.searchsorted(off, j, side="right") - 1
        d3 = dist.copy()
        d3[j] ^= 1
        _, dd = O.block_digests((off, hub, hop, d3), block=1)
        _, d0 = O.block_digests((off, hub, hop, dist), block=1)
        assert (dd != d0).sum() == 1 and dd[v] != d0[v]
        if off[v + 1] - off[v] >= 2:  # two labels of a vertex swapped
            a, b = off[v], off[v] + 1
            h4 = hub.copy(); h4[[a, b]] = h4[[b, a]]
            p4 = hop.copy(); p4[[a, b]] = p4[[b, a]]
            x4 = dist.copy(); x4[[a, b]] = x4[[b, a]]
            _, d4 = O.block_digests((off, h4, p4, x4), block=1)
            assert d4[v] != d0[v]


def test_threaded_dijkstra_timer_counts():
    """The time-boxed, threaded CPU_HopHL and 2023WWW restatements (the CPU
    baselines' timers) finish every hub without a limit, and on one thread
    make exactly the sequential generators' labels; with a limit they stop
    early with fewer hubs."""
    for g, K in tiny_graphs(61, 30, nmax=40):
        hubs, labels, _ = O.dijkstra_timed(g, K, threads=1)
        assert hubs == g.n and labels == O.dijkstra(g, K).total()
        hubs, labels, _ = O.dijkstra_timed(g, K, threads=4)
        assert hubs == g.n and labels >= O.canonical(g, K).total()
    for g, K in tiny_graphs(62, 30, nmax=40):
        hubs, labels, _ = O.www2023_timed(g, K, threads=1)
        assert hubs == g.n and labels == O.www2023(g, K).total()
    import gpu4hophl_code_amd as H
    r = H.Graph.rmat(14, 16, 3, 1, 100)
    off, col, w = r.csr()
    c = O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)
    hubs, labels, sec = O.dijkstra_timed(c, 5, threads=2, seconds=0.05)
    assert 0 < hubs < c.n and labels > 0
    hubs, labels, sec = O.www2023_timed(c, 5, threads=2, seconds=0.05)
    assert 0 < hubs < c.n and labels > 0
