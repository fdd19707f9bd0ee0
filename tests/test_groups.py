"""Vertex groups (hsdl_graph_groups) on the CPU: the id-order and shuffled
groupings, argument errors, and the pins of the grouping restatement in
tests/group_oracle.py that the GPU tests compare CDLP with."""
import numpy as np
import pytest

import gpu4hophl_code_amd as H
import group_oracle as G


def csr_lists(g):
    off, col, _ = g.csr()
    return [int(x) for x in off], [int(x) for x in col]


def check_partition(gid, ng, n, max_size):
    assert gid.shape == (n,) and ng >= 1
    cnt = np.bincount(gid, minlength=ng)
    assert len(cnt) == ng and cnt.min() >= 1 and cnt.max() <= max_size


@pytest.mark.parametrize("max_size", [1, 7, 100, 5000])
def test_contiguous_groups_are_id_ranges(max_size):
    g = H.Graph.rmat(10, 8, 11, 1, 10)
    gid, ng = g.groups(max_size, "contiguous")
    check_partition(gid, ng, g.n, max_size)
    assert np.array_equal(gid, np.arange(g.n) // max_size)


def test_random_groups_are_a_seeded_shuffle():
    g = H.Graph.rmat(10, 8, 11, 1, 10)
    a, na = g.groups(37, "random", seed=5)
    b, nb = g.groups(37, "random", seed=5)
    c, _ = g.groups(37, "random", seed=6)
    check_partition(a, na, g.n, 37)
    assert na == -(-g.n // 37) and np.array_equal(a, b) and not np.array_equal(a, c)
    # every group but the last is full
    assert (np.bincount(a)[:-1] == 37).all()


def test_group_argument_errors_and_no_cpu_fallback():
    g = H.Graph.rmat(6, 4, 1, 1, 10)
    with pytest.raises(H.HSDLError) as e:
        g.groups(0, "contiguous")
    assert e.value.code == -1  # HSDL_EINVAL
    out = np.zeros(g.n, np.uint32)
    import ctypes
    cnt = ctypes.c_uint32()
    assert H.lib().hsdl_graph_groups(g.h, 9, 4, 1, 0, -1, out, ctypes.byref(cnt)) == -1
    with pytest.raises(H.HSDLError) as e:
        g.groups(8, "cdlp")  # label propagation runs on the GPU only
    assert e.value.code == -8  # HSDL_ENODEV


def test_merge_rule_by_hand():
    # sizes 5 4 3 3 1, max 8: 5 + 3 (id 3, the largest of size <= 3) = 8,
    # final; 4 takes the group ordered below itself (3, id 2) = 7, then 1
    groups = [[0] * 5, [1] * 4, [2] * 3, [3] * 3, [4]]
    out = G.merge(groups, 8)
    assert out == [[0] * 5 + [3] * 3, [1] * 4 + [2] * 3 + [4]]
    # nothing fits beside the largest: every group stays alone
    assert G.merge([[0] * 6, [1] * 5], 8) == [[0] * 6, [1] * 5]


def test_propagation_finds_disjoint_cliques():
    """Cliques of at most max_size vertices, joined by nothing: each ends as
    one community (never split), whatever the merge then combines."""
    rng = np.random.default_rng(3)
    for trial in range(20):
        sizes = rng.integers(1, 9, rng.integers(2, 12))
        perm = rng.permutation(int(sizes.sum()))
        n = len(perm)
        adj = [[] for _ in range(n)]
        start = 0
        cliques = []
        for s in sizes:
            mem = [int(x) for x in perm[start:start + s]]
            cliques.append(mem)
            for a in mem:
                adj[a] += [b for b in mem if b != a]
            start += s
        off = [0]
        col = []
        for a in range(n):
            col += sorted(adj[a])
            off.append(len(col))
        label = G.propagate(off, col, 8, 50)
        for mem in cliques:
            assert len({label[v] for v in mem}) == 1, (trial, mem, [label[v] for v in mem])
        groups = G.cdlp_groups(off, col, 8, 50)
        gid = G.group_of(groups, n)
        for mem in cliques:
            assert len({gid[v] for v in mem}) == 1
        assert sorted(v for gr in groups for v in gr) == list(range(n))
        assert max(len(gr) for gr in groups) <= 8
        # ordered by total degree, largest first
        deg = [sum(off[v + 1] - off[v] for v in gr) for gr in groups]
        assert deg == sorted(deg, reverse=True)


def test_propagation_respects_the_size_cap():
    g = H.Graph.rmat(9, 8, 2, 1, 10)
    off, col = csr_lists(g)
    for cap in (3, 20):
        groups = G.cdlp_groups(off, col, cap, 30)
        assert max(len(x) for x in groups) <= cap
        assert sorted(v for x in groups for v in x) == list(range(g.n))
