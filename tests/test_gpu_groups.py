"""CDLP vertex groups on the GPU (hsdl_graph_groups) against the grouping
restatement in tests/group_oracle.py: the same groups in the same order."""
import numpy as np
import pytest

import gpu4hophl_code_amd as H
import group_oracle as G

pytestmark = pytest.mark.gpu


def graphs():
    yield "rmat9", H.Graph.rmat(9, 8, 2, 1, 10)
    yield "rmat11", H.Graph.rmat(11, 4, 5, 1, 10)
    yield "grid20", H.Graph.grid(20, 20, 3, 1, 10)
    # a star plus isolated vertices and a path
    n = 60
    u = [0] * 30 + list(range(40, 49))
    v = list(range(1, 31)) + list(range(41, 50))
    yield "star", H.Graph.from_edges(n, np.array(u, np.uint32), np.array(v, np.uint32),
                                     np.ones(len(u), np.int32))


@pytest.mark.parametrize("max_size", [1, 6, 40, 100000])
@pytest.mark.parametrize("rounds", [1, 4, 100])
def test_cdlp_groups_equal_the_restatement(max_size, rounds):
    for name, g in graphs():
        off, col, _ = g.csr()
        want = G.cdlp_groups([int(x) for x in off], [int(x) for x in col], max_size, rounds)
        gid, ng = g.groups(max_size, "cdlp", rounds=rounds, device=0)
        print(f"cdlp {name} max_size {max_size} rounds {rounds}: {ng} groups", flush=True)
        assert ng == len(want), name
        assert list(gid) == G.group_of(want, g.n), name
