"""hsdl_query_noalloc: the query entry point that allocates nothing (its
requests and answers are staged in the library's fixed device array) gives
hsdl_query's answers, call after call."""
import os

import numpy as np
import pytest

import gpu4hophl_code_amd as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_index():
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    return H.hop_constrained_two_hop_labels_generation(g, 4, clean=True)


def queries(n, count, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, count).astype(np.uint32)
    t = rng.integers(0, n, count).astype(np.uint32)
    k = rng.integers(-1, 7, count).astype(np.int8)
    return s, t, k


def test_noalloc_equals_query_on_consecutive_calls(c1_index):
    """Two calls in a row, the second smaller (the array keeps the first
    call's tail), then the largest allowed batch, then a small one again:
    every answer equals hsdl_query's."""
    idx = c1_index
    for count, seed in ((200000, 1), (777, 2), (H.NOALLOC_MAX_QUERIES, 3), (5, 4)):
        s, t, k = queries(4096, count, seed)
        want = idx.query(s, t, k)
        got = idx.query_noalloc(s, t, k)
        assert np.array_equal(got, want), count
    # the same small batch twice gives the same answers
    s, t, k = queries(4096, 1000, 5)
    assert np.array_equal(idx.query_noalloc(s, t, k), idx.query_noalloc(s, t, k))


def test_noalloc_limits_and_errors(c1_index):
    idx = c1_index
    n = H.NOALLOC_MAX_QUERIES + 1
    s = np.zeros(n, np.uint32)
    k = np.ones(n, np.int8)
    with pytest.raises(H.HSDLError) as e:
        idx.query_noalloc(s, s, k)
    assert e.value.code == -1
    with pytest.raises(H.HSDLError) as e:
        idx.query_noalloc(np.array([4096], np.uint32), np.array([0], np.uint32), np.array([1], np.int8))
    assert e.value.code == -1
    assert len(idx.query_noalloc(s[:0], s[:0], k[:0])) == 0
