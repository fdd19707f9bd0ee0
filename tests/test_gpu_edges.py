"""GPU edge cases: the exact 32-bit label packing, prune block sizes, batch
graph replays, HBM budget failures, server back-pressure and the label
dump's byte layout.
Every comparison with the oracle is exact (integer work)."""
import numpy as np
import pytest

import gpu4hophl_code_amd as H
import oracle_lib as O

pytestmark = pytest.mark.gpu


def csr_of(g):
    off, col, w = g.csr()
    return O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)


def same_labels(idx, want):
    off, hub, hop, dist = idx.labels()
    woff, whub, whop, wdist = want
    assert np.array_equal(off.astype(np.int64), np.asarray(woff, np.int64))
    assert np.array_equal(hub, whub)
    assert np.array_equal(hop, whop)
    assert np.array_equal(dist, wdist)


def bits(x):
    """Bits of an unsigned field that holds 0..x (at least 1)."""
    b = 1
    while (1 << b) <= x:
        b += 1
    return b


@pytest.fixture(scope="module")
def wide32():
    """R-MAT scale 12, edge factor 8, K=5, weights U[20000, 26000]: 12 hub
    bits + 3 hop bits + 17 distance bits = exactly 32, both for the
    generator's 32-bit pool copy (K * max weight = 130,000 < 2^17) and for
    the clean's 32-bit copy (every label of 4 or more hops is at least
    80,000 >= 2^16, so the largest label distance needs 17 bits) -- the
    packing R-MAT scale 20 at K=5 uses (20 + 3 + 9)."""
    g = H.Graph.rmat(12, 8, 0x5EED0032, 20000, 26000)
    c = csr_of(g)
    K = 5
    kr = O.kround(c, K)
    want = kr.arrays()
    assert int(c.w.max()) * K < (1 << 17) and int(c.w.max()) * K >= (1 << 16)
    assert bits(c.n - 1) + bits(K) + bits(int(c.w.max()) * K) == 32
    assert bits(c.n - 1) + bits(K) + bits(int(want[3].max())) == 32
    return g, c, K, kr, want


@pytest.mark.parametrize("rows", ["1", "2", "0"])
@pytest.mark.parametrize("mode,hub_batch", [("batched", 0), ("batched", 333), ("plain", 0)])
def test_exact_32bit_packing_generation_clean_queries(monkeypatch, wide32, rows, mode, hub_batch):
    """At exactly 32 packed bits the prune streams the 32-bit pool copy
    (prune_label_bytes 4) and the clean streams its 32-bit copy
    (clean_label_bytes 4); labels, cleaned labels and queries equal the
    oracle's."""
    monkeypatch.setenv("HSDL_PRUNE_ROWS", rows)
    g, c, K, kr, want = wide32
    m = H.MODE_PLAIN if mode == "plain" else H.MODE_BATCHED
    idx = H.hop_constrained_two_hop_labels_generation(g, K, mode=m, hub_batch=hub_batch)
    st = idx.stats()
    assert st["prune_label_bytes"] == 4
    same_labels(idx, want)
    idx.clean()
    assert idx.stats()["clean_label_bytes"] == 4
    cl = O.clean(kr)
    same_labels(idx, cl.arrays())
    rng = np.random.default_rng(5)
    s = rng.integers(0, c.n, 20000).astype(np.uint32)
    t = rng.integers(0, c.n, 20000).astype(np.uint32)
    k = rng.integers(-1, K + 2, 20000).astype(np.int8)
    assert np.array_equal(idx.query(s, t, k), O.query(cl, s, t, k.astype(np.int32)))


def test_exact_32bit_packing_clean_of_redundant_labels(wide32):
    """The clean at 32 packed bits on labels it must thin: the reference's
    vertex-group generation (pruning sees only same-group hubs) imported and
    cleaned on the GPU equals the oracle's clean of the same labels."""
    g, c, K, kr, want = wide32
    red = O.kround_grouped(c, K, 256)
    off, hub, hop, dist = red.arrays()
    assert len(hub) > len(want[1])
    idx = H.Index.from_labels(c.n, K, off.astype(np.uint64), hub, hop, dist)
    idx.clean()
    assert idx.stats()["clean_label_bytes"] == 4
    same_labels(idx, O.clean(red).arrays())


@pytest.fixture(scope="module")
def ragged_grid():
    """A 61 x 83 grid (5,063 vertices: no multiple of any vertex block) at
    K=6 with its oracle labels."""
    g = H.Graph.grid(61, 83, 0x5EED0061, 1, 1000)
    return g, O.kround(csr_of(g), 6).arrays()


@pytest.fixture(scope="module")
def heavy13():
    g = H.Graph.rmat(13, 16, 77, 1, 100)
    return g, O.kround(csr_of(g), 4).arrays()


@pytest.mark.parametrize("block", ["1024", "4096"])
@pytest.mark.parametrize("rows", ["1", "2"])
def test_row_block_sizes(monkeypatch, heavy13, ragged_grid, block, rows):
    """HSDL_ROW_BLOCK (vertices per y-block of a row item) at 1024 and 4096,
    beside the default 2048 the other tests use: on R-MAT scale 13 and on a
    grid whose vertex count is no multiple of the block, with row items where
    dense and in every round."""
    monkeypatch.setenv("HSDL_ROW_BLOCK", block)
    monkeypatch.setenv("HSDL_PRUNE_ROWS", rows)
    g, want = heavy13
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4), want)
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4, hub_batch=100), want)
    gg, gwant = ragged_grid
    same_labels(H.hop_constrained_two_hop_labels_generation(gg, 6), gwant)


@pytest.mark.parametrize("graph", ["1", "0"])
def test_batch_graphs_replay_exactly(monkeypatch, heavy13, ragged_grid, graph):
    """Every batch runs as one replay of a captured HIP graph (HSDL_GRAPH=1,
    the default) or as separate launches (0): the same labels on R-MAT scale
    13 and the ragged grid, over many replays (batches of 7 and 100 hubs,
    the autotuned schedule, the plain form), with the host's reads and
    synchronisations between replays, and across re-captures after the
    candidate and frontier lists overflow and grow (HSDL_LIST_INIT)."""
    monkeypatch.setenv("HSDL_GRAPH", graph)
    g, want = heavy13
    for kw in ({}, {"hub_batch": 7}, {"hub_batch": 100}, {"mode": H.MODE_PLAIN}):
        same_labels(H.hop_constrained_two_hop_labels_generation(g, 4, **kw), want)
    gg, gwant = ragged_grid
    same_labels(H.hop_constrained_two_hop_labels_generation(gg, 6, hub_batch=500), gwant)
    monkeypatch.setenv("HSDL_LIST_INIT", "4096")
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4, hub_batch=300), want)


@pytest.mark.parametrize("claim", ["1", "64"])
def test_prune_claim_sizes(monkeypatch, heavy13, ragged_grid, claim):
    """A prune wave claims 1 or 64 candidates at once (default 8), with row
    items and list items: the same labels."""
    monkeypatch.setenv("HSDL_CLAIM_MAX", claim)
    g, want = heavy13
    for rows in ("1", "2", "0"):
        monkeypatch.setenv("HSDL_PRUNE_ROWS", rows)
        same_labels(H.hop_constrained_two_hop_labels_generation(g, 4), want)
    gg, gwant = ragged_grid
    same_labels(H.hop_constrained_two_hop_labels_generation(gg, 6, hub_batch=300), gwant)


def test_hbm_budget_failures_are_clean(golden_dir):
    """A build under a device budget too small for it returns HSDL_ENOMEM
    (at set-up, or when the label pool, arena or lists must grow past it)
    and leaves the graph usable: over a sweep of budgets every build either
    fails with HSDL_ENOMEM or gives the golden labels, both outcomes occur,
    and a build without a budget afterwards is exact."""
    import os
    z = np.load(os.path.join(golden_dir, "c1_rmat12_k4.npz"))
    want = (z["lab_off"], z["hub"], z["hop"], z["dist"])
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    outcomes = []
    # set-up needs about 440 MB here (the dense slots, lists, tile arena and
    # heavy-piece list); budgets just above it fail later, when the label
    # pool must grow
    for mb in [1, 96] + list(range(430, 472, 2)) + [512]:
        try:
            idx = H.hop_constrained_two_hop_labels_generation(
                g, 4, hbm_budget=mb << 20, pool_init_labels=4096 * 16)
        except H.HSDLError as e:
            assert e.code == -2, (mb, e)
            outcomes.append("enomem")
            continue
        same_labels(idx, want)
        outcomes.append("ok")
        idx.free()
    assert "enomem" in outcomes and "ok" in outcomes, outcomes
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4), want)


@pytest.fixture(scope="module")
def c1_index(golden_dir):
    import os
    z = np.load(os.path.join(golden_dir, "c1_rmat12_k4.npz"))
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    idx = H.hop_constrained_two_hop_labels_generation(g, 4, clean=True)
    return idx, z


def test_server_queue_full_and_backlog_larger_than_queue(c1_index):
    """Back-pressure: a submit that does not fit the queue is refused whole
    (HSDL_EBUSY, nothing enqueued); a replay whose due requests exceed the
    queue capacity (all 20,000 arrive at once into a queue of 1,000)
    finishes, feeding the queue as results drain, with every answer equal to
    hsdl_query's."""
    idx, z = c1_index
    n = 4096
    rng = np.random.default_rng(9)
    s = rng.integers(0, n, 20000).astype(np.uint32)
    t = rng.integers(0, n, 20000).astype(np.uint32)
    k = rng.integers(1, 5, 20000).astype(np.int8)
    want = idx.query(s, t, k)
    srv = H.Server(idx, max_batch=256, max_wait_us=50, inflight=2, queue_capacity=1000)
    try:
        assert srv.submit(s[:1001], t[:1001], k[:1001]) is None  # EBUSY
        first = srv.submit(s[:1000], t[:1000], k[:1000])
        assert first == 0
        assert srv.submit(s[:1], t[:1], k[:1]) is None  # full until polled
        srv.drain()
        ids, ans, _ = srv.poll(2000)
        order = np.argsort(ids)
        assert np.array_equal(ids[order], np.arange(1000))
        assert np.array_equal(ans[order], want[:1000])
        got, lat = srv.replay(s, t, k, np.zeros(len(s), np.uint64))
        assert np.array_equal(got, want)
        assert (lat > 0).all()
    finally:
        srv.close()


def test_server_one_slot_many_small_batches(c1_index):
    """Slot starvation: one in-flight slot and batches of 64 -- requests wait
    for the slot, none is lost or answered twice, and the batches number at
    least requests / 64."""
    idx, z = c1_index
    rng = np.random.default_rng(10)
    nq = 30000
    s = rng.integers(0, 4096, nq).astype(np.uint32)
    t = rng.integers(0, 4096, nq).astype(np.uint32)
    k = rng.integers(0, 6, nq).astype(np.int8)
    want = idx.query(s, t, k)
    srv = H.Server(idx, max_batch=64, max_wait_us=0, inflight=1, queue_capacity=4096)
    try:
        gaps = rng.exponential(200.0, nq)  # ~5M q/s offered
        got, _ = srv.replay(s, t, k, np.cumsum(gaps).astype(np.uint64))
        st = srv.stats()
        assert np.array_equal(got, want)
        assert st["completed"] == nq and st["batches"] >= nq // 64
    finally:
        srv.close()


def _dump_bytes(off, hub, hop, dist, par=None):
    """The reference's label dump (Clean_pc/main.cu:11-57) built with
    numpy: size_t n; per vertex size_t count, then count records of four
    int32 {hub, parent, hop, dist}."""
    n = len(off) - 1
    parts = [np.array([n], np.uint64).tobytes()]
    rec = np.zeros((len(hub), 4), np.int32)
    rec[:, 0] = hub
    rec[:, 1] = 0 if par is None else par
    rec[:, 2] = hop
    rec[:, 3] = dist
    for v in range(n):
        parts.append(np.array([off[v + 1] - off[v]], np.uint64).tobytes())
        parts.append(rec[off[v]:off[v + 1]].tobytes())
    return b"".join(parts)


def test_label_dump_byte_layout(tmp_path, c1_index):
    """hsdl_save_labels writes the reference's dump byte for byte (a numpy
    packing of the exported labels, parent 0 without parents), and a dump
    written by numpy in that layout -- records in reverse order, as the
    reference's uncleaned export leaves them -- loads back sorted."""
    idx, z = c1_index
    off, hub, hop, dist = idx.labels()
    off = off.astype(np.int64)
    p = tmp_path / "c1.bin"
    idx.save(str(p))
    assert p.read_bytes() == _dump_bytes(off, hub, hop, dist)
    # reversed per vertex, with parents 0
    rh, rp, rd = hub.copy(), hop.copy(), dist.copy()
    for v in range(len(off) - 1):
        a, b = off[v], off[v + 1]
        rh[a:b], rp[a:b], rd[a:b] = hub[a:b][::-1], hop[a:b][::-1], dist[a:b][::-1]
    q = tmp_path / "rev.bin"
    q.write_bytes(_dump_bytes(off, rh, rp, rd))
    back = H.Index.load(str(q), 4)
    same_labels(back, (off, hub, hop, dist))


def _union_arrays(n, parts):
    """Per-vertex union of several label exports, sorted by (hub, hop, dist)."""
    vs, hs, ps, ds = [], [], [], []
    for off, hub, hop, dist in parts:
        off = off.astype(np.int64)
        vs.append(np.repeat(np.arange(n, dtype=np.int64), np.diff(off)))
        hs.append(hub)
        ps.append(hop)
        ds.append(dist)
    v, hub, hop, dist = (np.concatenate(x) for x in (vs, hs, ps, ds))
    order = np.lexsort((dist, hop, hub, v))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(v, minlength=n))
    return off, hub[order], hop[order], dist[order]


@pytest.mark.parametrize("grouping", ["contiguous", "random"])
def test_group_builds_match_the_oracle_grouping(monkeypatch, grouping):
    """Per-group builds (hsdl_generate_group, the reference's label_gen per
    vertex group): each group's labels and the union over the groups equal
    the oracle's group-restricted K-round set (contiguous groups of 1000
    hubs as the reference's driver cuts them, and random groups), through
    the batched, small-batch, overflow-redo and plain forms; the GPU clean of
    the union gives the oracle's clean of it."""
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    c = csr_of(g)
    K, n = 4, c.n
    rng = np.random.default_rng(12)
    gid = (np.arange(n) // 1000 if grouping == "contiguous" else rng.integers(0, 5, n)).astype(np.int32)
    want = O.kround_gid(c, K, gid)
    parts = []
    for i, grp in enumerate(np.unique(gid)):
        hubs = np.flatnonzero(gid == grp).astype(np.uint32)
        kw = [dict(), dict(hub_batch=7), dict(pool_init_labels=n * 16 + 1, arena_init_words=1),
              dict(mode=H.MODE_PLAIN)][i % 4]
        idx = H.hop_constrained_two_hop_labels_generation(g, K, hubs=hubs, **kw)
        off, hub, hop, dist = idx.labels()
        assert np.isin(hub, hubs).all()
        parts.append((off, hub, hop, dist))
    uoff, uhub, uhop, udist = _union_arrays(n, parts)
    woff, whub, whop, wdist = want.arrays()
    assert np.array_equal(uoff, woff)
    assert np.array_equal(uhub, whub) and np.array_equal(uhop, whop) and np.array_equal(udist, wdist)
    idx = H.Index.from_labels(n, K, uoff.astype(np.uint64), uhub, uhop, udist)
    idx.clean()
    same_labels(idx, O.clean(want).arrays())


def test_group_argument_errors():
    g = H.Graph.rmat(8, 4, 1)
    for bad in ([3, 2], [1, 1], [0, 256], []):
        with pytest.raises(H.HSDLError) as e:
            H.hop_constrained_two_hop_labels_generation(g, 3, hubs=np.array(bad, np.uint32))
        assert e.value.code == -1, bad


@pytest.mark.parametrize("zones,div,slice_", [("0", None, None), ("1", "1", None), ("1", "8", None),
                                              ("1", "1048576", None), ("1", "8", "64")])
@pytest.mark.parametrize("rows,hub_batch", [("1", 0), ("2", 7), ("0", 100)])
def test_hop_zones(monkeypatch, golden_dir, heavy13, ragged_grid, zones, div, slice_, rows, hub_batch):
    """Two-zone label lists (a hop-sorted zone with per-hop counts, then the
    labels appended since; a round's query reads hops below its own and the
    append zone): re-sorted after every batch (divisor 1), at the default 1/8
    threshold, almost never, with lists longer than a 64-label slice left
    unsorted, and switched off -- the labels equal the golden and oracle
    sets in every case."""
    import os
    monkeypatch.setenv("HSDL_HOP_ZONES", zones)
    monkeypatch.setenv("HSDL_PRUNE_ROWS", rows)
    if div:
        monkeypatch.setenv("HSDL_RESORT_DIV", div)
    if slice_:
        monkeypatch.setenv("HSDL_SORT_SLICE", slice_)
    z = np.load(os.path.join(golden_dir, "c1_rmat12_k4.npz"))
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4, hub_batch=hub_batch),
                (z["lab_off"], z["hub"], z["hop"], z["dist"]))
    g, want = heavy13
    same_labels(H.hop_constrained_two_hop_labels_generation(g, 4, hub_batch=hub_batch), want)
    gg, gwant = ragged_grid
    same_labels(H.hop_constrained_two_hop_labels_generation(gg, 6, hub_batch=hub_batch), gwant)


def test_hop_zones_with_overflow_redo_and_32bit(monkeypatch, wide32):
    """Hop zones through the pool and arena overflow path (a batch rolled
    back and run again after earlier batches were re-sorted) and at the
    exact 32-bit packing."""
    monkeypatch.setenv("HSDL_RESORT_DIV", "1")
    g, c, K, kr, want = wide32
    for kw in (dict(hub_batch=50, pool_init_labels=4096 * 16 + 1, arena_init_words=1),
               dict(hub_batch=13), dict()):
        same_labels(H.hop_constrained_two_hop_labels_generation(g, K, **kw), want)


def test_clean_of_a_vertex_with_more_than_65536_labels():
    """A vertex whose list outgrows the LDS keep bits of the sequential clean
    pass (70,010 labels: hop-1 labels of hubs 0..9 and hop-2 labels of 70,000
    further hubs, some dominated through the first ten and some by the same
    hub at a lower hop) is cleaned through the global keep-bit variant, equal
    to the oracle's clean; so is the rest of the index."""
    rng = np.random.default_rng(70)
    W, U = 10, 70000
    n = W + U + 1
    v = n - 1
    offs, hubs, hops, dists = [0], [], [], []
    for x in range(n - 1):
        hs, ps, ds = [], [], []
        if x >= W:  # L(u): hop-1 labels of the ten top hubs, then the self label
            for w in range(W):
                hs.append(w); ps.append(1); ds.append(int(rng.integers(1, 50)))
        hs.append(x); ps.append(0); ds.append(0)
        hubs += hs; hops += ps; dists += ds
        offs.append(offs[-1] + len(hs))
    # L(v): hop-1 labels of the top hubs, hop-2 labels of every other hub
    # (some dominated through a top hub), hop-3 repeats of some (dominated by
    # the same hub's hop-2 label), and the self label
    for w in range(W):
        hubs.append(w); hops.append(1); dists.append(int(rng.integers(1, 50)))
    for u in range(W, W + U):
        d = int(rng.integers(1, 120))
        hubs.append(u); hops.append(2); dists.append(d)
        if u % 7 == 0:
            hubs.append(u); hops.append(3); dists.append(d + int(rng.integers(0, 5)))
    hubs.append(v); hops.append(0); dists.append(0)
    offs.append(len(hubs))
    off = np.array(offs, np.uint64)
    hub = np.array(hubs, np.uint32)
    hop = np.array(hops, np.uint8)
    dist = np.array(dists, np.int32)
    assert off[-1] - off[-2] > 65536
    want = O.clean(O.labels_from_arrays(n, off.astype(np.int64), hub, hop, dist))
    assert want.total() < len(hub)
    idx = H.Index.from_labels(n, 3, off, hub, hop, dist)
    idx.clean()
    assert idx.stats()["clean_passes"] == 2
    same_labels(idx, want.arrays())


@pytest.mark.parametrize("hop_sort", ["0", "1"])
def test_clean_hop_sorted_stream(monkeypatch, wide32, hop_sort):
    """The clean's first pass streams each head's list from a 32-bit copy
    sorted by hop, only up to the hops the group can use
    (HSDL_CLEAN_HOP_SORT=1, the default), or whole in index order (0): on
    redundant labels (vertex groups of 64 hubs) and on a grid at K=8 the
    cleaned labels equal the oracle's."""
    monkeypatch.setenv("HSDL_CLEAN_HOP_SORT", hop_sort)
    g, c, K, kr, want = wide32
    red = O.kround_grouped(c, K, 64)
    off, hub, hop, dist = red.arrays()
    idx = H.Index.from_labels(c.n, K, off.astype(np.uint64), hub, hop, dist)
    idx.clean()
    same_labels(idx, O.clean(red).arrays())
    gg = H.Graph.grid(30, 41, 0x5EED0099, 1, 1000)
    gc = csr_of(gg)
    gred = O.kround_grouped(gc, 8, 100)
    off, hub, hop, dist = gred.arrays()
    idx = H.Index.from_labels(gc.n, 8, off.astype(np.uint64), hub, hop, dist)
    idx.clean()
    same_labels(idx, O.clean(gred).arrays())


@pytest.mark.parametrize("seed", [1, 2])
def test_clean_of_imported_labels_with_hubs_below_their_vertex(seed):
    """Imported label sets may hold hubs ranked below their vertex (the
    reference's multi-threaded 2023WWW generation makes them).  Start from
    the K-round labels of R-MAT scale 9 and add, for random labels (y,h,d) of
    L(s) with y < s, the mirror label (s,h,d) to L(y), and some with a larger
    distance: every added label is redundant, and lists then end in hubs
    above their vertex, which the clean's own-hub lookup in L(u) must pass
    over.  hsdl_clean equals the normative clean reading its input L(u)."""
    g = H.Graph.rmat(9, 8, 5, 1, 20)
    c = csr_of(g)
    K = 4
    off, hub, hop, dist = O.kround(c, K).arrays()
    rng = np.random.default_rng(seed)
    lists = [[(int(hub[i]), int(hop[i]), int(dist[i])) for i in range(off[v], off[v + 1])]
             for v in range(c.n)]
    added = 0
    for s in range(c.n):
        for (y, h, d) in list(lists[s]):
            if y < s and rng.random() < 0.3:
                lists[y].append((s, h, d + int(rng.integers(0, 3))))
                added += 1
    assert added > 100
    lists = [sorted(set(l)) for l in lists]
    # one label per (hub, hop): keep the smallest distance
    lists = [[x for i, x in enumerate(l) if i == 0 or l[i - 1][:2] != x[:2]] for l in lists]
    assert any(l and l[-1][0] > v for v, l in enumerate(lists))
    noff = np.zeros(c.n + 1, np.uint64)
    noff[1:] = np.cumsum([len(l) for l in lists])
    flat = [x for l in lists for x in l]
    nh = np.array([x[0] for x in flat], np.uint32)
    nhop = np.array([x[1] for x in flat], np.uint8)
    nd = np.array([x[2] for x in flat], np.int32)
    idx = H.Index.from_labels(c.n, K, noff, nh, nhop, nd)
    idx.clean()
    want = O.clean(O.labels_from_arrays(c.n, noff.astype(np.int64), nh, nhop, nd), lu_mode=1)
    same_labels(idx, want.arrays())
    assert want.total() == off[-1]
