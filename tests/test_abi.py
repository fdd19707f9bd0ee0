"""C ABI and host-side graph logic (CPU only: no compute call needs a GPU)."""
import os
import re

import numpy as np
import pytest

import gpu4hophl_code_amd as H
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "hsdl.h")).read()
    return sorted(set(re.findall(r"\b(hsdl_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = H.lib()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(H.EXPORTED) == decl


def test_library_is_built_from_the_tree_sources():
    mk = open(os.path.join(ROOT, "gpu4hophl_code_amd", "csrc", "Makefile")).read()
    lists = {k: re.search(rf"^{k} := (.*)$", mk, re.M).group(1).split()
             for k in ("HIP_SRCS", "CPP_SRCS", "DEPS")}
    assert H.HASHED_SOURCES == lists["HIP_SRCS"] + lists["CPP_SRCS"] + lists["DEPS"]
    assert re.fullmatch(r"[0-9a-f]{64}", H.build_id())
    assert H.build_id() == H.source_hash(), "libhsdl.so is older than its sources"


def test_strerror_codes():
    L = H.lib()
    assert L.hsdl_strerror(0) == b"ok"
    for code in range(-1, -10, -1):
        assert L.hsdl_strerror(code) != b"unknown error"


def csr_from(g):
    off, col, w = g.csr()
    return off.astype(np.int64), col.astype(np.int32), w


@pytest.mark.parametrize("name", ["simple_iterative_tests.txt", "simple_iterative_tests_2023www.txt",
                                  "euroroad2.txt"])
def test_txt_loader_matches_reference_rules(golden_dir, name):
    p = os.path.join(golden_dir, name)
    g = H.Graph.load_txt(p)
    og = O.load_txt(p)
    off, col, w = csr_from(g)
    assert np.array_equal(off, og.off) and np.array_equal(col, og.col) and np.array_equal(w, og.w)
    # degree renaming matches the oracle's restatement
    n2o = g.reorder_by_degree()
    og2, on2o = O.reorder(og)
    off, col, w = csr_from(g)
    assert np.array_equal(n2o.astype(np.int32), on2o)
    assert np.array_equal(off, og2.off) and np.array_equal(col, og2.col) and np.array_equal(w, og2.w)


def test_from_edges_last_wins_and_self_loops(tmp_path):
    rng = np.random.default_rng(3)
    n, m = 50, 400
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    w = rng.integers(1, 20, m)
    p = tmp_path / "multi.txt"
    with open(p, "w") as f:
        f.write(f"|V|= {n}\n|E|= {m}\n\n")
        for a, b, c in zip(u, v, w):
            f.write(f"Edge {a} {b} {c}\n")
    og = O.load_txt(str(p))
    g = H.Graph.from_edges(n, u, v, w)
    off, col, ww = csr_from(g)
    assert np.array_equal(off, og.off) and np.array_equal(col, og.col) and np.array_equal(ww, og.w)
    g2 = H.Graph.load_txt(str(p))
    assert all(np.array_equal(a, b) for a, b in zip(csr_from(g2), (og.off, og.col, og.w)))
    # flags: drop self loops, keep minimum duplicate weight
    g3 = H.Graph.from_edges(n, u, v, w, H.DROP_SELF_LOOPS | H.DEDUP_MIN)
    off3, col3, w3 = csr_from(g3)
    best = {}
    for a, b, c in zip(u, v, w):
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        best[key] = min(best.get(key, 1 << 30), c)
    got = {}
    for x in range(n):
        for e in range(off3[x], off3[x + 1]):
            assert col3[e] != x
            got[(min(x, col3[e]), max(x, col3[e]))] = w3[e]
    assert got == best


def test_save_and_reload_txt(tmp_path, golden_dir):
    g = H.Graph.load_txt(os.path.join(golden_dir, "euroroad2.txt"))
    p = tmp_path / "round.txt"
    g.save_txt(str(p))
    g2 = H.Graph.load_txt(str(p))
    for a, b in zip(csr_from(g), csr_from(g2)):
        assert np.array_equal(a, b)


def test_edgelist_loader(tmp_path):
    p = tmp_path / "x.e"
    p.write_text("Endpoint1 Endpoint2 Random-Weight (vertex ID from 0 to |V|-1; |V|= 6; |E|= 4)\n"
                 "0 1 35\n1 2 48\n4 5 7\n1 0 9\n")
    g = H.Graph.load_edgelist(str(p))
    off, col, w = csr_from(g)
    assert len(off) == 7
    assert list(zip(col[off[0]:off[1]], w[off[0]:off[1]])) == [(1, 9)]
    assert list(zip(col[off[1]:off[2]], w[off[1]:off[2]])) == [(0, 9), (2, 48)]
    assert off[4] - off[3] == 0  # vertex 3 isolated


def check_degree_order(off, col, w):
    n = len(off) - 1
    deg = np.diff(off)
    assert np.all(deg[:-1] >= deg[1:])
    for v in range(0, n, max(1, n // 200)):
        nb = col[off[v]:off[v + 1]]
        assert np.all(np.diff(nb.astype(np.int64)) > 0)  # sorted, no duplicates
        assert not np.any(nb == v)


def test_rmat_generator_matches_committed_fixture(golden_dir):
    z = np.load(os.path.join(golden_dir, "c1_rmat12_k4.npz"))
    g = H.Graph.rmat(12, 8, 0x5EED0001, 1, 100)
    off, col, w = csr_from(g)
    assert np.array_equal(off, z["off"]) and np.array_equal(col, z["col"]) and np.array_equal(w, z["w"])
    check_degree_order(off, col, w)
    assert w.min() >= 1 and w.max() <= 100
    assert len(col) // 2 <= 8 * 4096
    # symmetric
    pairs = set()
    for v in range(len(off) - 1):
        for e in range(off[v], off[v + 1]):
            pairs.add((v, int(col[e]), int(w[e])))
    assert all((b, a, c) in pairs for a, b, c in pairs)


def test_grid_generator():
    g = H.Graph.grid(16, 20, 7, 1, 1000)
    off, col, w = csr_from(g)
    assert len(off) - 1 == 320
    assert len(col) == 2 * (2 * 16 * 20 - 16 - 20)
    check_degree_order(off, col, w)
    deg = np.diff(off)
    assert (deg == 4).sum() == 14 * 18 and (deg == 2).sum() == 4


def test_compat_driver_is_built_and_links():
    """The C++ shim driver exists, links libhsdl.so and, without a GPU, stops
    at the first compute call with HSDL_ENODEV (no CPU fallback)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "compat_driver")
    assert os.path.exists(exe), "run make -C gpu4hophl_code_amd/csrc"
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present: the GPU tests run the driver")
    except ImportError:
        pass
    out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "euroroad2.txt"), "3", "5"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 3 and "no usable GPU" in out.stdout


def test_hybrid_compat_driver_is_built_and_links():
    """The hybrid-flow C++ driver (label_gen per vertex group, gpu_clean)
    links libhsdl.so and, without a GPU, stops at toCSR's upload with
    HSDL_ENODEV (no CPU fallback)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "compat_hybrid_driver")
    assert os.path.exists(exe), "run make -C gpu4hophl_code_amd/csrc"
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present: the GPU tests run the driver")
    except ImportError:
        pass
    out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "euroroad2.txt"), "3", "300",
                          "5", "/dev/null"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 3 and "no usable GPU" in out.stdout


def test_dijkstra_golden_equals_kround_golden(golden_dir):
    """The committed CPU_HopHL goldens (Dijkstra + clean) equal the K-round
    goldens (K-round + clean) on every graph and K they share: the two
    reference generators clean to the same labels."""
    z = np.load(os.path.join(golden_dir, "dijkstra_k.npz"))
    f = np.load(os.path.join(golden_dir, "fixtures_k.npz"))
    c1 = np.load(os.path.join(golden_dir, "c1_rmat12_k4.npz"))
    names = {"off": "lab_off", "hub": "hub", "hop": "hop", "dist": "dist"}
    assert len(z.files) == 4 * 19
    for k in z.files:
        want = c1[names[k.split("|")[-1]]] if k.startswith("rmat12|") else f[k]
        assert np.array_equal(z[k].astype(np.int64), want.astype(np.int64)), k


def test_generation_needs_gpu_no_cpu_fallback():
    import ctypes
    L = H.lib()
    n = ctypes.c_int()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    g = H.Graph.rmat(8, 4, 1)
    with pytest.raises(H.HSDLError) as e:
        H.hop_constrained_two_hop_labels_generation(g, 3)
    assert e.value.code == -8


@pytest.mark.parametrize("name", ["hsdl_opts", "hsdl_stats", "hsdl_server_opts",
                                  "hsdl_server_stats"])
def test_ctypes_structs_match_the_header(tmp_path, name):
    """The Python mirrors of the ABI structs have the C header's size and
    field offsets (a C program compiled against include/hsdl.h prints them)."""
    import ctypes
    import subprocess
    st = getattr(H, name)
    fields = [f for f, _ in st._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "hsdl.h"', "int main(void) {",
             f'  printf("size %zu\\n", sizeof({name}));']
    lines += [f'  printf("{f} %zu\\n", offsetof({name}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    assert int(out["size"]) == ctypes.sizeof(st)
    for f in fields:
        assert int(out[f]) == getattr(st, f).offset, f



def test_argument_errors_come_before_the_device():
    """Bad arguments are refused before any device work, with the codes the
    header documents (so these run without a GPU): K outside 1..15, an empty
    graph, K · max weight beyond the 27-bit label distance, and edge lists
    with an endpoint out of range or a weight below 1."""
    one = lambda a: np.asarray(a)
    g = H.Graph.from_edges(4, one([0, 1]), one([1, 2]), one([3, 5]))
    for K in (0, 16):
        with pytest.raises(H.HSDLError) as e:
            H.hop_constrained_two_hop_labels_generation(g, K)
        assert e.value.code == -1
    with pytest.raises(H.HSDLError) as e:
        H.hop_constrained_two_hop_labels_generation(H.Graph.from_edges(0, [], [], []), 3)
    assert e.value.code == -1
    wide = H.Graph.from_edges(3, one([0]), one([1]), one([(1 << 26) + 1]))
    with pytest.raises(H.HSDLError) as e:
        H.hop_constrained_two_hop_labels_generation(wide, 2)
    assert e.value.code == -7
    for u, v, w in (([0], [4], [1]), ([4], [0], [1]), ([0], [1], [0]), ([0], [1], [-3])):
        with pytest.raises(H.HSDLError) as e:
            H.Graph.from_edges(4, one(u), one(v), one(w))
        assert e.value.code == -1


def test_label_import_and_handle_errors(tmp_path):
    """Imported labels are validated on the host before any device work
    (offsets, hub range, hop ≤ K, 27-bit distances), missing files give
    HSDL_EIO, and null handles HSDL_EINVAL, all without a GPU."""
    import ctypes
    L = H.lib()
    off = np.array([0, 1, 3], np.uint64)
    hub = np.array([0, 0, 1], np.uint32)
    hop = np.array([0, 1, 0], np.uint8)
    dist = np.array([0, 4, 0], np.int32)

    def code(fn, *a):
        with pytest.raises(H.HSDLError) as e:
            fn(*a)
        return e.value.code

    I = H.Index.from_labels
    assert code(I, 2, 0, off, hub, hop, dist) == -1                                   # K = 0
    assert code(I, 2, 16, off, hub, hop, dist) == -1                                  # K > 15
    assert code(I, 2, 3, np.array([1, 1, 3], np.uint64), hub, hop, dist) == -1        # off[0] != 0
    assert code(I, 2, 3, np.array([0, 2, 1], np.uint64), hub, hop, dist) == -1        # decreasing
    assert code(I, 2, 3, off, np.array([0, 2, 1], np.uint32), hop, dist) == -1        # hub >= n
    assert code(I, 2, 1, off, hub, np.array([0, 2, 0], np.uint8), dist) == -7         # hop > K
    assert code(I, 2, 3, off, hub, hop, np.array([0, -1, 0], np.int32)) == -7         # dist < 0
    assert code(I, 2, 3, off, hub, hop, np.array([0, 1 << 27, 0], np.int32)) == -7    # dist >= 2^27
    assert code(H.Index.load, str(tmp_path / "missing"), 3) == -3
    assert code(H.Graph.load_txt, str(tmp_path / "missing.txt")) == -3
    nul = ctypes.c_void_p()
    assert L.hsdl_clean(None) == -1
    u32, i8, i32 = np.zeros(1, np.uint32), np.zeros(1, np.int8), np.zeros(1, np.int32)
    assert L.hsdl_query(None, 1, u32, u32, i8, i32) == -1
    assert L.hsdl_extract_paths(None, 1, u32, u32, i8, 0, None, u32) == -1
    assert L.hsdl_server_create(None, None, ctypes.byref(nul)) == -1


def _dump_bytes(rows):
    """The reference's label dump (Clean_pc/main.cu:11-57) written with
    numpy: size_t n, then per vertex size_t count and count records of four
    int32 {hub, parent, hop, dist}."""
    out = [np.array([len(rows)], np.uint64).tobytes()]
    for r in rows:
        out.append(np.array([len(r)], np.uint64).tobytes())
        if r:
            out.append(np.array(r, np.int32).reshape(-1).tobytes())
    return b"".join(out)


def test_label_dump_errors_before_the_device(tmp_path):
    """Malformed dumps are refused on the host: a hop above K (a hop of 32
    or more would otherwise spill into the packed hub field) gives
    HSDL_ERANGE, a count larger than the file or a truncated record
    HSDL_EIO -- never an abort from a huge allocation."""
    def code(path, K=5):
        with pytest.raises(H.HSDLError) as e:
            H.Index.load(str(path), K)
        return e.value.code

    ok_rows = [[(0, 0, 0, 0)], [(0, 0, 1, 7), (1, 0, 0, 0)]]
    p = tmp_path / "hop33.bin"
    p.write_bytes(_dump_bytes([[(0, 0, 0, 0)], [(0, 0, 33, 7), (1, 0, 0, 0)]]))
    assert code(p) == -7
    p = tmp_path / "hop6.bin"
    p.write_bytes(_dump_bytes([[(0, 0, 0, 0)], [(0, 0, 6, 7), (1, 0, 0, 0)]]))
    assert code(p) == -7
    # a per-vertex count of 2^39 records in a file of a few bytes
    good = _dump_bytes(ok_rows)
    bad = bytearray(good)
    bad[8:16] = np.array([1 << 39], np.uint64).tobytes()
    p = tmp_path / "count.bin"
    p.write_bytes(bytes(bad))
    assert code(p) == -3
    # a vertex count beyond what the file can hold
    bad = bytearray(good)
    bad[0:8] = np.array([1 << 31], np.uint64).tobytes()
    p = tmp_path / "n.bin"
    p.write_bytes(bytes(bad))
    assert code(p) == -3
    p = tmp_path / "trunc.bin"
    p.write_bytes(good[:-3])
    assert code(p) == -3
    p = tmp_path / "empty.bin"
    p.write_bytes(b"")
    assert code(p) == -3
    with pytest.raises(H.HSDLError) as e:
        H.Index.load(str(tmp_path / "hop6.bin"), 0)
    assert e.value.code == -1


def test_reorder_tie_rule_is_degree_desc_then_id():
    """The degree renaming (reference
    graph_v_of_v_update_vertexIDs_by_degrees_large_to_small.h:14-40, whose
    std::sort leaves ties unordered) is fixed here as (degree desc, old id
    asc): the oracle's and the library's new->old maps both equal numpy's
    lexsort of (id, -degree), on graphs with many equal degrees."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(2, 60))
        m = int(rng.integers(0, 3 * n))
        u = rng.integers(0, n, m)
        v = rng.integers(0, n, m)
        keep = u != v
        u, v = u[keep], v[keep]
        w = rng.integers(1, 9, len(u))
        g = H.Graph.from_edges(n, u, v, w, H.DROP_SELF_LOOPS)
        off, col, ww = g.csr()
        deg = np.diff(off.astype(np.int64))
        want = np.lexsort((np.arange(n), -deg))
        og = O.CSR(n, off.astype(np.int64), col.astype(np.int32), ww)
        _, on2o = O.reorder(og)
        assert np.array_equal(on2o, want.astype(np.int32)), trial
        n2o = g.reorder_by_degree()
        assert np.array_equal(n2o.astype(np.int64), want), trial
        # a planted mistake (ties by id descending) differs on these graphs
        if trial == 0:
            assert not np.array_equal(np.lexsort((-np.arange(n), -deg)), want) or n < 3


@pytest.mark.parametrize("text", ["|V|= 99999999999999\n", "|V|= -4\n0 1 1\n", "garbage\n0 1 1\n",
                                  "x |V|= 3\n0 1 1e30\n", "x |V|= 3\n0 1 nan\n",
                                  "x |V|= 3\n0 5 2\n"])
def test_edgelist_header_and_weight_errors(tmp_path, text):
    """The .e loader refuses a header without a readable vertex count in
    0..2^32-1 (the reference's reader exits on it), weights outside int32
    (whose int conversion would be undefined) and endpoints outside |V|,
    with HSDL_EIO and without sizing anything from the bad values."""
    p = tmp_path / "bad.e"
    p.write_text(text)
    with pytest.raises(H.HSDLError) as e:
        H.Graph.load_edgelist(str(p))
    assert e.value.code == -3
