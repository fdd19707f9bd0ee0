#!/usr/bin/env python3
"""HSDL benchmark on one MI355X: cleaned-index build (batched vs plain),
device-resident query throughput and open-loop serving latency.

A step is one build of the cleaned index of the workload graph: K-round
generation (batched, autotuned hub batches) followed by the canonical clean,
starting from the CSR already resident in HBM and ending with the cleaned
index resident in HBM.  `value` is cleaned labels per second of build time.

Default workload: R-MAT scale 18, edge factor 16, (a,b,c,d) =
(0.57,0.19,0.19,0.05), weights U[1,100], degree order, K = 5.

Prints one JSON line.  Extra fields: build_s (batched), plain_build_s and
batched_vs_plain, query throughput on device buffers, serving p50/p99 at the
requested Poisson rates, the roofline of the prune kernel (the build's
dominant kernel) against the reference algorithm's bytes (9 B per label of
L_{<h}(y) read per candidate and of L_{<h}(s) per active hub, counted by the
oracle and committed with the full-size fixtures), the CPU baseline (the
reference's threaded CPU_HopHL generation on the same graph, time-boxed) and,
by default, an R-MAT scale-20 leg (batched build + clean, labels before and
after, serving at the same rates on that index, and a deadline-bounded plain
build for a lower bound on the speedup).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HSDL build time + queries/s, p50/p99; batched vs plain on 1 MI355X"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "rmat18": dict(kind="rmat", scale=18, ef=16, K=5, wmax=100, seed=0x5EED0002,
                   desc="R-MAT scale 18 (262,144 vertices), edge factor 16, "
                        "(0.57,0.19,0.19,0.05), weights U[1,100], degree order, K=5"),
    "rmat12": dict(kind="rmat", scale=12, ef=8, K=4, wmax=100, seed=0x5EED0001,
                   desc="R-MAT scale 12 (4,096 vertices), edge factor 8, weights U[1,100], "
                        "degree order, K=4"),
    "rmat20": dict(kind="rmat", scale=20, ef=16, K=5, wmax=100, seed=0x5EED0004,
                   desc="R-MAT scale 20 (1,048,576 vertices), edge factor 16, weights U[1,100], "
                        "degree order, K=5"),
    "grid1024": dict(kind="grid", rows=1024, cols=1024, K=8, wmax=1000, seed=0x5EED0003,
                     desc="1024x1024 4-neighbour grid, weights U[1,1000], degree order, K=8"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_graph(H, cfg):
    if cfg["kind"] == "rmat":
        return H.Graph.rmat(cfg["scale"], cfg["ef"], cfg["seed"], 1, cfg["wmax"])
    return H.Graph.grid(cfg["rows"], cfg["cols"], cfg["seed"], 1, cfg["wmax"])


def build(H, g, K, mode, timing):
    t0 = time.perf_counter()
    idx = H.hop_constrained_two_hop_labels_generation(g, K, mode=mode, kernel_timing=timing)
    idx.clean()
    t1 = time.perf_counter()
    return idx, t1 - t0


def query_throughput(H, idx, n, nq, torch, reps=5):
    """Queries from device buffers on torch's current stream, timed with
    torch.cuda events on that stream."""
    rng = np.random.default_rng(0x5EED0200)
    s = rng.integers(0, n, nq).astype(np.uint32)
    t = rng.integers(0, n, nq).astype(np.uint32)
    k = rng.integers(1, 6, nq).astype(np.int8)
    ds = torch.from_numpy(s.view(np.int32)).cuda()
    dt = torch.from_numpy(t.view(np.int32)).cuda()
    dk = torch.from_numpy(k).cuda()
    dout = torch.empty(nq, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    idx.query_device(ds.data_ptr(), dt.data_ptr(), dk.data_ptr(), dout.data_ptr(), nq,
                     stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        idx.query_device(ds.data_ptr(), dt.data_ptr(), dk.data_ptr(), dout.data_ptr(), nq,
                         stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    return s, t, k, dout.cpu().numpy(), sec


def label_lengths(idx):
    return np.diff(idx.offsets().astype(np.int64))


def serving(H, idx, n, rates, nq):
    rng = np.random.default_rng(0x5EED0205)
    s = rng.integers(0, n, nq).astype(np.uint32)
    t = rng.integers(0, n, nq).astype(np.uint32)
    k = rng.integers(1, 6, nq).astype(np.int8)
    alone = idx.query(s, t, k)
    out = []
    for rate in rates:
        srv = H.Server(idx, max_batch=65536, max_wait_us=200, inflight=4)
        gaps = rng.exponential(1e9 / rate, nq)
        arr = np.cumsum(gaps).astype(np.uint64)
        t0 = time.perf_counter()
        ans, lat = srv.replay(s, t, k, arr)
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        same = bool(np.array_equal(ans, alone))
        out.append({"offered_qps": rate, "achieved_qps": nq / wall,
                    "p50_us": float(np.percentile(lat, 50)) / 1e3,
                    "p99_us": float(np.percentile(lat, 99)) / 1e3,
                    "batches": int(st["batches"]), "mean_batch": round(st["mean_batch"], 1),
                    "answers_equal_alone": same})
        log(f"serving {rate/1e6:.0f}M q/s: {out[-1]}")
    return out


def plain_per_query(H, idx, n, torch, count=2000):
    """One launch per query with a sync after each (the reference's query
    pattern, Query_mdnd/src/gpu_query.cu:139-155)."""
    rng = np.random.default_rng(0x5EED0206)
    s = torch.from_numpy(rng.integers(0, n, count).astype(np.int32)).cuda()
    t = torch.from_numpy(rng.integers(0, n, count).astype(np.int32)).cuda()
    k = torch.from_numpy(rng.integers(1, 6, count).astype(np.int8)).cuda()
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    lat = np.zeros(count)
    for i in range(count):
        t0 = time.perf_counter()
        idx.query_device(s.data_ptr() + 4 * i, t.data_ptr() + 4 * i, k.data_ptr() + i,
                         out.data_ptr() + 4 * i, 1, stream)
        torch.cuda.synchronize()
        lat[i] = time.perf_counter() - t0
    return {"qps": count / lat.sum(), "p50_us": float(np.percentile(lat, 50)) * 1e6,
            "p99_us": float(np.percentile(lat, 99)) * 1e6, "queries": count}


def oracle_csr(g):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    off, col, w = g.csr()
    return O, O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)


def cpu_baseline(g, K, n, seconds=20.0):
    """The reference's CPU path (CPU_HopHL: threaded per-hub Dijkstra
    generation, restated by the oracle) on the bench graph itself, stopped
    after `seconds`: labels per second over the hubs it finished.  The full
    generation does not finish in a bench's time (a DNF at the deadline, with
    the fraction of hubs done); the K-round twin's full build + clean at
    R-MAT scale 15 is reported beside it."""
    O, c = oracle_csr(g)
    threads = cpu_threads()
    hubs, labels, sec = O.dijkstra_timed(c, K, threads=threads, seconds=seconds)
    out = {"value": labels / sec, "unit": "labels/s", "cores": threads, "kind": "port",
           "sample": f"CPU_HopHL threaded Dijkstra generation on the bench graph (n={n}), "
                     f"stopped at {seconds:.0f} s: {hubs} of {n} hubs in id order "
                     f"({100.0 * hubs / n:.2f}%, DNF), {labels} uncleaned labels in {sec:.2f} s",
           "hubs_done": hubs, "hubs_total": n, "completed": hubs == n,
           "nproc": os.cpu_count(), "cpu_model": cpu_model()}
    # the reference's other CPU generation, 2023WWW (per-hub level search
    # without rank pruning), on the same graph and the same time box
    wh, wl, ws = O.www2023_timed(c, K, threads=threads, seconds=seconds / 2)
    out["www2023"] = {"labels_per_s": wl / ws, "hubs_done": wh, "hubs_total": n,
                      "labels": wl, "seconds": ws, "threads": threads,
                      "sample": f"2023WWW threaded generation on the bench graph, stopped at "
                                f"{seconds / 2:.0f} s: {wh} of {n} hubs ({100.0 * wh / n:.2f}%)"}
    # the K-round twin (the GPU's own algorithm on the CPU) on a graph of the
    # same family small enough to finish
    g15 = __import__("gpu4hophl_code_amd").Graph.rmat(15, 16, 0x5EED0002, 1, 100)
    _, c15 = oracle_csr(g15)
    t0 = time.perf_counter()
    lab = O.kround(c15, K, threads=threads)
    cl = O.clean(lab, 1, threads=threads)
    s15 = time.perf_counter() - t0
    out["kround_scale15"] = {"labels_per_s": cl.total() / s15, "seconds": s15,
                             "labels": cl.total(), "threads": threads,
                             "sample": "R-MAT scale 15 (same generator, edge factor 16, K=5): "
                                       "full K-round build + clean by the oracle"}
    return out


def cpu_threads():
    return min(16, os.cpu_count() or 1)  # the box's CPU share for one GPU


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_query_baseline(idx, n, K):
    """The oracle's merge-join query (restated reference CPU query) on the same
    cleaned index, one query per thread: 1 thread and the box's share."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    off, hub, hop, dist = idx.labels()
    L = O.labels_from_arrays(n, off.astype(np.int64), hub, hop, dist)
    rng = np.random.default_rng(0x5EED0207)
    out = {"threads": cpu_threads(), "kind": "port"}
    for nthr, nq in ((1, 100_000), (cpu_threads(), 1_000_000)):
        s = rng.integers(0, n, nq).astype(np.uint32)
        t = rng.integers(0, n, nq).astype(np.uint32)
        k = rng.integers(1, K + 1, nq).astype(np.int32)
        t0 = time.perf_counter()
        O.query(L, s, t, k, threads=nthr)
        out[f"qps_{nthr}_threads" if nthr > 1 else "qps_1_thread"] = nq / (time.perf_counter() - t0)
    return out


# The reference algorithm's per-build counts (the oracle's K-round build of
# the same graph, committed with the full-size fixtures): frontier items,
# their full adjacency, candidates, labels of L_{<h}(y) per candidate and of
# L_{<h}(s) per active hub, accepted labels.
ALG_FIXTURES = {"rmat18": "full_c2_rmat18_k5.npz", "grid1024": "full_c3_grid1024_k8.npz"}


def ref_counters(config):
    name = ALG_FIXTURES.get(config)
    path = os.path.join(ROOT, "tests", "golden", name) if name else None
    if not path or not os.path.exists(path):
        return None
    z = np.load(path)
    return {k: int(z[k]) for k in ("frontier_items", "scanned_edges", "candidates",
                                   "cand_label_reads", "hub_label_reads", "accepted")}


def clean_reference_bytes(H, g, K):
    """Algorithmic bytes of the reference clean on this graph's labels: 9 B
    per label of L(v), 9 B per label of L(u) for every label (u,.,.) of L(v)
    with u != v (the reference reads L(u) once per label), 9 B per kept
    label.  From an untimed build's labels."""
    idx = H.hop_constrained_two_hop_labels_generation(g, K)
    off, hub = idx.hubs()
    off = off.astype(np.int64)
    lens = np.diff(off)
    own = np.repeat(np.arange(len(lens), dtype=np.int64), lens)
    reads = int(lens[hub].astype(np.int64)[hub.astype(np.int64) != own].sum())
    idx.clean()
    kept = idx.label_count()
    idx.free()
    return 9 * int(off[-1]) + 9 * reads + 9 * kept


def profile_traffic(name, config, K):
    """HBM bytes per build of a kernel, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes on this workload (profiles/<kernel>_<config>_k<K>_traffic.json,
    each counter scaled by its reading on a copy of known size)."""
    path = os.path.join(ROOT, "profiles", f"{name}_{config}_k{K}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def rooflines(H, stats, config, K, ref, clean_bytes):
    """Rooflines of the build kernels, timed with HIP events on the build
    stream (summed over a build's launches).  `achieved` divides the
    reference algorithm's bytes per launch (ref: the oracle's counts) by the
    average launch time:
      k_prune:  9 B x (labels of L_{<h}(y) per candidate + labels of
                L_{<h}(s) per active hub)  -- the reference's prune reads
                every label of both lists; this kernel stops at the first
                witness and serves repeated lists from L2, so its rate can
                pass the HBM peak
      k_expand: 8 B per frontier item + 8 B per entry of its whole adjacency
      clean:    9 B per label of L(v) + 9 B x |L(u)| per label (u,.,.) of
                L(v), u != v + 9 B per kept label
    Without reference counts (no fixture for the workload) the kernels' own
    streamed bytes are used.  `kernel_stream` gives those bytes for every
    kernel, and `traffic` the HBM bytes per launch measured by rocprofv3."""
    launches = max(1, stats["prune_launches"])  # one expand and one prune per round
    own = {
        # 4 or 8 B per label of L(y) streamed (up to the step with the
        # witness) + 8 B per label of L(s) staged + 8 B per candidate + 4 B
        # per dense slot row items scanned
        "k_prune": (stats.get("prune_label_bytes") or 8) * stats["prune_label_reads"]
        + 8 * stats["prune_tile_reads"] + 8 * stats["candidates"]
        + 4 * stats.get("prune_row_slots", 0),
        # 8 B per frontier label (its label and its share of the group
        # record) + 12 B per relaxed (label, edge) pair (8 B (col, w), which
        # the labels of a group share, so an upper bound, + the 4 B dense
        # slot) + 8 B per candidate appended
        "k_expand": 8 * stats["frontier_items"] + 12 * stats["scanned_edges"]
        + 8 * stats["candidates"],
    }
    alg = None
    if ref:
        alg = {"k_prune": 9 * (ref["cand_label_reads"] + ref["hub_label_reads"]),
               "k_expand": 8 * (ref["frontier_items"] + ref["scanned_edges"])}
    secs = {"k_prune": stats["t_prune_s"], "k_expand": stats["t_expand_s"]}
    rl = {}
    for name in ("k_prune", "k_expand"):
        nbytes = alg[name] if alg else own[name]
        t = secs[name]
        achieved = nbytes / t / 1e9 if t > 0 else 0.0
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
             "kernel": name, "bytes": "reference algorithm" if alg else "kernel stream",
             "bytes_per_launch": nbytes / launches, "avg_launch_us": t / launches * 1e6,
             "launches": launches, "seconds": t,
             "kernel_stream": {"bytes_per_launch": own[name] / launches,
                               "achieved": round(own[name] / t / 1e9, 1) if t > 0 else None,
                               "frac": round(own[name] / t / 1e9 / HBM_PEAK_GBS, 4)
                               if t > 0 else None}}
        tr = profile_traffic(name, config, K)
        if tr and tr.get("hbm_bytes_per_build"):
            r["traffic"] = tr["hbm_bytes_per_build"] / tr.get("launches_per_build", launches)
            r["traffic_per_reference_byte"] = round(r["traffic"] / r["bytes_per_launch"], 4)
        rl[name] = r
    top = max(rl, key=lambda x: rl[x]["seconds"])
    roofline = rl[top]
    other = {k: v for k, v in rl.items() if k != top}
    # achievable copy bandwidth on this GPU (STREAM-style, read + write) and
    # the rate of random returning 32-bit atomicMin over 16 GB (the dense
    # slots' update; the expansion reads a slot first and updates it only
    # when the edge lowers it)
    import ctypes
    gbps = ctypes.c_double()
    if H.lib().hsdl_device_copy_bench(1 << 30, 5, ctypes.byref(gbps)) == 0:
        roofline["copy_gbps"] = round(gbps.value, 1)
    gops = ctypes.c_double()
    if H.lib().hsdl_device_atomic_bench(16 << 30, 1 << 30, ctypes.byref(gops)) == 0:
        ex = rl["k_expand"]
        ex["atomic_peak_gops"] = round(gops.value, 2)
        ex["slot_gops"] = round(stats["scanned_edges"] / stats["t_expand_s"] / 1e9, 2) \
            if stats["t_expand_s"] > 0 else None
    if stats["t_clean_s"] > 0:
        cown = 8 * stats["labels_uncleaned"] + \
            (stats.get("clean_label_bytes") or 8) * stats["clean_label_reads"]
        cb = clean_bytes if clean_bytes else cown
        ca = cb / stats["t_clean_s"] / 1e9
        other["clean"] = {
            "bound": "hbm", "achieved": round(ca, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ca / HBM_PEAK_GBS, 4), "bytes": "reference algorithm" if clean_bytes
            else "kernel stream", "bytes_per_build": cb, "seconds": stats["t_clean_s"],
            "kernel_stream": {"bytes_per_build": cown,
                              "frac": round(cown / stats["t_clean_s"] / 1e9 / HBM_PEAK_GBS, 4)},
            "note": "host-timed clean (both passes, compaction)"}
        tr = profile_traffic("k_clean_tile", config, K)
        if tr and tr.get("hbm_bytes_per_build"):
            other["clean"]["traffic"] = tr["hbm_bytes_per_build"]
    return roofline, other


def plain_build(H, g, K, deadline, labels):
    """The plain one-hub-per-launch form of the same build (generation +
    clean), stopped at `deadline` seconds (0: to the end); a stopped build
    gives a lower bound."""
    os.environ.setdefault("HSDL_PROGRESS", "30")  # the plain build reports as it goes
    t0 = time.perf_counter()
    try:
        pidx = H.hop_constrained_two_hop_labels_generation(g, K, mode=H.MODE_PLAIN,
                                                           deadline_s=deadline)
        pidx.clean()
        psec = time.perf_counter() - t0
        out = {"build_s": psec, "labels": pidx.label_count(),
               "same_label_count": pidx.label_count() == labels, "completed": True}
        pidx.free()
    except H.HSDLError as e:
        if e.code != -5:  # HSDL_ETIMEDOUT
            raise
        out = {"build_s": None, "completed": False, "lower_bound_s": time.perf_counter() - t0}
    return out


def scale20_leg(H, torch, args, warmup=1, steps=3):
    """R-MAT scale 20, edge factor 16, weights U[1,100], K=5: batched build +
    clean (1 untimed + 3 timed), labels before and after the clean, the
    build rooflines, serving on the cleaned index at the same rates, and the
    plain form stopped at a deadline (a lower bound on batched vs plain)."""
    cfg = CONFIGS["rmat20"]
    K = cfg["K"]
    t0 = time.perf_counter()
    g = make_graph(H, cfg)
    n, nnz = g.info()
    g.upload(0)
    log(f"rmat20: graph n={n} nnz={nnz} in {time.perf_counter() - t0:.1f} s")
    for i in range(warmup):
        idx, sec = build(H, g, K, H.MODE_BATCHED, False)
        log(f"rmat20 warmup {i}: {sec:.3f} s")
        idx.free()
    times, stats = [], None
    for i in range(steps):
        torch.cuda.synchronize()
        idx, sec = build(H, g, K, H.MODE_BATCHED, True)
        torch.cuda.synchronize()
        times.append(sec)
        stats = idx.stats()
        log(f"rmat20 step {i}: {sec:.3f} s; stats {stats}")
        if i + 1 < steps:
            idx.free()
    build_s = float(np.mean(times))
    rl, rl_other = rooflines(H, stats, "rmat20", K, None, None)
    out = {"workload": cfg["desc"], "vertices": n, "csr_entries": nnz, "steps": steps,
           "warmup": warmup, "build_s": build_s, "build_s_each": times,
           "labels_uncleaned": stats["labels_uncleaned"],
           "labels_cleaned": stats["labels_cleaned"], "generate_s": stats["t_generate_s"],
           "clean_s": stats["t_clean_s"], "batches": stats["batches"],
           "labels_per_s": stats["labels_cleaned"] / build_s, "roofline": rl,
           "roofline_other_kernels": rl_other}
    if args.serve == "on":
        rates = [float(x) for x in args.rates.split(",")]
        out["serving"] = {"rates": serving(H, idx, n, rates, args.queries),
                          "plain_per_query": plain_per_query(H, idx, n, torch)}
    labels = stats["labels_cleaned"]
    idx.free()
    if args.plain == "on":
        p = plain_build(H, g, K, args.s20_plain_deadline, labels)
        out["plain"] = p
        out["batched_vs_plain"] = p["build_s"] / build_s if p["completed"] else None
        out["batched_vs_plain_at_least"] = (p["lower_bound_s"] / build_s
                                            if not p["completed"] else None)
        log(f"rmat20 plain build: {p}")
    g.release_device()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="rmat18", choices=sorted(CONFIGS))
    ap.add_argument("--plain", default="on", choices=["on", "off"])
    ap.add_argument("--serve", default="on", choices=["on", "off"])
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--paths", default="on", choices=["on", "off"],
                    help="also build with parents and time hop-constrained path extraction")
    ap.add_argument("--rates", default="10e6,30e6,60e6")
    ap.add_argument("--queries", type=int, default=1 << 22)
    ap.add_argument("--plain-deadline", type=float, default=0.0,
                    help="stop the plain build after this many seconds (0: run to the end); "
                         "a stopped build reports a lower bound")
    ap.add_argument("--scale20", default="on", choices=["on", "off"],
                    help="also measure R-MAT scale 20 (build, clean, serving, bounded plain "
                         "build) after the headline workload")
    ap.add_argument("--s20-plain-deadline", type=float, default=60.0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--reference-bytes", default="on", choices=["on", "off"],
                    help="count the reference clean's bytes on an extra untimed build "
                         "(off: one build only, e.g. under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this "
              "build runs the generation, clean and queries on a single GPU", file=sys.stderr)
        sys.exit(2)

    import torch
    import gpu4hophl_code_amd as H

    cfg = CONFIGS[args.config]
    K = cfg["K"]
    g = make_graph(H, cfg)
    n, nnz = g.info()
    g.upload(0)
    log(f"workload: {cfg['desc']}: n={n} nnz={nnz}")

    for i in range(args.warmup):
        idx, sec = build(H, g, K, H.MODE_BATCHED, False)
        log(f"warmup {i}: {sec:.3f} s, {idx.label_count()} labels")
        idx.free()

    times, stats = [], None
    torch.cuda.synchronize()
    for i in range(args.steps):
        torch.cuda.synchronize()
        idx, sec = build(H, g, K, H.MODE_BATCHED, True)
        torch.cuda.synchronize()
        times.append(sec)
        stats = idx.stats()
        log(f"step {i}: {sec:.3f} s; stats {stats}")
        if i + 1 < args.steps:
            idx.free()
    ms = 1e3 * float(np.mean(times))
    labels = stats["labels_cleaned"]
    value = labels / (ms * 1e-3)

    roofline, roofline_other = rooflines(
        H, stats, args.config, K, ref_counters(args.config),
        clean_reference_bytes(H, g, K)
        if args.config in ALG_FIXTURES and args.reference_bytes == "on" else None)

    # queries on the device-resident cleaned index
    lens = label_lengths(idx)
    s, t, k, ans, qsec = query_throughput(H, idx, n, args.queries, torch)
    qbytes = float((9 * (lens[s] + lens[t]) + 13)[s != t].sum() + 13 * (s == t).sum())
    query = {"queries": args.queries, "qps": args.queries / qsec, "kernel_ms": qsec * 1e3,
             "roofline": {"bound": "hbm", "achieved": round(qbytes / qsec / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(qbytes / qsec / 1e9 / HBM_PEAK_GBS, 4)}}
    log(f"queries: {query}")

    # hop-constrained shortest paths: a second build that also derives every
    # label's parent, then 2^18 path queries through the host-array API
    # (uploads and downloads included)
    paths = None
    if args.paths == "on":
        pidx = H.hop_constrained_two_hop_labels_generation(g, K, clean=True, parents=True)
        pst = pidx.stats()
        rng = np.random.default_rng(0x5EED0400)
        npq = 1 << 18
        ps = rng.integers(0, n, npq).astype(np.uint32)
        pt = rng.integers(0, n, npq).astype(np.uint32)
        pk = rng.integers(1, K + 1, npq).astype(np.int8)
        pidx.paths(ps[:1024], pt[:1024], pk[:1024], cap=K)  # warm-up
        t0 = time.perf_counter()
        _, pcounts = pidx.paths(ps, pt, pk, cap=K)
        psec = time.perf_counter() - t0
        paths = {"parents_s": pst["t_parents_s"], "queries": npq, "paths_per_s": npq / psec,
                 "ms": psec * 1e3, "mean_edges": float(pcounts.mean()),
                 "note": "host arrays in and out (PCIe copies included)"}
        pidx.free()
        log(f"paths: {paths}")

    serve = None
    if args.serve == "on":
        rates = [float(x) for x in args.rates.split(",")]
        serve = {"rates": serving(H, idx, n, rates, args.queries),
                 "plain_per_query": plain_per_query(H, idx, n, torch)}
        log(f"plain per-query launches: {serve['plain_per_query']}")

    cpu_q = cpu_query_baseline(idx, n, K) if args.cpu_baseline == "on" else None
    if cpu_q:
        log(f"cpu query baseline: {cpu_q}")

    plain = None
    if args.plain == "on":
        idx.free()
        plain = plain_build(H, g, K, args.plain_deadline, labels)
        log(f"plain build: {plain}")

    cpu = cpu_baseline(g, K, n, args.cpu_seconds) if args.cpu_baseline == "on" else None
    if cpu:
        cpu["query"] = cpu_q
        log(f"cpu baseline: {cpu}")

    s20 = None
    if args.scale20 == "on" and args.config == "rmat18":
        g.release_device()
        del g
        s20 = scale20_leg(H, torch, args)

    line = {
        "metric": METRIC, "value": value, "unit": "labels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": cfg["desc"], "vertices": n, "csr_entries": nnz, "K": K,
                   "parallelism": "single GPU"},
        "build_s": ms / 1e3, "build_s_each": times, "labels_uncleaned": stats["labels_uncleaned"],
        "labels_cleaned": labels, "generate_s": stats["t_generate_s"],
        "clean_s": stats["t_clean_s"], "hub_batch_final": stats["hub_batch"],
        "batches": stats["batches"],
        "plain_build_s": plain["build_s"] if plain else None,
        "batched_vs_plain": (plain["build_s"] / (ms / 1e3)) if plain and plain["completed"] else None,
        "batched_vs_plain_at_least": (plain["lower_bound_s"] / (ms / 1e3))
        if plain and not plain["completed"] else None,
        "query": query, "paths": paths, "serving": serve, "roofline": roofline,
        "roofline_other_kernels": roofline_other, "cpu_baseline": cpu, "rmat20": s20,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
