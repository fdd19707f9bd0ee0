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
dominant kernel) and the CPU baseline (the oracle's K-round build on a
bounded sample, timed on this host).
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


def cpu_baseline(scale=15, K=5):
    """The oracle's K-round build + clean (restated reference CPU twin) on a
    bounded sample: an R-MAT graph of the same family at a smaller scale."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import gpu4hophl_code_amd as H
    import oracle_lib as O
    g = H.Graph.rmat(scale, 16, 0x5EED0002, 1, 100)
    off, col, w = g.csr()
    c = O.CSR(len(off) - 1, off.astype(np.int64), col.astype(np.int32), w)
    threads = cpu_threads()
    t0 = time.perf_counter()
    lab = O.kround(c, K, threads=threads)
    cl = O.clean(lab, 1, threads=threads)
    sec = time.perf_counter() - t0
    return {"value": cl.total() / sec, "unit": "labels/s", "cores": threads, "kind": "port",
            "sample": f"R-MAT scale {scale} (same generator, edge factor 16, K={K}), full "
                      f"K-round build + clean by the oracle, {cl.total()} labels in {sec:.2f} s",
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


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

    # rooflines of the two build kernels (HIP events on the build stream,
    # summed over their launches); the slower one is reported as `roofline`.
    #   k_prune:  4 or 8 B per label streamed from L(y) (the stream's label
    #             width, prune_label_bytes) + 8 B per label staged from L(s)
    #             + 8 B per candidate (its list entry and dense slot)
    #   k_expand: 28 B per frontier item (item + its two CSR offsets)
    #             + 12 B per scanned edge ((col, w) + the 4 B dense slot,
    #             read, and updated by an atomic when the edge lowers it)
    #             + 8 B per candidate appended
    launches = max(1, stats["prune_launches"])  # one expand and one prune per round
    kern = {
        "k_prune": ((stats.get("prune_label_bytes") or 8) * stats["prune_label_reads"]
                    + 8 * stats["prune_tile_reads"] + 8 * stats["candidates"], stats["t_prune_s"]),
        "k_expand": (28 * stats["frontier_items"] + 12 * stats["scanned_edges"]
                     + 8 * stats["candidates"], stats["t_expand_s"]),
    }
    def traffic_ratio(name):
        # HBM bytes per algorithmic byte, measured with rocprofv3 FETCH_SIZE and
        # WRITE_SIZE passes on this workload (profiles/<kernel>_<config>_k<K>.md)
        tpath = os.path.join(ROOT, "profiles", f"{name}_{args.config}_k{K}_traffic.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                return json.load(f).get("traffic_per_algorithmic_byte")
        return None

    rl = {}
    for name, (nbytes, secs) in kern.items():
        achieved = nbytes / secs / 1e9 if secs > 0 else 0.0
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
             "kernel": name, "bytes_per_launch": nbytes / launches,
             "avg_launch_us": secs / launches * 1e6, "launches": launches, "seconds": secs}
        ratio = traffic_ratio(name)
        if ratio:
            r["traffic"] = ratio * r["bytes_per_launch"]
            r["traffic_per_algorithmic_byte"] = ratio
        rl[name] = r
    top = max(rl, key=lambda x: rl[x]["seconds"])
    roofline = rl[top]
    roofline_other = {k: v for k, v in rl.items() if k != top}
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
        # dense-slot accesses per second (one read per scanned edge, and an
        # atomic when the edge lowers the slot)
        ex["slot_gops"] = round(stats["scanned_edges"] / stats["t_expand_s"] / 1e9, 2) \
            if stats["t_expand_s"] > 0 else None
    # the clean's first pass: 8 B per label of L(v) + 4 or 8 B per label of
    # the heads' lists L(u) it streams (the stream's label width)
    if stats["t_clean_s"] > 0:
        cb = 8 * stats["labels_uncleaned"] + \
            (stats.get("clean_label_bytes") or 8) * stats["clean_label_reads"]
        ca = cb / stats["t_clean_s"] / 1e9
        cr = traffic_ratio("k_clean_tile")
        roofline_other["k_clean_tile"] = {
            "bound": "hbm", "achieved": round(ca, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ca / HBM_PEAK_GBS, 4), "traffic": cr * cb if cr else None, "bytes": cb,
            "seconds": stats["t_clean_s"], "note": "host-timed clean (both passes, compaction)"}
        if cr:
            roofline_other["k_clean_tile"]["traffic_per_algorithmic_byte"] = cr

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
        os.environ.setdefault("HSDL_PROGRESS", "30")  # the plain build reports as it goes
        t0 = time.perf_counter()
        try:
            pidx = H.hop_constrained_two_hop_labels_generation(
                g, K, mode=H.MODE_PLAIN, deadline_s=args.plain_deadline)
            pidx.clean()
            psec = time.perf_counter() - t0
            plain = {"build_s": psec, "labels": pidx.label_count(),
                     "same_label_count": pidx.label_count() == labels, "completed": True}
            pidx.free()
        except H.HSDLError as e:
            if e.code != -5:  # HSDL_ETIMEDOUT
                raise
            plain = {"build_s": None, "completed": False,
                     "lower_bound_s": time.perf_counter() - t0}
        log(f"plain build: {plain}")

    cpu = cpu_baseline() if args.cpu_baseline == "on" else None
    if cpu:
        cpu["query"] = cpu_q
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC, "value": value, "unit": "labels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": cfg["desc"], "vertices": n, "csr_entries": nnz, "K": K,
                   "parallelism": "single GPU"},
        "build_s": ms / 1e3, "labels_uncleaned": stats["labels_uncleaned"],
        "labels_cleaned": labels, "generate_s": stats["t_generate_s"],
        "clean_s": stats["t_clean_s"], "hub_batch_final": stats["hub_batch"],
        "batches": stats["batches"],
        "plain_build_s": plain["build_s"] if plain else None,
        "batched_vs_plain": (plain["build_s"] / (ms / 1e3)) if plain and plain["completed"] else None,
        "batched_vs_plain_at_least": (plain["lower_bound_s"] / (ms / 1e3))
        if plain and not plain["completed"] else None,
        "query": query, "paths": paths, "serving": serve, "roofline": roofline,
        "roofline_other_kernels": roofline_other, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
