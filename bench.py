#!/usr/bin/env python3
"""Benchmark of the continuous-subgraph-matching path on one MI355X.

A step is one update batch of the plain method.  The headline workload is the largest one-GPU
configuration: a seeded Chung-Lu graph of 4M vertices and ~35M edges (degree exponent 2.0, eight
planted hubs of expected degree 20,000, the other degrees capped at 2,000, 16 uniform vertex labels,
one edge label), 5% of the edges held out and streamed as 95% insertions / 5% deletions in batches
of 1,048,576 edges, with six standing queries of 8-10 vertices (2-3 tree leaves, every other vertex
closing a cycle) matched against every batch.  The same measurement on the 400K-vertex / 3.4M-edge
workload (batches of 65,536 edges, six 6-8-vertex queries) is reported beside it under "c2".
Inputs are resident in HBM before the timed region.  Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gcsm-bu_amd")
TOOL = os.path.join(PKG, "bin", "csm_tool")
ORACLE = os.path.join(ROOT, "oracle", "bin", "oracle")
METRIC = "incremental matches/s & update edges/s on 1 GPU; p50/p99 batch latency"
HBM_PEAK_GBS = 8000.0
FAMILIES = ["enumerate_count", "enumerate_write", "cartesian_product", "local_index", "batch_build", "apply_batch",
            "cartesian_prepare", "anchors", "level_count", "level_write", "level_prepare", "local_check", "merge_insert"]
# families whose kernels count algorithmic bytes on the device (the roofline candidates)
BYTE_FAMILIES = ["enumerate_count", "enumerate_write", "level_count", "level_write", "local_index", "local_check",
                 "merge_insert"]
# (the level families' launches are the device-chained kernels; levels the host resumes after a
# chain stop use k_level_flat<0>/<1>, the same kernel body, and are timed with them; the
# local_index family is one k_local_queue or k_local_expand launch)
KERNEL_OF = {"enumerate_count": "k_extend<0>", "enumerate_write": "k_extend<1>", "level_count": "k_level_flat_chain<0>",
             "level_write": "k_level_flat_chain<1>", "local_index": "k_local_queue", "local_check": "k_local_check",
             "merge_insert": "k_apply_insertions"}
SERVING_METRIC = "serving: update edges/s and p50/p99 latency of in-memory batches (host arrays, copy included)"
WORKLOADS = {
    "c3": "plain method, large/skewed: Chung-Lu 4M vertices / ~35M edges, gamma 2.0, 8 hubs of expected degree "
          "20,000 (others capped at 2,000), 16 uniform labels, 5% held out, 95/5 insert/delete stream, batches of "
          "1,048,576 edges, 6 queries of 8-10 vertices (2-3 tree leaves, the other vertices each closing a cycle)",
    "c2": "plain method on one GPU: Chung-Lu 400K vertices / ~3.4M edges, gamma 2.1, max degree ~3000, "
          "Zipf(1.0) labels over 32, 10% held out, 90/10 insert/delete stream, batches of 65,536 edges, "
          "6 queries of 6-8 vertices (3 sparse, 3 dense)",
    "c4": "multi-query: the c2 graph and stream (Chung-Lu 400K vertices / ~3.4M edges, batches of 65,536 edges) "
          "with 64 standing queries of 4-12 vertices (16 trees, 24 sparse cyclic, 24 dense)",
    "c5": "serving: directed, 3M vertices / ~25M arcs, lognormal(1.5, 1.2) out-degree, 4 vertex x 8 edge labels, "
          "10% held out, 85/15 insert/delete, lognormal batch sizes in [1K, 256K] (median 16K), "
          "32 queries of 4-8 vertices",
}


def header_default_threshold():
    """CSM_DEFAULT_SEMI_BFS_THRESHOLD from include/csm.h (the value a session uses when given 0)."""
    for line in open(os.path.join(ROOT, "include", "csm.h")):
        f = line.split()
        if len(f) == 3 and f[0] == "#define" and f[1] == "CSM_DEFAULT_SEMI_BFS_THRESHOLD":
            return int(f[2].rstrip("uU"))
    raise RuntimeError("CSM_DEFAULT_SEMI_BFS_THRESHOLD missing from include/csm.h")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_module(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG, name + ".py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def load_ffi():
    return load_module("csm_ffi")


def ensure_built():
    if not all(os.path.exists(p) for p in (TOOL, ORACLE, os.path.join(ROOT, "oracle", "bin", "baseline"),
                                            os.path.join(PKG, "lib", "libcsm.so"))):
        spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()


def workload(config, seed, data_dir, scale=1.0):
    meta_path = os.path.join(data_dir, "meta.txt")
    if not os.path.exists(meta_path):
        log(f"generating workload {config} seed {seed} scale {scale} into {data_dir}")
        t = time.time()
        subprocess.run([TOOL, "gen", config, data_dir, "--seed", str(seed), "--scale", str(scale)], check=True)
        log(f"generated in {time.time() - t:.1f}s")
    meta = dict(line.split() for line in open(meta_path))
    queries = [os.path.join(data_dir, f"q{k}.graph") for k in range(int(meta["queries"]))]
    return meta, queries


def open_session(ffi, data_dir, queries, batch, lanes=0, semi_bfs_threshold=0):
    s = ffi.Session(device=0, lanes=lanes, semi_bfs_threshold=semi_bfs_threshold)
    s.load_graph(os.path.join(data_dir, "data.graph"))
    for q in queries:
        s.add_query(q)
    nb = s.open_updates(os.path.join(data_dir, "update.stream"), batch)
    s.build()
    for q in s.queries:
        s.plan(q)
    return s, nb


def serving(ffi, data_dir, queries, meta, max_batches):
    """In-memory batches of the directed serving workload: lognormal batch sizes from batches.txt,
    each handed over as host arrays (csm_apply_edges: host-to-device copy, arena admission, index
    update, matching of every standing query).  Latency is wall clock per batch."""
    import numpy as np
    cols = {"src": [], "dst": [], "lab": [], "op": []}
    for line in open(os.path.join(data_dir, "update.stream")):
        t = line.split()
        if t and t[0] in ("e", "-e"):
            cols["src"].append(int(t[1])); cols["dst"].append(int(t[2])); cols["lab"].append(int(t[3]))
            cols["op"].append(1 if t[0] == "-e" else 0)
    src, dst = np.array(cols["src"], dtype=np.uint32), np.array(cols["dst"], dtype=np.uint32)
    lab, op = np.array(cols["lab"], dtype=np.uint32), np.array(cols["op"], dtype=np.uint8)
    sizes = [int(x) for x in open(os.path.join(data_dir, "batches.txt"))][:max_batches]
    s = ffi.Session(device=0, directed=True, max_batch_edges=int(meta["batch_max"]))
    s.load_graph(os.path.join(data_dir, "data.graph"))
    for q in queries:
        s.add_query(q)
    s.build()
    for q in s.queries:
        s.plan(q)
    lat, edges, matches, lo = [], 0, 0, 0
    for b, x in enumerate(sizes):
        t0 = time.perf_counter()
        res, info = s.apply_edges(src[lo:lo + x], dst[lo:lo + x], lab[lo:lo + x], op[lo:lo + x])
        s.synchronize()
        if b > 0:  # batch 0 warms up
            lat.append(time.perf_counter() - t0)
            edges += x
            matches += sum(r[0] + r[1] for r in res)
        lo += x
    used, cap = s.arena_usage()
    s.close()
    lat_ms = sorted(v * 1e3 for v in lat)
    total = sum(lat)
    return {"metric": SERVING_METRIC, "value": edges / total, "unit": "update edges/s", "batches": len(lat),
            "batch_sizes": {"min": min(sizes[1:]), "median": sorted(sizes[1:])[len(sizes[1:]) // 2], "max": max(sizes[1:])},
            "p50_batch_ms": statistics.median(lat_ms),
            "p99_batch_ms": lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))],
            "matches_per_s": matches / total, "queries": len(queries), "arena_words_used": used, "arena_words": cap,
            "config": {"workload": WORKLOADS[meta["config"]], "seed": int(meta["seed"])}}


def traffic_of(kernel, config):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (lines "traffic <kernel> <bytes per dispatch>"), or None."""
    path = os.path.join(ROOT, "profiles", f"{config}_kernel_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
            return int(f[2])
    return None


def roofline_pass(ffi, data_dir, queries, batch, config):
    """The dominant enumeration kernel timed on one lane (one stream, no other stream's packets
    between its HIP events): warm-up batch 0, back to the initial graph, then batches 0-1 timed;
    algorithmic bytes / event time."""
    mask = sum(1 << FAMILIES.index(k) for k in BYTE_FAMILIES)
    s, nb = open_session(ffi, data_dir, queries, batch, lanes=1)
    s.snapshot()
    s.profile_enable(mask)
    s.apply_batch(0)
    s.synchronize()
    s.restore()
    s.profile_enable(mask)  # counters from here on
    for b in range(min(2, nb)):
        s.apply_batch(b)
    s.synchronize()
    stats = {k: s.profile_read(FAMILIES.index(k)) for k in BYTE_FAMILIES}
    s.close()

    def gbs(k):
        _, ms, nbytes = stats[k]
        return nbytes / (ms / 1e3) / 1e9 if ms > 0 else 0.0

    dom = max(stats, key=lambda k: stats[k][1])
    launches, ms, nbytes = stats[dom]
    achieved = gbs(dom)
    traffic = traffic_of(KERNEL_OF[dom], config)
    others = {KERNEL_OF[k]: {"ms": round(stats[k][1], 3), "launches": stats[k][0], "achieved_GBs": round(gbs(k), 1),
                             "frac": round(gbs(k) / HBM_PEAK_GBS, 5)} for k in BYTE_FAMILIES if stats[k][0]}
    return {"bound": "hbm", "kernel": KERNEL_OF[dom], "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
            "avg_launch_us": round(ms * 1e3 / launches, 3) if launches else None, "launches": launches,
            "algorithmic_bytes_per_launch": nbytes // launches if launches else 0, "kernels": others,
            "measured": "one-lane session, batches 0-1 after a warm-up batch 0 and a restore, HIP events on the lane's stream",
            "traffic_source": f"profiles/{config}_kernel_traffic.md" if traffic is not None else None}


BASELINE = os.path.join(ROOT, "oracle", "bin", "baseline")


def cpu_baseline(data_dir, queries, batch, seconds):
    """The multithreaded CPU backtracking enumerator over the planner's matching orders (the
    PrintOrders text each session wrote beside the queries), on a uniform random sample of batch
    0's (phase, query, step, anchor) items: threads take items in a seeded random order and stop
    taking new ones after `seconds`.  Host cores: the box's CPU share (at most 16 threads)."""
    threads = min(16, os.cpu_count() or 1)
    cmd = [BASELINE, os.path.join(data_dir, "data.graph"), os.path.join(data_dir, "update.stream"), str(batch)]
    for q in queries:
        cmd += [q, q + ".orders"]
    cmd += ["--threads", str(threads), "--max-batches", "1", "--time-limit-ms", str(int(seconds * 1000))]
    out = subprocess.run(cmd, capture_output=True, text=True, check=True).stdout.split()
    # batch <b> items <done> of <total> matches <count> ms <wall>
    done, total, matches, ms = int(out[3]), int(out[5]), int(out[7]), float(out[9])
    return {"value": matches / (ms / 1e3) if ms > 0 else None, "unit": "matches/s", "cores": threads, "kind": "port",
            "sample": f"multithreaded backtracking (std::thread x{threads}) over the planner's matching orders, anchored "
                      f"at each step's update edge, on batch 0 of the same workload: {done} of {total} (phase, query, "
                      f"step, anchor) items taken in seeded random order until {seconds:.0f}s, {matches} matches in "
                      f"{ms / 1e3:.1f}s"}


def timed_batches(ffi, data_dir, queries, batch, steps, warmup, opts, profile_mask=0, label=""):
    """Warm-up batches untimed, then `steps` timed batches (wall clock around csm_apply_batch plus a
    synchronisation), cycling through the stream; when it runs out, the session returns to the
    initial graph (csm_restore, a device copy) outside the timed region."""
    t = time.time()
    s, nb = open_session(ffi, data_dir, queries, batch, **opts)
    s.snapshot()  # the initial graph, restored whenever the stream runs out
    for q, path in zip(s.queries, queries):  # the plans, for the CPU baseline's matching orders
        with open(path + ".orders", "w") as f:
            f.write(s.plan_text(q))
    log(f"{label}loaded + indexed + planned in {time.time() - t:.1f}s, {nb} batches, pool {s.pool_usage()[0] >> 20} MiB")
    s.profile_enable(profile_mask)
    nxt = 0

    def rewind_if_done():
        nonlocal nxt
        if nxt >= nb:
            s.restore()
            nxt = 0

    for w in range(warmup):
        rewind_if_done()
        t0 = time.perf_counter()
        s.apply_batch(nxt)
        s.synchronize()
        log(f"{label}warmup {w} (batch {nxt}) done in {(time.perf_counter() - t0) * 1e3:.1f} ms")
        nxt += 1
    s.profile_enable(profile_mask)  # reset counters: only timed batches below
    lat, matches, edges, per_query = [], 0, 0, [0] * len(queries)
    s.synchronize()
    for k in range(steps):
        rewind_if_done()
        b = nxt
        t0 = time.perf_counter()
        res, info = s.apply_batch(b)
        s.synchronize()
        dt = time.perf_counter() - t0
        nxt += 1
        lat.append(dt)
        for i, r in enumerate(res):
            matches += r[0] + r[1]
            per_query[i] += r[0] + r[1]
        edges += info.edges
        log(f"{label}step {k} (batch {b}): {dt * 1e3:.1f} ms, {sum(r[0] + r[1] for r in res)} matches, {info.edges} edges")
    stats = {FAMILIES[f]: s.profile_read(f) for f in range(len(FAMILIES)) if profile_mask >> f & 1}
    s.close()
    total = sum(lat)
    lat_ms = sorted(x * 1e3 for x in lat)
    return {"value": matches / total if total > 0 else 0.0, "ms_per_step": total * 1e3 / max(1, steps),
            "update_edges_per_s": edges / total if total > 0 else 0.0, "p50_batch_ms": statistics.median(lat_ms),
            "p99_batch_ms": lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))],
            "matches": matches, "per_query_matches": per_query, "batches_in_stream": nb,
            "kernel_ms": {k: round(v[1], 3) for k, v in stats.items()}}


def plain_path(ffi, data_dir, queries, batch):
    """The standing queries matched the reference's way, one query per session (per process in the
    reference: its descriptors are __constant__ globals, match_gpu.cu:20-24), against the batched
    form: one session holding every query.  Both apply warm-up batch 0 and then batches 1.. of the
    stream; a batch's plain latency is the sum of the single-query sessions' times for it.  Every
    query's per-batch counts and hashes must be identical in the two forms."""
    s, nb = open_session(ffi, data_dir, queries, batch)
    s.apply_batch(0)
    s.synchronize()
    multi_ms, multi = [], []
    for b in range(1, nb):
        t0 = time.perf_counter()
        res, _ = s.apply_batch(b)
        s.synchronize()
        multi_ms.append((time.perf_counter() - t0) * 1e3)
        multi.append(res)
    s.close()
    plain_ms = [0.0] * (nb - 1)
    for k, q in enumerate(queries):
        s, _ = open_session(ffi, data_dir, [q], batch)
        s.apply_batch(0)
        s.synchronize()
        for b in range(1, nb):
            t0 = time.perf_counter()
            res, _ = s.apply_batch(b)
            s.synchronize()
            plain_ms[b - 1] += (time.perf_counter() - t0) * 1e3
            if res[0] != multi[b - 1][k]:
                raise RuntimeError(f"query {k} batch {b}: one-query session {res[0]} != batched {multi[b - 1][k]}")
        s.close()
        log(f"plain path: query {k} done")
    matches = sum(r[0] + r[1] for res in multi for r in res)
    return {"batches": f"1-{nb - 1} after warm-up batch 0", "queries": len(queries),
            "batched_ms_per_batch": sum(multi_ms) / len(multi_ms), "plain_ms_per_batch": sum(plain_ms) / len(plain_ms),
            "batched_p99_ms": max(multi_ms), "plain_p99_ms": max(plain_ms),
            "batched_matches_per_s": matches / (sum(multi_ms) / 1e3), "plain_matches_per_s": matches / (sum(plain_ms) / 1e3),
            "speedup": sum(plain_ms) / sum(multi_ms), "results": "every query's per-batch counts and hashes identical"}


def config_block(config, meta, args, opts, scale):
    return {"workload": WORKLOADS[config], "config": config, "seed": args.seed, "scale": scale,
            "batch_edges": int(meta["batch_size"]), "queries": int(meta["queries"]),
            "technique": "GCSM-BU defaults (order 2, semi-BFS, local index, CP, heavy-vertex avoidance, relation switch)",
            "lanes": opts.get("lanes", int(os.environ.get("CSM_LANES", "8"))),
            "semi_bfs_threshold": opts.get("semi_bfs_threshold") or header_default_threshold()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c3", choices=["c3", "c2", "c4"],
                    help="c3: the headline workload (c2 is measured beside it unless --no-c2); c2 alone; c4: 64 "
                         "standing queries on the c2 graph and stream, also timed with one lane for comparison")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="graph and stream scale of the c4 workload (c2, c3 and serving always run at full size)")
    ap.add_argument("--data-dir", default=None)
    ap.add_argument("--no-c2", action="store_true", help="with --config c3: skip the c2 measurement beside it")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-mask", type=int, default=0,
                    help="kernel families to time with HIP events during the timed batches (off by default: "
                         "the roofline comes from a separate one-lane pass)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain copy kernel of 2^26 words (HBM traffic counter calibration)")
    ap.add_argument("--serving", action="store_true",
                    help="run the in-memory serving workload (c5) instead and print its own JSON line")
    ap.add_argument("--serving-batches", type=int, default=64)
    ap.add_argument("--autotune", action="store_true",
                    help="pick lanes and the semi-BFS threshold first: one fresh session per candidate on "
                         "batches 0-2 (batch 0 untimed), results checked equal; then measure with the fastest")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the one-lane roofline pass (what the committed kernel profiles cover)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU: the matching path is single-GPU "
                                   "(multi-GPU sharding by anchor edges is a design note, not built)"}))
        return 2
    ensure_built()
    ffi = load_ffi()
    if args.serving:
        args.config = "c5"
    if args.config != "c4":
        args.scale = 1.0

    def data_dir_of(config):
        tag = "" if args.scale == 1.0 else f"_x{args.scale:g}"
        if args.data_dir and config == args.config:
            return args.data_dir
        return os.path.join("/tmp", f"csm_bench_{config}_s{args.seed}{tag}")

    data_dir = data_dir_of(args.config)
    meta, queries = workload(args.config, args.seed, data_dir, args.scale)
    batch = int(meta["batch_size"])
    if args.serving:
        print(json.dumps(serving(ffi, data_dir, queries, meta, args.serving_batches + 1)))
        return 0
    if args.roofline_only:
        r = roofline_pass(ffi, data_dir, queries, batch, args.config)
        if args.calibrate:
            s, _ = open_session(ffi, data_dir, queries[:1], batch, lanes=1)
            r["calibration_copy_ms"] = s.calibration_copy(1 << 26)
            s.close()
        print(json.dumps({"roofline": r}))
        return 0

    opts, tuned = {}, None
    if args.autotune:
        at = load_module("autotune")
        t = time.time()
        best, table = at.tune(lambda o: open_session(ffi, data_dir, queries, batch, **o), warmup=1, batches=2)
        opts = best.options()
        tuned = {"chosen": opts, "candidates": table, "batches": "1-2 after warm-up batch 0"}
        log(f"autotuned in {time.time() - t:.1f}s: {opts}")
    m = timed_batches(ffi, data_dir, queries, batch, args.steps, args.warmup, opts, args.profile_mask)
    calib = None
    if args.calibrate:
        s, _ = open_session(ffi, data_dir, queries[:1], batch, lanes=1)
        words = 1 << 26
        ms = s.calibration_copy(words)
        s.close()
        calib = {"kernel": "k_copy_u32", "bytes_read": 4 * words, "bytes_written": 4 * words, "ms": ms,
                 "GB/s": round(8 * words / (ms / 1e3) / 1e9, 1)}
    one_lane = plain = None
    if args.config == "c4":  # the same batches with one lane: every (query, edge) step in turn
        m1 = timed_batches(ffi, data_dir, queries, batch, args.steps, args.warmup, {"lanes": 1}, label="one lane: ")
        one_lane = {"ms_per_step": m1["ms_per_step"], "p99_batch_ms": m1["p99_batch_ms"], "lanes": 1}
        plain = plain_path(ffi, data_dir, queries, batch)
    roofline = roofline_pass(ffi, data_dir, queries, batch, args.config)
    cpu = None if args.no_cpu_baseline else cpu_baseline(data_dir, queries, batch, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": m["value"], "unit": "matches/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": config_block(args.config, meta, args, opts, args.scale),
        "update_edges_per_s": m["update_edges_per_s"], "p50_batch_ms": m["p50_batch_ms"], "p99_batch_ms": m["p99_batch_ms"],
        "matches": m["matches"], "per_query_matches": m["per_query_matches"],
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if one_lane:
        out["one_lane"] = one_lane
    if plain:
        out["plain_path"] = plain
    if tuned:
        out["autotune"] = tuned
    if m["kernel_ms"]:
        out["kernel_ms"] = m["kernel_ms"]
    if calib:
        out["calibration_copy"] = calib
    if args.config == "c3" and not args.no_c2:  # the 400K-vertex workload, same measurement
        d2 = data_dir_of("c2")
        meta2, queries2 = workload("c2", args.seed, d2)
        b2 = int(meta2["batch_size"])
        m2 = timed_batches(ffi, d2, queries2, b2, args.steps, args.warmup, {}, label="c2: ")
        out["c2"] = {"value": m2["value"], "unit": "matches/s", "ms_per_step": m2["ms_per_step"],
                     "update_edges_per_s": m2["update_edges_per_s"], "p50_batch_ms": m2["p50_batch_ms"],
                     "p99_batch_ms": m2["p99_batch_ms"], "matches": m2["matches"],
                     "config": config_block("c2", meta2, args, {}, 1.0),
                     "roofline": roofline_pass(ffi, d2, queries2, b2, "c2")}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
