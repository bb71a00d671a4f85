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
# (the level families' launches are a team's device-chained multi-step kernels; levels the host
# resumes after a chain stop use k_level_flat<0>/<1>, the same kernel body, and are timed with
# them; the local_index family is one k_local_queue_multi launch per indexing position)
KERNEL_OF = {"enumerate_count": "k_extend<0>", "enumerate_write": "k_extend<1>", "level_count": "k_level_flat_chain_multi<0>",
             "level_write": "k_level_flat_chain_multi<1>", "local_index": "k_local_queue_multi",
             "local_check": "k_local_check_multi", "merge_insert": "k_apply_insertions"}
ROOFLINE_LANES = 8  # one team of the default size: one stream, so HIP events bracket one kernel
SERVING_METRIC = ("serving: update edges/s and p50/p99 arrival-to-completion latency of in-memory batches arriving as a "
                  "Poisson process (host arrays, copy included)")
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


# Queries the generator selected earlier for a configuration at full size (its exact, budgeted count
# per candidate takes minutes there); the graph and stream are regenerated from the seed.
SELECTED = {("c4", 1): os.path.join(PKG, "workloads", "c4_seed1")}


def workload(config, seed, data_dir, scale=1.0):
    meta_path = os.path.join(data_dir, "meta.txt")
    if not os.path.exists(meta_path):
        log(f"generating workload {config} seed {seed} scale {scale} into {data_dir}")
        t = time.time()
        cmd = [TOOL, "gen", config, data_dir, "--seed", str(seed), "--scale", str(scale)]
        if scale == 1.0 and (config, seed) in SELECTED:
            cmd += ["--queries-from", SELECTED[(config, seed)]]
        subprocess.run(cmd, check=True)
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


def serving(ffi, data_dir, queries, meta, max_batches, load=0.75, seed=7):
    """The directed serving workload's in-memory batches (lognormal sizes from batches.txt) arriving
    as a seeded Poisson process, served two ways on fresh sessions over the same graph:
      - request loop (continuous batching, csm_serve_*): each batch is submitted at its arrival
        time; the next request is validated, keyed and copied to the device while the current one
        matches, and enters a staging slot as soon as one frees up;
      - plain per-batch path: one csm_apply_edges call per batch, in arrival order, each starting
        when it has arrived and the previous call has returned.
    Latency is arrival to completion (queueing included).  The arrival rate is `load` times the
    plain path's service rate, measured on the first batches beforehand.  Every request's results
    must be identical in both runs.  Batch 0 warms each session up and is not timed."""
    import threading
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
    batches, lo = [], 0
    for x in sizes:
        batches.append((src[lo:lo + x], dst[lo:lo + x], lab[lo:lo + x], op[lo:lo + x]))
        lo += x

    def session():
        s = ffi.Session(device=0, directed=True, max_batch_edges=int(meta["batch_max"]))
        s.load_graph(os.path.join(data_dir, "data.graph"))
        for q in queries:
            s.add_query(q)
        s.build()
        for q in s.queries:
            s.plan(q)
        s.apply_edges(*batches[0])  # warm-up
        s.synchronize()
        return s

    work = batches[1:]
    # service time of the plain path on a prefix (its own session, discarded)
    s = session()
    t0 = time.perf_counter()
    k = min(8, len(work))
    for b in work[:k]:
        s.apply_edges(*b)
    s.synchronize()
    mean_service = (time.perf_counter() - t0) / k
    s.close()
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(mean_service / load, size=len(work))
    arrivals = np.cumsum(gaps) - gaps[0]  # seconds after the first arrival
    log(f"serving: {len(work)} requests, mean plain service {mean_service * 1e3:.2f} ms, Poisson arrivals at load {load}")

    # plain per-batch path
    s = session()
    plain_lat, plain_res = [], []
    start = time.perf_counter()
    for a, b in zip(arrivals, work):
        now = time.perf_counter() - start
        if now < a:
            time.sleep(a - now)
        res, _ = s.apply_edges(*b)
        s.synchronize()
        plain_lat.append((time.perf_counter() - start - a) * 1e3)
        plain_res.append(res)
    plain_wall = time.perf_counter() - start
    s.close()
    log(f"serving: plain path done in {plain_wall:.2f}s")

    # request loop
    s = session()
    s.serve_start()
    tickets = [None] * len(work)
    submit_at = [0.0] * len(work)
    start = time.perf_counter()

    def submitter():
        for i, (a, b) in enumerate(zip(arrivals, work)):
            now = time.perf_counter() - start
            if now < a:
                time.sleep(a - now)
            submit_at[i] = time.perf_counter() - start
            tickets[i] = s.serve_submit(*b)

    th = threading.Thread(target=submitter)
    th.start()
    loop_lat, loop_res, matches = [], [], 0
    for i in range(len(work)):
        while tickets[i] is None:
            time.sleep(0.0002)
        res, info, ms = s.serve_wait(tickets[i])
        # latency from arrival: the library's submit-to-done time plus any submission lag
        loop_lat.append(ms + (submit_at[i] - arrivals[i]) * 1e3)
        loop_res.append(res)
        matches += sum(r[0] + r[1] for r in res)
    loop_wall = time.perf_counter() - start
    th.join()
    s.serve_stop()
    used, cap = s.arena_usage()
    s.close()
    if loop_res != plain_res:
        raise RuntimeError("request loop results differ from the plain per-batch path")
    edges = sum(len(b[0]) for b in work)

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]

    return {"metric": SERVING_METRIC, "value": edges / loop_wall, "unit": "update edges/s", "requests": len(work),
            "arrivals": f"Poisson, rate {load} x the plain path's service rate ({1 / mean_service * load:.1f} batches/s), seed {seed}",
            "batch_sizes": {"min": min(sizes[1:]), "median": sorted(sizes[1:])[len(sizes[1:]) // 2], "max": max(sizes[1:])},
            "p50_batch_ms": pct(loop_lat, 0.5), "p99_batch_ms": pct(loop_lat, 0.99), "matches_per_s": matches / loop_wall,
            "plain_path": {"update_edges_per_s": edges / plain_wall, "p50_batch_ms": pct(plain_lat, 0.5),
                           "p99_batch_ms": pct(plain_lat, 0.99), "matches_per_s": matches / plain_wall},
            "results": "every request identical in both runs", "queries": len(queries),
            "arena_words_used": used, "arena_words": cap,
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
    """The dominant byte-counting kernel timed in a session of one team (eight lanes on one stream:
    the multi-step kernels of the timed run, with no other stream's packets between their HIP
    events): warm-up batch 0, back to the initial graph, then batches 0-1 timed; algorithmic bytes
    / event time."""
    mask = sum(1 << FAMILIES.index(k) for k in BYTE_FAMILIES)
    s, nb = open_session(ffi, data_dir, queries, batch, lanes=ROOFLINE_LANES)
    s.set_tunables(team=ROOFLINE_LANES)  # one team: the size of a team of the default 32 lanes
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
            "measured": "session of one team of eight lanes (one stream), batches 0-1 after a warm-up batch 0 and a restore, "
                        "HIP events on the team's stream",
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


def timed_batches(ffi, data_dir, queries, batch, steps, warmup, opts, profile_mask=0, label="", tune=None):
    """Warm-up batches untimed, then `steps` timed batches (wall clock around csm_apply_batch plus a
    synchronisation), cycling through the stream; when it runs out, the session returns to the
    initial graph (csm_restore, a device copy) outside the timed region.  With `tune` (a dict), the
    library's autotuner (csm_tune, coordinate search on batches 0-2, batch 0 untimed) first picks
    the session's launch settings; its evaluations are stored in the dict."""
    t = time.time()
    s, nb = open_session(ffi, data_dir, queries, batch, **opts)
    s.snapshot()  # the initial graph, restored whenever the stream runs out
    if tune is not None:
        t1 = time.time()
        trail, best = s.tune(0, 0, warmup=1, batches=min(2, nb - 1))  # leaves the initial graph, takes its snapshot
        tune.update({"chosen": s.tunables(), "seconds": round(time.time() - t1, 1),
                     "evaluations": [{**st, "ms_per_batch": round(ms, 3)} for st, ms in trail], "best": best})
        log(f"{label}library autotuner: {len(trail)} evaluations in {time.time() - t1:.1f}s, chose {s.tunables()}")
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
    settings = s.tunables()
    s.close()
    total = sum(lat)
    lat_ms = sorted(x * 1e3 for x in lat)
    return {"value": matches / total if total > 0 else 0.0, "ms_per_step": total * 1e3 / max(1, steps),
            "update_edges_per_s": edges / total if total > 0 else 0.0, "p50_batch_ms": statistics.median(lat_ms),
            "p99_batch_ms": lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))],
            "matches": matches, "per_query_matches": per_query, "batches_in_stream": nb, "launch_settings": settings,
            "kernel_ms": {k: round(v[1], 3) for k, v in stats.items()}}


def plain_path(ffi, data_dir, queries, batch, opts=None):
    """The standing queries matched the reference's way, one query per session (per process in the
    reference: its descriptors are __constant__ globals, match_gpu.cu:20-24), against the batched
    form: one session holding every query.  Both apply warm-up batch 0 and then batches 1.. of the
    stream; a batch's plain latency is the sum of the single-query sessions' times for it.  Every
    query's per-batch counts and hashes must be identical in the two forms.  The batched session
    takes the measured run's options (`opts`); the one-query sessions the library defaults."""
    s, nb = open_session(ffi, data_dir, queries, batch, **(opts or {}))
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
            "lanes": opts.get("lanes", int(os.environ.get("CSM_LANES", "32"))),
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
    ap.add_argument("--lanes", type=int, default=0,
                    help="step lanes of the measured session (0: the library default, 32; with --config c4, 64: "
                         "64 standing queries have about 500 steps per phase, and 64 lanes in teams of 16 take "
                         "75 against 82 ms per batch)")
    ap.add_argument("--profile-mask", type=int, default=0,
                    help="kernel families to time with HIP events during the timed batches (off by default: "
                         "the roofline comes from a separate one-lane pass)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain copy kernel of 2^26 words (HBM traffic counter calibration)")
    ap.add_argument("--serving", action="store_true",
                    help="run the in-memory serving workload (c5) instead and print its own JSON line")
    ap.add_argument("--serving-batches", type=int, default=96)
    ap.add_argument("--serving-load", type=float, default=0.75,
                    help="arrival rate of the serving run as a fraction of the plain path's service rate")
    ap.add_argument("--autotune", action="store_true",
                    help="pick lanes first (one fresh session per lane count on batches 0-2, batch 0 untimed, "
                         "results checked equal), then the launch settings with the library's autotuner (csm_tune) "
                         "in the measuring session; then measure")
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
        print(json.dumps(serving(ffi, data_dir, queries, meta, args.serving_batches + 1, load=args.serving_load)))
        return 0
    if args.roofline_only:
        r = roofline_pass(ffi, data_dir, queries, batch, args.config)
        if args.calibrate:
            s, _ = open_session(ffi, data_dir, queries[:1], batch, lanes=1)
            r["calibration_copy_ms"] = s.calibration_copy(1 << 26)
            s.close()
        print(json.dumps({"roofline": r}))
        return 0

    opts, tuned, lib_tune = {}, None, None
    lanes = args.lanes or (64 if args.config == "c4" else 0)
    if lanes:
        opts = {"lanes": lanes}
    if args.autotune:
        at = load_module("autotune")
        t = time.time()
        thr = header_default_threshold()
        nb = -(-int(meta["stream"]) // batch)  # batches in the update stream
        best, table = at.tune(lambda o: open_session(ffi, data_dir, queries, batch, **o),
                              [at.Candidate(lanes, thr) for lanes in (16, 32, 64)], warmup=1, batches=max(1, min(2, nb - 1)))
        opts = {"lanes": best.lanes}
        lib_tune = {}
        tuned = {"lanes": {"chosen": best.lanes, "candidates": table, "batches": "1-2 after warm-up batch 0"},
                 "launch_settings": lib_tune}
        log(f"lanes autotuned in {time.time() - t:.1f}s: {opts}")
    m = timed_batches(ffi, data_dir, queries, batch, args.steps, args.warmup, opts, args.profile_mask, tune=lib_tune)
    if lib_tune:
        opts = {**opts, "semi_bfs_threshold": lib_tune["chosen"]["semi_bfs_threshold"]}
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
        plain = plain_path(ffi, data_dir, queries, batch, opts)
    roofline = roofline_pass(ffi, data_dir, queries, batch, args.config)
    cpu = None if args.no_cpu_baseline else cpu_baseline(data_dir, queries, batch, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": m["value"], "unit": "matches/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {**config_block(args.config, meta, args, opts, args.scale), "launch_settings": m["launch_settings"]},
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
