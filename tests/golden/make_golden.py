"""Writes the golden fixtures under tests/golden from the reference-built planner/loader
(oracle/_ref/ref_plan, compiled from the reference's own graph.cpp/plan.cpp by oracle/Makefile.ref).

Run in the container that has the reference tree:  python tests/golden/make_golden.py
Outputs (committed):
  queries/*.graph        query fixtures: a sample of the reference's shipped query files (data) and
                         seeded random queries up to 16 vertices / 42 edges
  summaries/*.txt        cardinality / average-degree inputs for the planner ("r card avg_bits")
  plans.json.gz          ref_plan dump for every (query, summary, matching order)
  ref_log_az_dense_q0.txt  query-only lines of the reference's shipped log
                         (output/gamma_output_az_dense_q1.txt:3-8 eidx print, 15-67 Orders block)
  grouping/              a small insert-only workload and the reference's per-relation grouping dump
"""
import gzip
import json
import os
import random
import shutil
import struct
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
REF_PLAN = os.path.join(ROOT, "oracle", "_ref", "ref_plan")
TOOL = os.path.join(ROOT, "gcsm-bu_amd", "bin", "csm_tool")


def f32bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def write_summary(path, rng, kind):
    with open(path, "w") as f:
        for r in range(84):
            if kind == "ties":
                c, ns = rng.choice([5, 10]), rng.choice([1, 2, 5])
            elif kind == "zeros":
                c = 0 if rng.random() < 0.4 else rng.randint(1, 100)
                ns = 0 if c == 0 else rng.randint(1, c)
            elif kind == "large":
                c = rng.randint(1, 10 ** 9)
                ns = rng.randint(1, min(c, 10 ** 7))
            else:
                c, ns = rng.choice([1, 3, 2000, 5000, 3000000]), 1
            avg = float("nan") if ns == 0 else struct.unpack("<f", struct.pack("<f", float(c)))[0] / float(ns)
            avg = struct.unpack("<f", struct.pack("<f", avg))[0]
            f.write(f"{r} {c} {f32bits(avg):x}\n")


def random_query(rng, n, m, nvl, nel):
    edges = set()
    for v in range(1, n):  # random spanning tree
        edges.add((rng.randrange(v), v))
    all_pairs = [(a, b) for a in range(n) for b in range(a + 1, n) if (a, b) not in edges]
    rng.shuffle(all_pairs)
    for p in all_pairs[: max(0, m - (n - 1))]:
        edges.add(p)
    lines = [f"v {v} {rng.randrange(nvl)}" for v in range(n)]
    lines += [f"e {a} {b} {rng.randrange(nel)}" for a, b in sorted(edges)]
    return "t # 0\n" + "\n".join(lines) + "\n"


def main():
    if not os.path.exists(REF_PLAN):
        sys.exit("build oracle/_ref/ref_plan first (make -f oracle/Makefile.ref)")
    qdir = os.path.join(HERE, "queries")
    sdir = os.path.join(HERE, "summaries")
    for d in (qdir, sdir):
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
    rng = random.Random(20261202)
    names = []
    for ds, sub in [("amazon", "6"), ("livejournal", "30"), ("lsbench", "x1"), ("netflow", "")]:
        for shape in ["dense_6", "sparse_6", "tree_6"]:
            base = os.path.join(REF, "datasets", ds, sub, "query_graph", shape)
            for qname in ["Q_0", f"Q_{rng.randrange(1, 100)}"]:
                dst = f"{ds}_{shape}_{qname}.graph"
                shutil.copy(os.path.join(base, qname), os.path.join(qdir, dst))
                names.append(dst)
    for k in range(16):
        n = rng.randint(2, 16)
        m = rng.randint(n - 1, min(42, n * (n - 1) // 2))
        dst = f"random_{k}.graph"
        with open(os.path.join(qdir, dst), "w") as f:
            f.write(random_query(rng, n, m, rng.choice([1, 2, 3, 6]), rng.choice([1, 1, 2, 3])))
        names.append(dst)
    sums = ["-"]
    for kind in ["ties", "zeros", "large", "spread"]:
        p = f"{kind}.txt"
        write_summary(os.path.join(sdir, p), rng, kind)
        sums.append(p)
    plans = {}
    for q in names:
        for s in sums:
            spath = "-" if s == "-" else os.path.join(sdir, s)
            for order in range(4):
                r = subprocess.run([REF_PLAN, "plan", os.path.join(qdir, q), spath, str(order)], capture_output=True, text=True)
                plans[f"{q}|{s}|{order}"] = r.stdout if r.returncode == 0 else f"ERROR rc={r.returncode}"
    with gzip.open(os.path.join(HERE, "plans.json.gz"), "wt") as f:
        json.dump(plans, f, sort_keys=True)
    # query-only lines of the reference's shipped log
    log = open(os.path.join(REF, "output", "gamma_output_az_dense_q1.txt")).read().splitlines()
    with open(os.path.join(HERE, "ref_log_az_dense_q0.txt"), "w") as f:
        f.write("\n".join(log[2:8] + log[14:67]) + "\n")
    shutil.copy(os.path.join(REF, "datasets", "amazon", "6", "query_graph", "dense_6", "Q_0"), os.path.join(HERE, "az_dense_6_Q_0.graph"))
    # grouping fixture: a small insert-only workload
    gdir = os.path.join(HERE, "grouping")
    shutil.rmtree(gdir, ignore_errors=True)
    subprocess.run([TOOL, "gen", "tinyins", gdir, "--seed", "11"], check=True)
    out = {}
    for q in sorted(f for f in os.listdir(gdir) if f.startswith("q")):
        for b in [64, 300, 1000]:
            r = subprocess.run([REF_PLAN, "load", os.path.join(gdir, q), os.path.join(gdir, "data.graph"),
                                os.path.join(gdir, "update.stream"), str(b)], capture_output=True, text=True, check=True)
            out[f"{q}|{b}"] = r.stdout
    with open(os.path.join(gdir, "expected.json"), "w") as f:
        json.dump(out, f, sort_keys=True, indent=0)
    print(f"{len(plans)} plan dumps, {len(out)} grouping dumps")


if __name__ == "__main__":
    main()
