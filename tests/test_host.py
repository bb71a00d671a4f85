"""Host-side parsing and loading (H1-H4) and the workload generator."""
import filecmp
import json
import os
import subprocess

import pytest

import util

G = util.GOLDEN


def test_grouping_matches_reference_loader():
    """Per-relation grouping of the data and update files equals LoadInitial/LoadUpdate
    (graph.cpp:359-563) on an insert-only workload, for several batch sizes."""
    util.ensure_built()
    gd = os.path.join(G, "grouping")
    want = json.load(open(os.path.join(gd, "expected.json")))
    for key, dump in want.items():
        q, b = key.split("|")
        got = subprocess.run([util.TOOL, "load", os.path.join(gd, q), os.path.join(gd, "data.graph"),
                              os.path.join(gd, "update.stream"), b], capture_output=True, text=True, check=True).stdout
        assert got == dump, key


@pytest.mark.parametrize("text,err", [
    ("v 0 1\n", "less than 2"),
    ("v 0 1\nv 1 1\nv 2 1\ne 0 1 0\n", "connected"),
    ("v 0 1\nv 1 1\ne 0 0 0\n", "self-loop"),
    ("v 0 1\nv 1 1\ne 0 1 0\ne 1 0 0\n", "duplicate"),
    ("v 0 1\nv 2 1\ne 0 2 0\n", "contiguous"),
    ("v 0 1\nv 0 1\n", "declared twice"),
    ("".join(f"v {i} 0\n" for i in range(17)), "out of range"),
])
def test_query_errors(tmp_path, text, err):
    util.ensure_built()
    p = tmp_path / "q.graph"
    p.write_text(text)
    r = subprocess.run([util.TOOL, "meta", str(p)], capture_output=True, text=True)
    assert r.returncode == 1 and err in r.stderr


def test_query_limit_42_edges(tmp_path):
    util.ensure_built()
    n = 16
    lines = [f"v {i} 0\n" for i in range(n)]
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)][:43]
    lines += [f"e {a} {b} 0\n" for a, b in pairs]
    p = tmp_path / "q.graph"
    p.write_text("".join(lines))
    r = subprocess.run([util.TOOL, "meta", str(p)], capture_output=True, text=True)
    assert r.returncode == 1 and "at most 16 vertices and 42 edges" in r.stderr


def test_update_file_rejects_vertex_lines(tmp_path):
    util.ensure_built()
    gd = os.path.join(G, "grouping")
    u = tmp_path / "u.stream"
    u.write_text("e 1 2 0\nv 3 1\n")
    r = subprocess.run([util.TOOL, "load", os.path.join(gd, "q0.graph"), os.path.join(gd, "data.graph"), str(u), "10"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "vertex update" in r.stderr


def test_generator_is_seeded(tmp_path):
    util.ensure_built()
    a, b, c = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    util.gen("tiny", a, seed=7)
    util.gen("tiny", b, seed=7)
    util.gen("tiny", c, seed=8)
    for f in ["data.graph", "update.stream", "q0.graph", "q1.graph", "q2.graph"]:
        assert filecmp.cmp(a / f, b / f, shallow=False)
    assert not filecmp.cmp(a / "data.graph", c / "data.graph", shallow=False)


def test_generator_stream_preconditions(tmp_path):
    """Deletions remove existing edges, insertions add new ones, nothing is deleted and re-inserted."""
    util.ensure_built()
    meta, _ = util.gen("tiny", tmp_path, seed=9)
    edges = set()
    for line in open(tmp_path / "data.graph"):
        t = line.split()
        if t and t[0] == "e":
            edges.add((min(int(t[1]), int(t[2])), max(int(t[1]), int(t[2]))))
    seen = set()
    for line in open(tmp_path / "update.stream"):
        t = line.split()
        k = (min(int(t[1]), int(t[2])), max(int(t[1]), int(t[2])))
        assert k not in seen
        seen.add(k)
        if t[0] == "-e":
            assert k in edges
            edges.remove(k)
        else:
            assert k not in edges
            edges.add(k)
    assert int(meta["deletes"]) > 0 and int(meta["inserts"]) > 0
