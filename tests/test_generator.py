"""The workload generator's exact, budgeted embedding count (csm_tool count, the counter that
selects the C3 and C4 queries) against the oracle's brute force, on undirected, edge-labelled and
directed graphs; and the selection's recorded counts against a recount of the union graph."""
import os

import pytest

import util


def tool_count(data, query, *extra):
    out = util.run([util.TOOL, "count", str(data), str(query), "--threads", "4"] + list(extra)).split()
    assert out[0] == "within" and out[2] == "count"
    return int(out[1]), int(out[3])


@pytest.mark.parametrize("config,scale,seed", [("tiny", None, 1), ("tiny", None, 2), ("tinyins", None, 3), ("c1", None, 1),
                                               ("c2", 0.005, 2), ("c4", 0.004, 1), ("c5", 0.0005, 1)])
def test_count_equals_brute_force(tmp_path, config, scale, seed):
    util.ensure_built()
    meta, queries = util.gen(config, tmp_path, seed=seed, scale=scale)
    directed = meta.get("directed") == "1"
    data = tmp_path / "data.graph"
    for q in queries:
        within, got = tool_count(data, q, *(["--directed"] if directed else []))
        want, _ = util.oracle_count(data, q, directed=directed)
        assert within == 1 and got == want, f"{q}"


def test_cap_and_budget_stop_the_count(tmp_path):
    util.ensure_built()
    meta, queries = util.gen("c1", tmp_path, seed=1)
    data = tmp_path / "data.graph"
    _, full = tool_count(data, queries[0])
    assert full > 10
    within, partial = tool_count(data, queries[0], "--cap", str(full - 1))
    assert within == 0 and partial > full - 1
    within, _ = tool_count(data, queries[0], "--cap", str(full))
    assert within == 1
    within, _ = tool_count(data, queries[0], "--budget", "10")
    assert within == 0


def test_selected_queries_recount(tmp_path):
    """C3 at 1% scale: each kept query's recorded count (union of the initial graph and every
    insertion) is within the cap and equals a recount of that union graph."""
    util.ensure_built()
    meta, queries = util.gen("c3", tmp_path, seed=1, scale=0.01)
    union = tmp_path / "union.graph"
    with open(union, "w") as f:
        for line in open(tmp_path / "data.graph"):
            f.write(line)
        for line in open(tmp_path / "update.stream"):
            if line.startswith("e "):
                f.write(line)
    cap = float(meta["embedding_cap"])
    for k, q in enumerate(queries):
        rec = int(meta[f"union_embeddings_q{k}"])
        assert 0 < rec <= cap
        assert tool_count(union, q) == (1, rec)
