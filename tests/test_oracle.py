"""The oracle, pinned before it is trusted.

count:   brute force vs closed forms computed independently with numpy (edges, wedges, stars,
         triangles = trace(A^3), 4-cycles from trace(A^4)), the labelled edge count, and labelled
         5-10-vertex cycles (with and without a chord) against a plain depth-first walk.
stream:  per-batch anchored counts telescope to brute-force differences, with and without
         deletions, and the sum does not depend on the batch size.
summary: closed form == literal replay of the reference's offline UpdateGlobalIndex.
A planted mistake (a dropped query edge) is caught by the closed forms.
"""
import os

import numpy as np
import pytest

import util


def write_graph(path, n, edges, labels=None, elabels=None):
    with open(path, "w") as f:
        for v in range(n):
            f.write(f"v {v} {labels[v] if labels else 0}\n")
        for k, (a, b) in enumerate(edges):
            f.write(f"e {a} {b} {elabels[k] if elabels else 0}\n")


def write_query(path, n, edges, labels=None):
    write_graph(path, n, edges, labels)


def random_graph(rng, n, p):
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
    A = np.zeros((n, n), dtype=np.int64)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    return edges, A


@pytest.mark.parametrize("seed", range(4))
def test_closed_forms(tmp_path, seed):
    util.ensure_built()
    rng = np.random.default_rng(seed)
    n = 14
    edges, A = random_graph(rng, n, 0.35)
    g = tmp_path / "g.graph"
    write_graph(g, n, edges)
    d = A.sum(1)
    cases = {
        "edge": (2, [(0, 1)], 2 * len(edges)),
        "wedge": (3, [(0, 1), (0, 2)], int((d * (d - 1)).sum())),
        "star3": (4, [(0, 1), (0, 2), (0, 3)], int((d * (d - 1) * (d - 2)).sum())),
        "triangle": (3, [(0, 1), (1, 2), (0, 2)], int(np.trace(A @ A @ A))),
        "4cycle": (4, [(0, 1), (1, 2), (2, 3), (0, 3)],
                   int(np.trace(A @ A @ A @ A) - 2 * (d * d).sum() + 2 * len(edges))),
    }
    for name, (qn, qe, want) in cases.items():
        q = tmp_path / f"{name}.graph"
        write_query(q, qn, qe)
        got, _ = util.oracle_count(g, q)
        assert got == want, name


def test_labelled_edge_count(tmp_path):
    util.ensure_built()
    rng = np.random.default_rng(5)
    n = 30
    edges, _ = random_graph(rng, n, 0.2)
    labels = [int(x) for x in rng.integers(0, 3, n)]
    g = tmp_path / "g.graph"
    write_graph(g, n, edges, labels)
    q = tmp_path / "q.graph"
    write_query(q, 2, [(0, 1)], [0, 1])
    want = sum(1 for a, b in edges if {labels[a], labels[b]} == {0, 1})
    assert util.oracle_count(g, q)[0] == want


def _brute_cycle_embeddings(n, edges, labels, qlabels, chord=None):
    """Injective maps of the k-cycle q0-q1-...-q(k-1)-q0 (plus an optional chord between query
    vertices) into the graph that keep vertex labels: every sequence of distinct vertices, walked
    by plain depth-first search. Automorphic images count separately."""
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    k, total = len(qlabels), 0

    def dfs(path, used):
        nonlocal total
        v = path[-1]
        if len(path) == k:
            if path[0] in adj[v] and (chord is None or path[chord[1]] in adj[path[chord[0]]]):
                total += 1
            return
        for w in adj[v]:
            if not used >> w & 1 and labels[w] == qlabels[len(path)]:
                path.append(w)
                dfs(path, used | 1 << w)
                path.pop()

    for s in range(n):
        if labels[s] == qlabels[0]:
            dfs([s], 1 << s)
    return total


@pytest.mark.parametrize("k,seed", [(5, 0), (8, 1), (8, 2), (10, 3)])
def test_long_cycles_against_brute_force(tmp_path, k, seed):
    """8-10-vertex cycle queries with labels (the cycle workload's query shape): the oracle's count
    equals a brute-force walk over vertex sequences, with and without a chord."""
    util.ensure_built()
    rng = np.random.default_rng(seed)
    n = 18
    edges, _ = random_graph(rng, n, 0.4)
    labels = [int(x) for x in rng.integers(0, 2, n)]
    g = tmp_path / "g.graph"
    write_graph(g, n, edges, labels)
    qlabels = [int(x) for x in rng.integers(0, 2, k)]
    ring = [(i, (i + 1) % k) for i in range(k)]
    for chord in (None, (0, k // 2)):
        q = tmp_path / f"c{k}_{chord is not None}.graph"
        write_query(q, k, ring + ([chord] if chord else []), qlabels)
        want = _brute_cycle_embeddings(n, edges, labels, qlabels, chord)
        got, _ = util.oracle_count(g, q)
        assert got == want, f"k={k} chord={chord}"
        if chord is None:
            assert want > 0  # the case is not vacuous
    # planted mistake: the cycle with its closing edge dropped (an open path) must not match
    q = tmp_path / f"p{k}.graph"
    write_query(q, k, ring[:-1], qlabels)
    assert util.oracle_count(g, q)[0] != _brute_cycle_embeddings(n, edges, labels, qlabels)


def test_planted_mistake_is_caught(tmp_path):
    """Counting a 4-cycle with one query edge dropped (a 4-path) must disagree with the closed form."""
    util.ensure_built()
    rng = np.random.default_rng(1)
    edges, A = random_graph(rng, 12, 0.4)
    g = tmp_path / "g.graph"
    write_graph(g, 12, edges)
    d = A.sum(1)
    c4 = int(np.trace(A @ A @ A @ A) - 2 * (d * d).sum() + 2 * len(edges))
    q = tmp_path / "q.graph"
    write_query(q, 4, [(0, 1), (1, 2), (2, 3)])
    assert util.oracle_count(g, q)[0] != c4


@pytest.mark.parametrize("seed", [1, 2])
def test_stream_telescopes(tmp_path, seed):
    util.ensure_built()
    meta, queries = util.gen("tiny", tmp_path, seed=seed)
    data, upd = tmp_path / "data.graph", tmp_path / "update.stream"
    final = tmp_path / "final.graph"
    util.write_final_graph(data, upd, final)
    sums = {}
    for batch in [100, 23, 10 ** 9]:
        res, _ = util.oracle_stream(data, upd, batch, queries)
        for k, q in enumerate(queries):
            pos = sum(v[0] - v[1] for (b, kk), v in res.items() if kk == k)
            hsh = sum(v[2] - v[3] for (b, kk), v in res.items() if kk == k) % (1 << 64)
            sums.setdefault(k, set()).add((pos, hsh))
    for k, q in enumerate(queries):
        c0, h0 = util.oracle_count(data, q)
        c1, h1 = util.oracle_count(final, q)
        assert sums[k] == {(c1 - c0, (h1 - h0) % (1 << 64))}


def test_summary_closed_form_equals_literal_replay(tmp_path):
    util.ensure_built()
    meta, queries = util.gen("tiny", tmp_path, seed=3)
    for q in queries:
        a = util.run([util.ORACLE, "summary", str(tmp_path / "data.graph"), q])
        b = util.run([util.ORACLE, "summary", str(tmp_path / "data.graph"), q, "--literal"])
        assert a == b


def test_directed_telescopes(tmp_path):
    """Directed extension: per-batch anchored counts telescope on a directed reading of the files."""
    util.ensure_built()
    meta, queries = util.gen("tiny", tmp_path, seed=6)
    data, upd = tmp_path / "data.graph", tmp_path / "update.stream"
    final = tmp_path / "final.graph"
    util.write_final_graph(data, upd, final)
    res, _ = util.oracle_stream(data, upd, 50, queries, directed=True)
    for k, q in enumerate(queries):
        pos = sum(v[0] - v[1] for (b, kk), v in res.items() if kk == k)
        c0, _ = util.oracle_count(data, q, directed=True)
        c1, _ = util.oracle_count(final, q, directed=True)
        assert pos == c1 - c0


def test_variable_batch_sizes(tmp_path):
    """--batch-sizes: equal sizes reproduce the fixed-size run line for line; uneven sizes still
    telescope to the brute-force difference between the first and the final graph."""
    util.ensure_built()
    meta, queries = util.gen("tiny", tmp_path, seed=4)
    data, upd = tmp_path / "data.graph", tmp_path / "update.stream"
    n = sum(1 for line in open(upd) if line.split() and line.split()[0] in ("e", "-e"))
    fixed, _ = util.oracle_stream(data, upd, 37, queries)
    (tmp_path / "equal.txt").write_text("37\n" * (n // 37 + 1))
    same, _ = util.oracle_stream(data, upd, 1, queries, extra=["--batch-sizes", str(tmp_path / "equal.txt")])
    assert same == fixed
    sizes = [1, 50, 3, 90, 7, 200, 11, 5000]
    (tmp_path / "uneven.txt").write_text("".join(f"{x}\n" for x in sizes))
    res, _ = util.oracle_stream(data, upd, 1, queries, extra=["--batch-sizes", str(tmp_path / "uneven.txt")])
    final = tmp_path / "final.graph"
    util.write_final_graph(data, upd, final)
    for k, q in enumerate(queries):
        pos = sum(v[0] - v[1] for (b, kk), v in res.items() if kk == k)
        hsh = sum(v[2] - v[3] for (b, kk), v in res.items() if kk == k) % (1 << 64)
        c0, h0 = util.oracle_count(data, q)
        c1, h1 = util.oracle_count(final, q)
        assert (pos, hsh) == (c1 - c0, (h1 - h0) % (1 << 64))


def _random_digraph(rng, n, p, nlabels):
    """Arcs a -> b (a != b) drawn independently for every ordered pair, each with a uniform edge
    label; both a -> b and b -> a may be present.  Returns the arcs and one adjacency matrix per
    edge label."""
    arcs = [(a, b, int(rng.integers(0, nlabels))) for a in range(n) for b in range(n) if a != b and rng.random() < p]
    A = np.zeros((nlabels, n, n), dtype=np.int64)
    for a, b, l in arcs:
        A[l, a, b] = 1
    return arcs, A


def _write_arcs(path, n, arcs, labels=None):
    with open(path, "w") as f:
        for v in range(n):
            f.write(f"v {v} {labels[v] if labels else 0}\n")
        for a, b, l in arcs:
            f.write(f"e {a} {b} {l}\n")


@pytest.mark.parametrize("seed", range(3))
def test_directed_and_edge_labelled_closed_forms(tmp_path, seed):
    """Directed and multi-edge-label counting against matrix closed forms (A_l = arcs of label l):
    directed 3-cycle = tr(A^3); directed 2-path with labels (l1, l2) = sum(A_l1 A_l2) -
    tr(A_l1 A_l2) (the closed walks a -> b -> a are not injective); undirected triangle with three
    edge labels = tr(S_0 S_1 S_2) over the symmetric per-label adjacencies.  Planted mistakes (a
    reversed arc, a dropped edge label) change the count."""
    util.ensure_built()
    rng = np.random.default_rng(100 + seed)
    n = 16
    # directed, one edge label
    arcs, A = _random_digraph(rng, n, 0.3, 1)
    g = tmp_path / "d1.graph"
    _write_arcs(g, n, arcs)
    q = tmp_path / "dcycle3.graph"
    _write_arcs(q, 3, [(0, 1, 0), (1, 2, 0), (2, 0, 0)])
    A0 = A[0]
    assert util.oracle_count(g, q, directed=True)[0] == int(np.trace(A0 @ A0 @ A0))
    # planted mistake: one arc of the cycle reversed gives a different count
    q = tmp_path / "dcycle3_rev.graph"
    _write_arcs(q, 3, [(0, 1, 0), (2, 1, 0), (2, 0, 0)])
    assert util.oracle_count(g, q, directed=True)[0] != int(np.trace(A0 @ A0 @ A0))
    # directed, three edge labels: every labelled 2-path
    arcs, A = _random_digraph(rng, n, 0.35, 3)
    g = tmp_path / "d3.graph"
    _write_arcs(g, n, arcs)
    for l1 in range(3):
        for l2 in range(3):
            P = A[l1] @ A[l2]
            want = int(P.sum() - np.trace(P))
            q = tmp_path / f"dpath_{l1}{l2}.graph"
            _write_arcs(q, 3, [(0, 1, l1), (1, 2, l2)])
            assert util.oracle_count(g, q, directed=True)[0] == want, (l1, l2)
    # planted mistake: the same path read with its edge labels ignored (all arcs of any label)
    Aall = A.sum(0)
    Pall = Aall @ Aall
    q = tmp_path / "dpath_01.graph"
    assert util.oracle_count(g, q, directed=True)[0] != int(Pall.sum() - np.trace(Pall))
    # undirected, three edge labels: one label per unordered pair
    pairs = [(a, b, int(rng.integers(0, 3))) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.45]
    S = np.zeros((3, n, n), dtype=np.int64)
    for a, b, l in pairs:
        S[l, a, b] = S[l, b, a] = 1
    g = tmp_path / "u3.graph"
    _write_arcs(g, n, pairs)
    q = tmp_path / "tri012.graph"
    _write_arcs(q, 3, [(0, 1, 0), (1, 2, 1), (0, 2, 2)])
    want = int(np.trace(S[0] @ S[1] @ S[2]))
    assert want > 0
    assert util.oracle_count(g, q)[0] == want


def _brute_directed_cycles(n, arcs, labels, qlabels, qelabels):
    """Injective maps of the directed cycle q0 -> q1 -> ... -> q(k-1) -> q0 whose arc i carries edge
    label qelabels[i], keeping vertex labels: a plain depth-first walk over vertex sequences."""
    out = {}
    for a, b, l in arcs:
        out.setdefault(a, set()).add((b, l))
    k, total = len(qlabels), 0

    def dfs(path, used):
        nonlocal total
        v = path[-1]
        if len(path) == k:
            if (path[0], qelabels[k - 1]) in out.get(v, ()):
                total += 1
            return
        for w, l in out.get(v, ()):
            if l == qelabels[len(path) - 1] and not used >> w & 1 and labels[w] == qlabels[len(path)]:
                path.append(w)
                dfs(path, used | 1 << w)
                path.pop()

    for s in range(n):
        if labels[s] == qlabels[0]:
            dfs([s], 1 << s)
    return total


@pytest.mark.parametrize("k,seed", [(4, 0), (5, 1), (6, 2)])
def test_directed_labelled_cycles_against_brute_force(tmp_path, k, seed):
    """Directed cycles with vertex and edge labels (the serving workload's query kind): the oracle's
    directed count equals a depth-first walk over arc sequences; a query with one edge label
    changed is a planted mistake the walk tells apart."""
    util.ensure_built()
    rng = np.random.default_rng(200 + seed)
    n = 14
    arcs, _ = _random_digraph(rng, n, 0.45, 2)
    labels = [int(x) for x in rng.integers(0, 2, n)]
    g = tmp_path / "g.graph"
    _write_arcs(g, n, arcs, labels)
    qlabels = [int(x) for x in rng.integers(0, 2, k)]
    qel = [int(x) for x in rng.integers(0, 2, k)]
    q = tmp_path / "q.graph"
    _write_arcs(q, k, [(i, (i + 1) % k, qel[i]) for i in range(k)], qlabels)
    want = _brute_directed_cycles(n, arcs, labels, qlabels, qel)
    assert want > 0
    assert util.oracle_count(g, q, directed=True)[0] == want
    wrong = list(qel)
    wrong[0] ^= 1
    assert _brute_directed_cycles(n, arcs, labels, qlabels, wrong) != want
    q2 = tmp_path / "q2.graph"
    _write_arcs(q2, k, [(i, (i + 1) % k, wrong[i]) for i in range(k)], qlabels)
    assert util.oracle_count(g, q2, directed=True)[0] != want


def test_directed_edge_labelled_stream_telescopes(tmp_path):
    """The serving workload family (directed, 4 vertex x 8 edge labels, 85/15 insert/delete,
    variable batches) at a small scale: per-batch anchored counts and hashes telescope to the
    brute-force difference of the first and the final graph, for every query."""
    util.ensure_built()
    meta, queries = util.gen("c5", tmp_path, seed=2, scale=0.0005)
    data, upd = tmp_path / "data.graph", tmp_path / "update.stream"
    final = tmp_path / "final.graph"
    util.write_final_graph(data, upd, final, directed=True)
    res, _ = util.oracle_stream(data, upd, 1, queries, threads=8, directed=True,
                                extra=["--batch-sizes", str(tmp_path / "batches.txt")])
    assert any(v[1] for v in res.values())
    nonzero = 0
    for k, q in enumerate(queries[:12]):
        pos = sum(v[0] - v[1] for (b, kk), v in res.items() if kk == k)
        hsh = sum(v[2] - v[3] for (b, kk), v in res.items() if kk == k) % (1 << 64)
        c0, h0 = util.oracle_count(data, q, directed=True)
        c1, h1 = util.oracle_count(final, q, directed=True)
        assert (pos, hsh) == (c1 - c0, (h1 - h0) % (1 << 64)), f"query {k}"
        nonzero += pos != 0
    assert nonzero > 0
