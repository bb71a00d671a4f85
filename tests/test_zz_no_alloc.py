"""csm_apply_range_static (partial-result rings in the library's static device ring, no allocation)
against csm_apply_range: two sessions over the same graph, queries and stream take the same
ranges, the first through the existing entry point, the second through the static one; every
per-query count and hash must be identical, on two calls in a row (the second smaller), then on a
normal call after them.
"""
import sys

import pytest

import util

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def open_session(ffi, d, queries, batch, **opts):
    s = ffi.Session(device=0, **opts)
    s.load_graph(d / "data.graph")
    for q in queries:
        s.add_query(q)
    s.open_updates(d / "update.stream", batch)
    s.build()
    for q in s.queries:
        s.plan(q)
    return s


@pytest.mark.parametrize("lanes", [0, 16])
def test_static_ring_matches_pool_ring(tmp_path, lanes):
    util.ensure_built()
    ffi = util.ffi()
    meta, queries = util.gen("c2", tmp_path, seed=4, scale=0.02)
    batch = int(meta["batch_size"])
    stream = int(meta["stream"])
    n1 = min(stream // 2, 3 * batch)
    n2 = n1 // 3
    assert 0 < n2 < n1 and n1 + 2 * n2 <= stream
    say(f"c2 x0.02 seed 4: lanes {lanes}, ranges [0,{n1}) [{n1},{n1 + n2}) [{n1 + n2},{n1 + 2 * n2})")
    a = open_session(ffi, tmp_path, queries, batch, lanes=lanes, max_batch_edges=n1)
    b = open_session(ffi, tmp_path, queries, batch, lanes=lanes, max_batch_edges=n1)
    try:
        want1, _ = a.apply_range(0, n1)
        got1, info1 = b.apply_range_static(0, n1)
        assert got1 == want1
        assert info1.edges == n1 and info1.deletes > 0
        assert any(r[0] for r in got1) and any(r[1] for r in got1)
        want2, _ = a.apply_range(n1, n2)
        got2, _ = b.apply_range_static(n1, n2)
        assert got2 == want2
        want3, _ = a.apply_range(n1 + n2, n2)
        got3, _ = b.apply_range(n1 + n2, n2)  # the session's own rings are back in place
        assert got3 == want3
        with pytest.raises(ffi.CsmError) as e:
            b.apply_range_static(stream, 1)
        assert e.value.status == 1  # ARGUMENT: range beyond the stream, nothing changed
    finally:
        a.close()
        b.close()
