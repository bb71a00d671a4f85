"""The CPU baseline (oracle/bin/baseline: backtracking over the planner's matching orders, timed by
bench.py) must count what the oracle counts: per batch, its matches equal the oracle's positive
plus negative counts summed over the queries, with deletions in the stream and any plan."""
import os

import pytest

import util

BASELINE = os.path.join(util.ROOT, "oracle", "bin", "baseline")


def orders_of(query, order):
    out = util.run([util.TOOL, "plan", str(query), "-", str(order)])
    return out[out.index("PrintOrders:"):]


@pytest.mark.parametrize("config,scale,seed,batch,order", [
    ("tiny", None, 1, 100, 2), ("tiny", None, 2, 37, 1), ("tinyins", None, 3, 100, 0),
    ("c2", 0.01, 2, 655, 2), ("c3", 0.01, 1, 10485, 2), ("c3", 0.01, 1, 10485, 3),
])
def test_baseline_counts_equal_oracle(tmp_path, config, scale, seed, batch, order):
    util.ensure_built()
    meta, queries = util.gen(config, tmp_path, seed=seed, scale=scale)
    data, upd = tmp_path / "data.graph", tmp_path / "update.stream"
    args = []
    for q in queries:
        with open(q + ".orders", "w") as f:
            f.write(orders_of(q, order))
        args += [q, q + ".orders"]
    want, _ = util.oracle_stream(data, upd, batch, queries, threads=8)
    out = util.run([BASELINE, str(data), str(upd), str(batch)] + args + ["--threads", "8"])
    got = {}
    for line in out.splitlines():
        t = line.split()  # batch <b> items <done> of <total> matches <count> ms <wall>
        assert t[0] == "batch" and t[3] == t[5]
        got[int(t[1])] = int(t[7])
    nb = 1 + max(b for b, _ in want)
    assert got == {b: sum(v[0] + v[1] for (bb, _), v in want.items() if bb == b) for b in range(nb)}
    assert any(got.values())
