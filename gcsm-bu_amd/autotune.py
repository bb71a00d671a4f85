"""Choice of a session's lane count for one workload, by measurement.

The reference fixes its tunables at compile time. The semi-BFS levels stop expanding at
MIN_NUM_RESULTS_TO_GPU = 65,536 partial results (config.h:47, used at match_gpu.cu:1961). One CUDA
stream runs the steps of a batch one after another. Here both are session options
(`csm_options.semi_bfs_threshold`, `csm_options.lanes`). The lane count sizes the session's scratch,
so `tune` opens one fresh session per candidate and applies the same prefix of the stream: warm-up
batches untimed, then timed batches. It keeps the candidate with the least time per timed batch.
The settings a session can change between batches (semi-BFS threshold, waves per anchor, DFS stack
depth, grids, lanes per team) are tuned inside one session by the library itself (`csm_tune`,
`Session.tune`).

The settings change scheduling only, never results. Every candidate's per-batch results (counts and
hashes of every query) must equal the first candidate's, and a difference raises `TuneMismatch`.
"""
import dataclasses
import os
import time

MAX_LANES = 64  # kMaxLanes of the library


@dataclasses.dataclass(frozen=True)
class Candidate:
    lanes: int               # concurrent step lanes (each with its own scratch), 1..64
    semi_bfs_threshold: int  # records at a Cartesian-product depth at which the levels stop expanding

    def options(self):
        return {"lanes": self.lanes, "semi_bfs_threshold": self.semi_bfs_threshold}


DEFAULT_CANDIDATES = tuple(Candidate(lanes, t) for lanes in (16, 24, 32) for t in (1 << 14, 1 << 16, 1 << 18, 1 << 20))


class TuneMismatch(RuntimeError):
    """Two candidates gave different results for the same batches (a library bug, never a tuning
    outcome)."""


def _check(candidates, warmup, batches):
    if not candidates:
        raise ValueError("no candidates")
    for c in candidates:
        if not 1 <= c.lanes <= MAX_LANES:
            raise ValueError(f"lanes {c.lanes} outside 1..{MAX_LANES}")
        if not 1 <= c.semi_bfs_threshold < 1 << 32:
            raise ValueError(f"semi_bfs_threshold {c.semi_bfs_threshold} outside 1..2^32-1")
    if warmup < 0 or batches < 1:
        raise ValueError("need warmup >= 0 and batches >= 1")
    if os.environ.get("CSM_LANES") and len({c.lanes for c in candidates}) > 1:
        # the library reads CSM_LANES at session creation and it overrides csm_options.lanes
        raise ValueError("CSM_LANES is set: it would override every candidate's lanes")


def tune(open_session, candidates=DEFAULT_CANDIDATES, warmup=1, batches=2, clock=time.perf_counter):
    """Time every candidate on batches [warmup, warmup + batches) of the stream.

    `open_session(options)` returns `(session, number of batches)` for a session created with
    `options` (keyword arguments of `csm_ffi.Session`), with the graph loaded, the queries planned
    and the stream open. `tune` closes it. Returns `(best candidate, table)`, where each row of
    `table` is `{"lanes", "semi_bfs_threshold", "ms_per_batch"}` in candidate order. Ties go to the
    earlier candidate.
    """
    candidates = tuple(candidates)
    _check(candidates, warmup, batches)
    table, first = [], None
    for c in candidates:
        s, nb = open_session(c.options())
        try:
            if nb < warmup + batches:
                raise ValueError(f"stream has {nb} batches, tuning needs {warmup + batches}")
            results = []
            for b in range(warmup):
                results.append(s.apply_batch(b)[0])
            s.synchronize()
            elapsed = 0.0
            for b in range(warmup, warmup + batches):
                t0 = clock()
                res, _ = s.apply_batch(b)
                s.synchronize()
                elapsed += clock() - t0
                results.append(res)
        finally:
            s.close()
        if first is None:
            first = results
        elif results != first:
            b = next(k for k, (x, y) in enumerate(zip(results, first)) if x != y)
            raise TuneMismatch(f"{c} differs from {candidates[0]} at batch {b}: {results[b]} != {first[b]}")
        table.append({**c.options(), "ms_per_batch": elapsed * 1e3 / batches})
    best = min(range(len(candidates)), key=lambda k: (table[k]["ms_per_batch"], k))
    return candidates[best], table
