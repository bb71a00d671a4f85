"""The autotuner's host logic, with stand-in sessions (no GPU): it keeps the fastest candidate,
times only the batches after the warm-up, closes every session, and raises when two candidates'
results differ."""
import pytest

import util

at = util.autotune()


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeSession:
    """apply_batch advances the clock by cost[batch] seconds and returns fixed results."""

    def __init__(self, clock, cost, results, log):
        self.clock, self.cost, self.results, self.log = clock, cost, results, log

    def apply_batch(self, b):
        self.clock.t += self.cost[b]
        self.log.append(("apply", b))
        return self.results[b], None

    def synchronize(self):
        pass

    def close(self):
        self.log.append(("close",))


def factory(clock, costs, results, nb=4, log=None):
    log = [] if log is None else log

    def open_session(opts):
        c = at.Candidate(**opts)
        r = results(c) if callable(results) else results
        return FakeSession(clock, costs[c], r, log), nb

    return open_session, log


def test_picks_fastest_and_times_only_timed_batches():
    a, b, c = at.Candidate(4, 1 << 14), at.Candidate(8, 1 << 16), at.Candidate(12, 1 << 18)
    # the warm-up batch 0 is slowest for b: it must not count
    costs = {a: [0.001, 0.010, 0.010, 9], b: [5.0, 0.004, 0.006, 9], c: [0.001, 0.007, 0.007, 9]}
    clock = FakeClock()
    res = [[(1, 0, 7, 0)], [(2, 1, 8, 9)], [(0, 0, 0, 0)], [(5, 5, 5, 5)]]
    open_session, log = factory(clock, costs, res)
    best, table = at.tune(open_session, [a, b, c], warmup=1, batches=2, clock=clock)
    assert best == b
    assert [row["lanes"] for row in table] == [4, 8, 12]
    assert table[0]["ms_per_batch"] == pytest.approx(10.0)
    assert table[1]["ms_per_batch"] == pytest.approx(5.0)
    assert table[2]["ms_per_batch"] == pytest.approx(7.0)
    assert log.count(("close",)) == 3
    assert ("apply", 3) not in log  # only batches 0..2


def test_tie_goes_to_earlier_candidate():
    a, b = at.Candidate(8, 1 << 16), at.Candidate(4, 1 << 16)
    clock = FakeClock()
    open_session, _ = factory(clock, {a: [0, 1.0], b: [0, 1.0]}, [[(1, 0, 0, 0)]] * 2, nb=2)
    best, _ = at.tune(open_session, [a, b], warmup=1, batches=1, clock=clock)
    assert best == a


def test_mismatch_raises_and_names_the_batch():
    a, b = at.Candidate(8, 1 << 16), at.Candidate(8, 1 << 14)
    clock = FakeClock()
    good = [[(1, 0, 3, 0)], [(2, 0, 4, 0)], [(3, 0, 5, 0)]]
    bad = [[(1, 0, 3, 0)], [(2, 0, 4, 0)], [(3, 0, 6, 0)]]  # hash of batch 2 differs
    open_session, log = factory(clock, {a: [0] * 3, b: [0] * 3}, lambda c: good if c == a else bad, nb=3)
    with pytest.raises(at.TuneMismatch, match="batch 2"):
        at.tune(open_session, [a, b], warmup=1, batches=2, clock=clock)
    assert log.count(("close",)) == 2


def test_short_stream_and_bad_candidates_rejected():
    a = at.Candidate(8, 1 << 16)
    clock = FakeClock()
    open_session, log = factory(clock, {a: [0] * 2}, [[(0, 0, 0, 0)]] * 2, nb=2)
    with pytest.raises(ValueError, match="2 batches"):
        at.tune(open_session, [a], warmup=1, batches=2, clock=clock)
    assert log == [("close",)]
    for bad in ([], [at.Candidate(0, 1)], [at.Candidate(65, 1)], [at.Candidate(8, 0)], [at.Candidate(8, 1 << 32)]):
        with pytest.raises(ValueError):
            at.tune(open_session, bad, clock=clock)
    with pytest.raises(ValueError):
        at.tune(open_session, [a], warmup=0, batches=0, clock=clock)


def test_default_candidates_cover_the_library_defaults():
    assert at.Candidate(32, 1 << 16) in at.DEFAULT_CANDIDATES  # the library's defaults (32 lanes, 65,536)
    assert all(1 <= c.lanes <= at.MAX_LANES for c in at.DEFAULT_CANDIDATES)


def test_lanes_environment_override_rejected(monkeypatch):
    a, b = at.Candidate(4, 1 << 16), at.Candidate(8, 1 << 16)
    clock = FakeClock()
    open_session, log = factory(clock, {a: [0] * 2, b: [0] * 2}, [[(0, 0, 0, 0)]] * 2, nb=2)
    monkeypatch.setenv("CSM_LANES", "4")
    with pytest.raises(ValueError, match="CSM_LANES"):
        at.tune(open_session, [a, b], warmup=1, batches=1, clock=clock)
    assert log == []  # no session opened
    best, _ = at.tune(open_session, [a], warmup=1, batches=1, clock=clock)  # one lane count: allowed
    assert best == a
