"""CPU: the sketch server's scheduling (admission order, HBM budget, lanes)
driven by fake lanes whose work completes after a set number of polls."""
from __future__ import annotations

import math

import numpy as np
import pytest

from shgemm_amd.serve import SketchRequest, SketchServer, percentile, poisson_offsets, run_arrivals, sketch_shapes


class FakeLanes:
    """Work of request r finishes after `polls(r)` completion checks."""

    def __init__(self, polls):
        self.polls = polls
        self.launched = []

    def launch(self, req, slot):
        self.launched.append((req.rid, slot))
        return {"left": self.polls(req)}

    def launch_batch(self, reqs, slot):
        self.launched.append((tuple(r.rid for r in reqs), slot))
        return {"left": max(self.polls(r) for r in reqs)}

    @staticmethod
    def finished(token):
        token["left"] -= 1
        return token["left"] < 0

    def close(self):
        pass


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4
        return self.t


def _req(i, m=100, k=100, l=64):
    return SketchRequest(rid=i, m=m, k=k, l=l, seed=i, a=0, lda=m)


def test_fifo_budget_and_lanes():
    reqs = [_req(i, m=100 * (1 + i % 5), k=200, l=64) for i in range(40)]
    budget = 3 * max(r.footprint for r in reqs)
    lanes = FakeLanes(lambda r: 1 + r.rid % 7)
    srv = SketchServer(hbm_budget=budget, slots=4, lanes=lanes, clock=FakeClock())
    for r in reqs:
        srv.submit(r)
    seen_in_use = []
    done = []
    while not srv.idle():
        done += srv.poll()
        srv.admit()
        seen_in_use.append(srv.in_use)
        assert len(srv.inflight) <= 4
    assert sorted(r.rid for r in done) == list(range(40))
    assert srv.admit_order == list(range(40))            # first come, first served
    assert max(seen_in_use) <= budget and srv.peak_in_use <= budget
    assert srv.in_use == 0 and len(srv.free) == 4
    # no lane held two requests at once
    busy = {}
    for r in sorted(done, key=lambda r: r.admitted):
        assert busy.get(r.slot, -1.0) <= r.admitted
        busy[r.slot] = r.done
    assert all(r.done >= r.admitted >= r.arrival for r in done)


def test_budget_blocks_head_of_line():
    big = _req(0, m=1000, k=1000, l=64)
    small = _req(1, m=10, k=10, l=64)
    srv = SketchServer(hbm_budget=big.footprint + 10, slots=4, lanes=FakeLanes(lambda r: 3), clock=FakeClock())
    srv.submit(big)
    srv.submit(_req(2, m=1000, k=1000, l=64))  # does not fit beside the first
    srv.submit(small)
    srv.admit()
    assert [rid for rid in srv.admit_order] == [0]  # the second waits, and the small one behind it too (FIFO)
    srv.drain()
    assert srv.admit_order == [0, 2, 1]


def test_oversized_request_rejected():
    srv = SketchServer(hbm_budget=1000, slots=2, lanes=FakeLanes(lambda r: 0), clock=FakeClock())
    with pytest.raises(ValueError):
        srv.submit(_req(0, m=100, k=100, l=64))


def test_run_arrivals_stats():
    rng = np.random.default_rng(0)
    reqs = [_req(i) for i in range(200)]
    offs = poisson_offsets(len(reqs), 1000.0, rng)
    srv = SketchServer(hbm_budget=1e12, slots=3, lanes=FakeLanes(lambda r: 2), clock=FakeClock())
    st = run_arrivals(srv, reqs, offs)
    assert st["requests"] == 200
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["peak_inflight"] <= 3


def test_percentile_nearest_rank():
    xs = list(range(1, 101))
    assert percentile(xs, 50) == 50 and percentile(xs, 99) == 99 and percentile(xs, 100) == 100
    assert percentile([5.0], 99) == 5.0
    assert math.isnan(percentile([], 50))


def test_poisson_and_shapes():
    rng = np.random.default_rng(1)
    offs = poisson_offsets(20000, 500.0, rng)
    gaps = np.diff(offs)
    assert abs(gaps.mean() - 1 / 500) < 0.05 / 500 and offs[0] == 0.0
    sh = sketch_shapes(5000, rng)
    ms, ks, ls = zip(*sh)
    assert 1024 <= min(ms) and max(ms) <= 32768 and 1024 <= min(ks) and max(ks) <= 16384
    assert set(ls) <= {64, 128, 256, 512, 1024}
    # log-uniform: the median is near the geometric mean of the range
    assert abs(math.log(np.median(ms)) - 0.5 * (math.log(1024) + math.log(32768))) < 0.15


def test_batched_admission_fifo_budget_and_completion():
    """max_batch > 1: the requests waiting at a step go to one free lane as
    one launch (up to max_batch, within the budget, in arrival order); a
    batch's requests complete together and free their budget then."""
    reqs = [_req(i, m=100 * (1 + i % 3), k=200, l=64) for i in range(30)]
    budget = 5 * max(r.footprint for r in reqs)
    lanes = FakeLanes(lambda r: 2 + r.rid % 4)
    srv = SketchServer(hbm_budget=budget, slots=2, lanes=lanes, clock=FakeClock(), max_batch=4)
    for r in reqs:
        srv.submit(r)
    done = []
    while not srv.idle():
        done += srv.poll()
        srv.admit()
        assert srv.in_use <= budget
        assert len(srv.inflight) <= 2
        assert all(len(b) <= 4 for b, _ in srv.inflight.values())
    assert sorted(r.rid for r in done) == list(range(30))
    order = [rid for batch, _ in lanes.launched for rid in batch]
    assert order == list(range(30))                         # first come, first served
    assert max(len(b) for b, _ in lanes.launched) == 4      # a backlog is batched
    # every request of a launch completes at the same time
    for batch, _ in lanes.launched:
        t = {next(r for r in done if r.rid == rid).done for rid in batch}
        assert len(t) == 1
    st_sizes = srv.batch_sizes
    assert sum(st_sizes) == 30 and len(st_sizes) == len(lanes.launched)


def test_batching_off_launches_one_request_per_lane():
    lanes = FakeLanes(lambda r: 1)
    srv = SketchServer(hbm_budget=1e12, slots=3, lanes=lanes, clock=FakeClock())
    for i in range(7):
        srv.submit(_req(i))
    srv.drain()
    assert all(isinstance(rid, int) for rid, _ in lanes.launched)
