"""The serving loop's scheduling logic on the CPU, with a stand-in device
(the planner is the real host planner; counting is replaced by a
deterministic function of the schedule), and the synthetic load generators."""
import threading
import time
import zlib

import numpy as np

import graphpi_amd as gp
from graphpi_amd.serve import CountServer, Request, percentile, poisson_arrivals, random_pattern, run_load


class FakeDevice:
    """A count over TASKS tasks whose raw sum is unit(s) x redundancy per
    task (so the count alone is value(s) = TASKS x unit(s)).  count_batch_progress
    advances each schedule by step(s) tasks per call, so some requests take
    several time slices; it takes time and records how much footprint was in
    flight at once."""
    TASKS = 100
    lock = threading.Lock()

    def __init__(self, host, log):
        self.host = host
        self.log = log

    def triangle_counting(self):
        return 1000

    def footprint(self, scheds, tuning=None):
        return 100 * len(scheds)

    def count_tasks(self, s):
        return FakeDevice.TASKS

    def count_batch_progress(self, scheds, begin, tuning=None, abort_s=None):
        with FakeDevice.lock:
            self.log["in_flight"] += self.footprint(scheds)
            self.log["peak"] = max(self.log["peak"], self.log["in_flight"])
            self.log["sizes"].append(len(scheds))
        time.sleep(0.002 * len(scheds))
        # a "heavy" schedule's task outlasts the per-task limit a light batch
        # runs with (the server's probe): abandoned without progress there
        probe = tuning is not None and tuning.task_abort_s > 0
        aborted = [probe and self.heavy(s) for s in scheds]
        done = [b if ab else min(FakeDevice.TASKS, b + self.step(s)) for s, b, ab in zip(scheds, begin, aborted)]
        raws = [(d - b) * self.unit(s) * (s.redundancy or 1) for s, b, d in zip(scheds, begin, done)]
        with FakeDevice.lock:
            self.log["in_flight"] -= self.footprint(scheds)
            if len(scheds) == 1 and self.heavy(scheds[0]):
                self.log.setdefault("heavy_alone", 0)
                self.log["heavy_alone"] += 1
        return done, raws, gp.Stats(), all(d == FakeDevice.TASKS for d in done), aborted

    @staticmethod
    def unit(s):
        d = s.to_dict()
        return zlib.crc32(repr((d["adj"], d["restrict_pair"], d["size"])).encode()) & 0xFFFF

    @staticmethod
    def heavy(s):
        return FakeDevice.unit(s) % 4 == 0

    @staticmethod
    def step(s):
        return 40 + 30 * (FakeDevice.unit(s) % 3)  # 1, 2 or 3 slices

    @staticmethod
    def value(s):
        return FakeDevice.TASKS * FakeDevice.unit(s)


def _server(log, **kw):
    graphs = {"a": gp.HostGraph.complete(8), "b": gp.HostGraph.complete(9)}
    return CountServer(graphs, device_factory=lambda h: FakeDevice(h, log), **kw)


def test_random_patterns_connected_and_deterministic():
    for k in range(3, 8):
        a1 = random_pattern(np.random.default_rng(k), k)
        a2 = random_pattern(np.random.default_rng(k), k)
        assert (a1 == a2).all() and (a1 == a1.T).all() and not a1.diagonal().any()
        seen, stack = {0}, [0]
        while stack:
            v = stack.pop()
            for u in np.nonzero(a1[v])[0]:
                if u not in seen:
                    seen.add(int(u))
                    stack.append(int(u))
        assert len(seen) == k


def test_poisson_arrivals_rate():
    off = poisson_arrivals(np.random.default_rng(1), 200.0, 4000)
    assert abs(len(off) / off[-1] - 200.0) < 20.0
    assert (np.diff(off) > 0).all()


def test_results_same_alone_and_in_the_mix():
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    rng = np.random.default_rng(3)
    reqs = [Request(i, "ab"[i % 2], random_pattern(rng, int(rng.integers(3, 7)))) for i in range(120)]
    with _server(log, slots=3, max_batch=8, mem_budget=1500) as srv:
        offs = poisson_arrivals(rng, 2000.0, len(reqs))
        run_load(srv, reqs, offs)
        for r in reqs:
            alone = FakeDevice.value(srv.schedule_for(r.graph, r.adj, r.settings))
            assert r.wait() == alone
            assert r.latency >= 0
    assert max(log["sizes"]) <= 8
    assert log["peak"] <= 1500
    assert max(log["sizes"]) > 1  # requests were batched
    lat = [r.latency for r in reqs]
    assert percentile(lat, 50) <= percentile(lat, 99)


def test_bad_pattern_fails_alone():
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    disconnected = np.zeros((4, 4), dtype=np.int32)
    disconnected[0, 1] = disconnected[1, 0] = 1
    disconnected[2, 3] = disconnected[3, 2] = 1
    with _server(log, slots=1) as srv:
        bad = srv.submit(Request(0, "a", disconnected))
        good = srv.submit(Request(1, "a", gp.PATTERNS["triangle"]))
        assert good.wait(10) is not None
        try:
            bad.wait(10)
            assert False, "a disconnected pattern must fail"
        except ValueError:
            pass


def test_footprint_error_fails_requests_and_keeps_serving():
    # a footprint query that raises fails the requests it covered; the slot
    # survives and serves the requests that come after
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    calls = {"n": 0}

    class Flaky(FakeDevice):
        def footprint(self, scheds, tuning=None):
            calls["n"] += 1
            if calls["n"] == 1:
                raise gp.GPError(-3, "schedule count out of range")
            return super().footprint(scheds, tuning)

    graphs = {"a": gp.HostGraph.complete(8)}
    with CountServer(graphs, slots=1, device_factory=lambda h: Flaky(h, log)) as srv:
        first = srv.submit(Request(0, "a", gp.PATTERNS["triangle"]))
        try:
            first.wait(10)
            assert False, "the first request must fail"
        except gp.GPError:
            pass
        second = srv.submit(Request(1, "a", gp.PATTERNS["rectangle"]))
        assert second.wait(10) == FakeDevice.value(srv.schedule_for("a", second.adj, second.settings))


def test_max_batch_limit():
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    for bad in (0, 257):
        try:
            _server(log, max_batch=bad)
            assert False, bad
        except ValueError:
            pass


def test_long_counts_resume_and_deadlines_fail_alone():
    """Counts that need several time slices resume where they stopped and
    return the same value as alone; a request whose deadline has passed when
    its slice ends fails with RequestTimeout and the others are unaffected."""
    from graphpi_amd.serve import RequestTimeout
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    rng = np.random.default_rng(5)
    reqs = [Request(i, "ab"[i % 2], random_pattern(rng, int(rng.integers(3, 7)))) for i in range(40)]
    with _server(log, slots=2, max_batch=4, mem_budget=10000) as srv:
        run_load(srv, reqs, poisson_arrivals(rng, 1000.0, len(reqs)))
        for r in reqs:
            s = srv.schedule_for(r.graph, r.adj, r.settings)
            assert r.wait() == FakeDevice.value(s) and r.done == r.total == FakeDevice.TASKS
            assert r.heavy or r.slices == -(-FakeDevice.TASKS // FakeDevice.step(s))
            assert r.heavy == (FakeDevice.heavy(s) and r.slices > -(-FakeDevice.TASKS // FakeDevice.step(s)))
        assert max(r.slices for r in reqs) > 1
        # requests whose tasks outlast the probe were run alone and still exact
        assert any(r.heavy for r in reqs) and log.get("heavy_alone", 0) > 0
        # a deadline already passed when the request is taken: it fails
        # without using the device, and the others are unaffected
        name = next(n for n in ("house", "pentagon", "4clique", "rectangle", "5clique", "hourglass")
                    if FakeDevice.step(srv.schedule_for("a", np.asarray(gp.PATTERNS[n], dtype=np.int32),
                                                        (1, 1, 1))) < FakeDevice.TASKS)
        late = Request(100, "a", gp.PATTERNS[name], deadline=0.0)
        srv.submit(late)
        ok = srv.submit(Request(101, "a", gp.PATTERNS[name]))
        try:
            late.wait(10)
            assert False, "the deadline must fail the request"
        except RequestTimeout as e:
            assert late.timed_out and e.done == 0 and late.slices == 0
        assert ok.wait(10) == FakeDevice.value(ok.sched)
        # a deadline that passes during a count that needs more than one
        # slice fails it after a slice, with its progress so far
        slow = Request(102, "a", gp.PATTERNS[name], deadline=0.001)
        srv.submit(slow)
        try:
            slow.wait(10)
            assert False, "the deadline must fail the request"
        except RequestTimeout as e:
            assert slow.timed_out and e.done < e.total


def test_new_requests_order_deadlines_smallest_first():
    """The new-request queue: requests already past their deadline fail
    without a slice; those with a deadline are taken smallest pattern first;
    those without one follow in arrival order."""
    from graphpi_amd.serve import RequestTimeout
    log = {"in_flight": 0, "peak": 0, "sizes": []}
    graphs = {"a": gp.HostGraph.complete(8)}
    srv = CountServer(graphs, slots=0, device_factory=lambda h: FakeDevice(h, log))  # no slots: the test drives
    t = time.perf_counter()
    big = Request(0, "a", gp.PATTERNS["house"], t_arrival=t - 3.0, deadline=100.0)
    small = Request(1, "a", gp.PATTERNS["triangle"], t_arrival=t - 2.0, deadline=100.0)
    free_old = Request(2, "a", gp.PATTERNS["rectangle"], t_arrival=t - 5.0)
    free_new = Request(3, "a", gp.PATTERNS["4clique"], t_arrival=t - 1.0)
    late = Request(4, "a", gp.PATTERNS["triangle"], t_arrival=t - 2.0, deadline=1.0)
    for r in (big, small, free_old, free_new, late):
        srv.submit(r)
    order = []
    with srv._cv:
        while srv._pending:
            batch = srv._take_new(time.perf_counter())
            order.append([r.rid for r in batch])
    assert late.timed_out and late.slices == 0 and late.done == 0
    try:
        late.wait(0)
        assert False
    except RequestTimeout:
        pass
    flat = [rid for b in order for rid in b]
    assert sorted(flat) == [0, 1, 2, 3]
    assert flat.index(1) < flat.index(0) < flat.index(2) < flat.index(3)
    srv.close()
