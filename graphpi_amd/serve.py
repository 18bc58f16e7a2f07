"""Serving loop: resident graphs, counting queries admitted as they arrive.

A CountServer keeps several graphs resident in HBM and answers counting
requests (a pattern, its planner settings and a graph name).  Worker slots
take requests continuously: a slot that frees up immediately gathers the
requests waiting for one graph (those with deadlines first, the smallest
patterns first among them, then the others oldest first), up to a batch size
and within the device-memory budget (each batch's footprint comes from
gp_count_footprint), and counts them together with
gp_count_batch_progress.  Slots run on their own threads and HIP streams, so
a batch that arrives while others are on the GPU starts as soon as a slot and
memory are free and its kernels overlap with theirs.

Counts run in time slices: a batch call stops after about quantum_s (its
launches stop fetching at a deadline, so each has counted a prefix of every
request's tasks), finished requests are answered, and the others continue in
later calls from the task where they stopped, shortest estimated remaining
time first (CountServer explains the two queues).  A long count therefore
never holds the device for long and short requests that arrive behind it
are not stuck.  A request that cannot finish before its deadline (seconds
after arrival) fails with RequestTimeout, as soon as that is known.  Every
answered result is the exact count: the raw sums of consecutive task ranges
add up, so slicing, batching and concurrency change when a request runs,
never what it returns.

Schedules are built on the host (the planner) and cached per (graph, pattern,
settings), as a server would cache them per canonical pattern.
"""
import collections
import threading
import time

import numpy as np

from . import DeviceGraph, build_schedule, device_memory, kernel_plan


class RequestTimeout(TimeoutError):
    """The request's deadline passed before its count was complete; .done of
    .total tasks had been counted."""

    def __init__(self, rid, done, total):
        super().__init__(f"request {rid}: deadline passed after {done} of {total} tasks")
        self.done = done
        self.total = total


class Request:
    __slots__ = ("rid", "graph", "adj", "settings", "deadline", "t_arrival", "t_start", "t_done", "count", "raw",
                 "error", "sched", "total", "done", "raw_acc", "slices", "heavy", "gpu_s", "cls", "_event")

    def __init__(self, rid, graph, adj, settings=(1, 1, 1), t_arrival=None, deadline=None):
        self.rid = rid
        self.graph = graph
        self.adj = np.asarray(adj, dtype=np.int32)
        self.settings = tuple(settings)
        self.deadline = deadline  # seconds after arrival; None: the server's default
        self.t_arrival = time.perf_counter() if t_arrival is None else t_arrival
        self.t_start = self.t_done = None
        self.count = self.raw = None
        self.error = None
        self.sched = None
        self.total = 0      # tasks of the count
        self.done = 0       # tasks counted so far (a prefix)
        self.raw_acc = 0    # u64 raw sum of those tasks
        self.slices = 0     # batch calls the request took part in
        self.heavy = False  # a task ran past the probe time in a new requests' batch
        self.gpu_s = 0.0    # device time spent on it (a call's time shared by its requests)
        self.cls = 0        # launch class of its kernel plan (launch_class)
        self._event = threading.Event()

    @property
    def timed_out(self):
        return isinstance(self.error, RequestTimeout)

    def wait(self, timeout=None):
        if not self._event.wait(timeout):
            raise TimeoutError(f"request {self.rid} not finished")
        if self.error is not None:
            raise self.error
        return self.count

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


def launch_class(sched):
    """The class of LDS size a batch call launches the schedule's plan in
    (pair level / batched level with at most 3 blocks / more / neither), as
    gp_count_batch and gp_count_batch_progress group plans."""
    p = kernel_plan(sched)
    return 0 if p["pair"] else (1 if p["blocks"] <= 3 else 2) if p["batched"] else 3


class CountServer:
    """graphs: {name: HostGraph}.  slots: concurrent batch calls on the
    device (slot 0 serves new requests first, the others continuing ones
    first).  max_batch: schedules per call.  mem_budget: device bytes the
    in-flight calls may use besides the graphs (a call larger than the budget
    still runs, alone).  quantum_s: a call's time slice.  deadline_s: default
    deadline of a request, in seconds after its arrival (None: none).
    probe_s: longest a single task may run in a new requests' batch (default
    2 x quantum).

    Scheduling.  New requests of one graph are counted together in a first
    slice with the per-task probe (requests with deadlines first, the
    smallest patterns first among them): most small counts finish there.  A request
    still unfinished (its slice ran out, or one of its tasks ran past the
    probe, which abandons it for that call without touching the others)
    joins the continuing queue with its progress.  That queue runs shortest
    estimated remaining time first (remaining tasks x the time per task seen
    so far), one request per call with its deadline as the hard stop, and a
    request whose estimate exceeds the time left before its deadline is
    failed at once instead of holding the device until it runs out."""

    MAX_BATCH = 256  # schedules gp_count_batch takes per call

    def __init__(self, graphs, slots=2, max_batch=32, mem_budget=None, tuning=None, device_factory=None,
                 quantum_s=0.05, deadline_s=None, probe_s=None, order="edf"):
        if not 1 <= max_batch <= self.MAX_BATCH:
            raise ValueError(f"max_batch must be 1..{self.MAX_BATCH}")
        if order not in ("edf", "srpt"):
            raise ValueError("order must be 'edf' or 'srpt'")
        self._order = order
        self._quantum = quantum_s
        self._deadline = deadline_s
        self._probe = probe_s if probe_s is not None else (2 * quantum_s if quantum_s else None)
        self._slots = slots
        make = device_factory or DeviceGraph
        self._dev = {}
        self._host = {}
        self._tri = {}
        for name, host in graphs.items():
            d = make(host)
            self._dev[name] = d
            self._host[name] = host
            self._tri[name] = d.triangle_counting()
        # stream launches by default: graph launches (tuning.use_graph = 1)
        # gave the same counts but higher median latency on these small
        # requests (measured: p50 at 200 q/s 8-10 ms against 2.3-2.6 ms)
        self._tuning = tuning
        # continuing requests: the time slice, the deadline as the hard stop
        self._slice_tuning = self._tuning_for_slice()
        # new requests: also the per-fetch probe
        self._new_tuning = self._tuning_for_slice()
        if self._probe:
            self._new_tuning.task_abort_s = self._probe
            # and no deferred heavy tasks: their parts run after the slice's
            # deadline, each under the probe, so a request of many heavy
            # tasks would hold the device until its abort time; in place, a
            # task that runs past the probe marks its request heavy at once
            self._new_tuning.split_min = -1
        self._max_batch = max_batch
        if mem_budget is None:
            # most of what is free once the graphs are resident (the rest is
            # left to the runtime and to other users of the device)
            free = device_memory()[0] if device_factory is None else 32 << 30
            mem_budget = int(free * 0.85)
        self._budget = mem_budget
        self._in_flight = 0
        self._pending = collections.deque()  # new requests, arrival order
        self._long = []                      # continuing requests
        self._cv = threading.Condition()
        self._closed = False
        self._sched_cache = {}
        self.batches = []  # (graph, size, footprint, seconds, kernel ms, launches) per batch call
        self._workers = [threading.Thread(target=self._run, args=(i,), name=f"count-slot-{i}", daemon=True)
                         for i in range(slots)]
        for w in self._workers:
            w.start()

    # -- client side --
    def submit(self, req):
        """Plan the request (on the caller's thread, cached) and queue it.  A
        pattern the planner rejects fails at once without reaching a slot."""
        if req.graph not in self._dev:
            raise KeyError(f"unknown graph {req.graph}")
        try:
            req.sched = self.schedule_for(req.graph, req.adj, req.settings)
            req.total = self._dev[req.graph].count_tasks(req.sched)
            req.cls = launch_class(req.sched)
            if req.deadline is None:
                req.deadline = self._deadline
        except Exception as e:
            req.error = e
            req.t_done = time.perf_counter()
            req._event.set()
            return req
        with self._cv:
            if self._closed:
                raise RuntimeError("server closed")
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def close(self):
        """Finish the queued requests, stop the slots and free the device
        graphs (and their call buffers) now, rather than whenever the
        garbage collector reaches them: a free synchronises the device and
        would stall whatever runs on it then."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        self._workers = []
        for d in self._dev.values():
            if hasattr(d, "free"):
                d.free()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scheduling --
    def schedule_for(self, graph, adj, settings):
        key = (graph, adj.tobytes(), adj.shape[0], settings)
        with self._cv:
            s = self._sched_cache.get(key)
        if s is None:
            h = self._host[graph]
            pm, rt, iep = settings
            s, valid = build_schedule(adj, pm, rt, iep, h.v_cnt, h.e_cnt, self._tri[graph])
            if not valid:
                raise ValueError("pattern has a disconnected prefix under these settings")
            with self._cv:
                self._sched_cache[key] = s
        return s

    def _tuning_for_slice(self):
        from . import Tuning
        t = Tuning()
        if self._tuning is not None:
            import ctypes
            ctypes.memmove(ctypes.byref(t), ctypes.byref(self._tuning), ctypes.sizeof(Tuning))
        if t.waves_per_cu <= 0:
            # the slots' launches run side by side: each gets its share of the
            # persistent waves (and of the per-wave scratch they need)
            t.waves_per_cu = max(4, 24 // max(1, self._slots))
        if self._quantum:
            t.time_limit_s = self._quantum
            t.slice_ms = max(1, int(self._quantum * 1000.0))
        return t

    def _order_key(self, r, now):
        """Continuing requests with a finite estimate (they have counted some
        tasks) come before those with none (a task of theirs ran past the
        probe); within each, earliest deadline first (order = "edf", the
        default) or shortest estimated remaining time first ("srpt")."""
        est = self._estimate(r, now)
        if self._order == "srpt":
            return (est == float("inf"), est)
        return (est == float("inf"), r.t_arrival + r.deadline if r.deadline is not None else float("inf"))

    @staticmethod
    def _estimate(r, now):
        """Seconds of device time the request still needs, from its own rate
        so far (infinite when it has counted nothing yet)."""
        if r.done <= 0:
            return float("inf")
        return (r.total - r.done) * r.gpu_s / r.done

    def _fail_timeout(self, r, now):
        r.error = RequestTimeout(r.rid, r.done, r.total)
        r.t_done = now
        r._event.set()

    def _take_new(self, now):
        """Called with the lock held: new requests of one graph and launch
        class, up to max_batch.  Requests with a deadline come first, the
        one with the least static cost (pattern vertices, then the graph's
        tasks) picking the group: shortest job first, since under load the
        small requests are the ones that can still meet their deadlines.
        Requests without one follow in arrival order (shortest first would
        only hold the large ones back).  Pending requests whose deadline has
        already passed fail here without a slice.  (A batch call runs one
        launch per class of LDS size in turn; with one class per call its
        whole time slice goes to the one launch, and no request waits behind
        another class's launch for a slice that never reaches it.)"""
        live = collections.deque()
        for r in self._pending:
            if r.deadline is not None and now - r.t_arrival >= r.deadline:
                self._fail_timeout(r, now)
            else:
                live.append(r)
        self._pending = live
        if not live:
            return []
        cost = lambda r: ((0, r.adj.shape[0], r.total, r.t_arrival) if r.deadline is not None
                          else (1, 0, 0, r.t_arrival))
        head = min(live, key=cost)
        graph, cls = head.graph, head.cls
        batch = sorted((r for r in live if r.graph == graph and r.cls == cls), key=cost)[:self._max_batch]
        taken = set(id(r) for r in batch)
        self._pending = collections.deque(r for r in live if id(r) not in taken)
        return batch

    def _take_long(self, now):
        """Called with the lock held: the continuing request with the least
        estimated remaining time; requests whose estimate exceeds the time
        left before their deadline fail here."""
        keep = []
        for r in self._long:
            left = r.t_arrival + r.deadline - now if r.deadline is not None else float("inf")
            if left <= 0 or (r.done > 0 and self._estimate(r, now) > left):
                self._fail_timeout(r, now)
            else:
                keep.append(r)
        self._long = keep
        if not keep:
            return []
        best = min(keep, key=lambda r: (self._order_key(r, now), r.t_arrival))
        keep.remove(best)
        return [best]

    def _take_batch(self, slot):
        """Called with the lock held: (graph, requests, schedules, footprint,
        new) or None when nothing fits.  Slot 0 takes new requests first, the
        others continuing ones first."""
        now = time.perf_counter()
        new_first = slot == 0 or not self._long
        batch, new = [], False
        if new_first and self._pending:
            batch, new = self._take_new(now), True
        if not batch and self._long:
            batch, new = self._take_long(now), False
        if not batch and self._pending:
            batch, new = self._take_new(now), True
        if not batch:
            return None
        graph = batch[0].graph
        scheds = [r.sched for r in batch]
        dev = self._dev[graph]
        tun = self._new_tuning if new else self._slice_tuning
        try:
            fp = dev.footprint(scheds, tun)
            # admit the longest prefix that fits the free budget (at least one
            # request when nothing else is in flight)
            while batch and self._in_flight + fp > self._budget and (self._in_flight > 0 or len(batch) > 1):
                r = batch.pop()
                (self._pending.appendleft if new else self._long.append)(r)
                scheds.pop()
                fp = dev.footprint(scheds, tun) if scheds else 0
        except Exception as e:
            # the requests taken out of the queue fail; the slot keeps serving
            for r in batch:
                r.error = e
                r.t_done = now
                r._event.set()
            return None
        if not batch:
            return None
        self._in_flight += fp
        return (graph, batch, scheds, fp, new)

    def _run(self, slot):
        while True:
            with self._cv:
                item = None
                while item is None:
                    if self._closed and not self._pending and not self._long:
                        return
                    item = self._take_batch(slot)
                    if item is None:
                        self._cv.wait(0.01)
            graph, reqs, scheds, fp, new = item
            t0 = time.perf_counter()
            for r in reqs:
                if r.t_start is None:
                    r.t_start = t0
            # each request's deadline is its count's abort time: a launch still
            # on its tasks then abandons them, its launches before stand and
            # the batch's other requests carry on
            abort = [max(r.t_arrival + r.deadline - t0, 1e-3) if r.deadline is not None else 0.0 for r in reqs]
            tun = self._new_tuning if new else self._slice_tuning
            try:
                done, raws, st, _, aborted = self._dev[graph].count_batch_progress(scheds, [r.done for r in reqs],
                                                                                   tun, abort)
                err = None
            except Exception as e:
                done = raws = aborted = st = None
                err = e
            t1 = time.perf_counter()
            finished, again, unstarted = [], [], []
            for i, r in enumerate(reqs):
                r.slices += 1
                if err is not None:
                    r.error = err
                    finished.append(r)
                    continue
                progressed = done[i] > r.done
                if progressed or aborted[i]:
                    r.gpu_s += (t1 - t0) / len(reqs)
                r.raw_acc = (r.raw_acc + raws[i]) % (1 << 64)
                r.done = done[i]
                if r.done >= r.total:
                    signed = r.raw_acc - (1 << 64) if r.raw_acc >= (1 << 63) else r.raw_acc
                    red = r.sched.redundancy or 1
                    q = abs(signed) // abs(red)
                    r.count = q if (signed >= 0) == (red > 0) else -q
                    r.raw = r.raw_acc
                    finished.append(r)
                elif r.deadline is not None and t1 - r.t_arrival > r.deadline:
                    r.error = RequestTimeout(r.rid, r.done, r.total)
                    finished.append(r)
                elif new and not progressed and not aborted[i]:
                    # the slice ended before reaching its tasks: still new
                    unstarted.append(r)
                else:
                    if aborted[i]:
                        r.heavy = True
                    again.append(r)
            with self._cv:
                self._in_flight -= fp
                self.batches.append((graph, len(reqs), fp, t1 - t0, st.kernel_ms if st is not None else 0.0,
                                     st.launches if st is not None else 0))
                self._long.extend(again)
                self._pending.extendleft(reversed(unstarted))
                self._cv.notify_all()
            for r in finished:
                r.t_done = t1
                r._event.set()


# ---------------------------------------------------------------------------
# synthetic load
# ---------------------------------------------------------------------------

def random_pattern(rng, k, p_extra=0.3):
    """A connected k-vertex pattern: a random spanning tree (each vertex i > 0
    joins a uniformly chosen earlier vertex, then the labels are shuffled)
    plus every other pair independently with probability p_extra."""
    adj = np.zeros((k, k), dtype=np.int32)
    for i in range(1, k):
        j = int(rng.integers(0, i))
        adj[i, j] = adj[j, i] = 1
    for i in range(k):
        for j in range(i + 1, k):
            if not adj[i, j] and rng.random() < p_extra:
                adj[i, j] = adj[j, i] = 1
    perm = rng.permutation(k)
    return adj[np.ix_(perm, perm)]


def poisson_arrivals(rng, rate, n):
    """n arrival offsets (seconds) of a Poisson process with the given rate."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def run_load(server, requests, offsets):
    """Submit requests at their arrival offsets (seconds from now) and wait for
    all of them.  Returns the wall time from the first arrival to the last
    completion."""
    t0 = time.perf_counter()
    for r, off in zip(requests, offsets):
        now = time.perf_counter() - t0
        if off > now:
            time.sleep(off - now)
        r.t_arrival = time.perf_counter()
        server.submit(r)
    for r in requests:
        r._event.wait()
    return max(r.t_done for r in requests) - t0
