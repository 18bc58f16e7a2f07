"""Serving on the device: requests of mixed pattern sizes (3-7 vertices, the
served configuration's mix) on two resident R-MAT graphs (edge factor 16),
admitted while others run; every result equals the same count made alone
through gp_count and the oracle's restatement of the reference loop nest
(src/graph.cpp:199-245), exact integers, tolerance 0."""
import numpy as np
import pytest

import graphpi_amd as gp
import oracle_binding as ob
from graphpi_amd.serve import CountServer, Request, percentile, poisson_arrivals, random_pattern, run_load

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone_and_oracle():
    graphs = {"s8": gp.HostGraph.rmat(8, 16, seed=5), "s9": gp.HostGraph.rmat(9, 16, seed=6)}
    rng = np.random.default_rng(11)
    reqs = [Request(i, ["s8", "s9"][int(rng.integers(0, 2))], random_pattern(rng, int(rng.integers(3, 8))))
            for i in range(60)]
    with CountServer(graphs, slots=3, max_batch=16) as srv:
        wall = run_load(srv, reqs, poisson_arrivals(rng, 400.0, len(reqs)))
        lat = [r.latency * 1e3 for r in reqs]
        print(f"{len(reqs)} requests in {wall:.3f} s, {len(srv.batches)} launches, p50 {percentile(lat, 50):.2f} ms, "
              f"p99 {percentile(lat, 99):.2f} ms", flush=True)
        assert max(b[1] for b in srv.batches) > 1  # requests shared launches
        assert {r.adj.shape[0] for r in reqs} == {3, 4, 5, 6, 7}
        alone = {name: gp.DeviceGraph(h) for name, h in graphs.items()}
        for r in reqs:
            got = r.wait()
            s = srv.schedule_for(r.graph, r.adj, r.settings)
            one = alone[r.graph].count(s)
            want = ob.pattern_matching(graphs[r.graph], s, threads=8)
            assert (got, r.raw) == (one.count, one.raw) == (want["count"], want["raw"]), r.rid


def test_time_sliced_requests_resume_exactly_and_deadlines():
    """Short time slices (2 ms): longer counts are served over several batch
    calls, resuming from the task where the last slice stopped, and still
    equal the count made alone; a request with a deadline too short for its
    count fails with RequestTimeout while the requests around it finish."""
    from graphpi_amd.serve import RequestTimeout
    graphs = {"s12": gp.HostGraph.rmat(12, 16, seed=7)}
    rng = np.random.default_rng(13)
    reqs = [Request(i, "s12", random_pattern(rng, int(rng.integers(3, 7)))) for i in range(24)]
    with CountServer(graphs, slots=2, max_batch=8, quantum_s=0.002) as srv:
        run_load(srv, reqs, poisson_arrivals(rng, 200.0, len(reqs)))
        alone = gp.DeviceGraph(graphs["s12"])
        for r in reqs:
            got = r.wait()
            one = alone.count(srv.schedule_for(r.graph, r.adj, r.settings))
            assert (got, r.raw) == (one.count, one.raw), r.rid
        print(f"slices per request: max {max(r.slices for r in reqs)}, "
              f"{sum(r.slices > 1 for r in reqs)} of {len(reqs)} took more than one", flush=True)
        assert max(r.slices for r in reqs) > 1
        # a 6-vertex motif whose count on this graph takes over a second
        late = srv.submit(Request(100, "s12", gp.motifs(6)[77], deadline=0.001))
        tri = srv.submit(Request(101, "s12", gp.PATTERNS["triangle"]))
        with pytest.raises(RequestTimeout):
            late.wait(60)
        assert late.timed_out and late.done < late.total
        assert tri.wait(60) == alone.count(srv.schedule_for("s12", tri.adj, tri.settings)).count
