"""Serving on the device: requests of mixed pattern sizes on two resident
graphs, admitted while others run; every result equals the same count made
alone through gp_count (exact integers, tolerance 0)."""
import numpy as np
import pytest

import graphpi_amd as gp
from graphpi_amd.serve import CountServer, Request, percentile, poisson_arrivals, random_pattern, run_load

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    graphs = {"g10": gp.HostGraph.rmat(10, 8, seed=5), "g12": gp.HostGraph.rmat(12, 8, seed=6)}
    rng = np.random.default_rng(11)
    reqs = [Request(i, ["g10", "g12"][int(rng.integers(0, 2))], random_pattern(rng, int(rng.integers(3, 7))))
            for i in range(80)]
    with CountServer(graphs, slots=3, max_batch=16) as srv:
        wall = run_load(srv, reqs, poisson_arrivals(rng, 400.0, len(reqs)))
        lat = [r.latency * 1e3 for r in reqs]
        print(f"80 requests in {wall:.3f} s, {len(srv.batches)} launches, p50 {percentile(lat, 50):.2f} ms, "
              f"p99 {percentile(lat, 99):.2f} ms", flush=True)
        assert len(srv.batches) < len(reqs)
        alone = {name: gp.DeviceGraph(h) for name, h in graphs.items()}
        for r in reqs:
            got = r.wait()
            one = alone[r.graph].count(srv.schedule_for(r.graph, r.adj, r.settings))
            assert (got, r.raw) == (one.count, one.raw), r.rid
