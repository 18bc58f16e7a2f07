"""The heavy-skew configuration's patterns (4-clique, 5-clique and two
triangles joined by an edge) on a reduced Chung-Lu graph with the same Pareto
exponent (2.1) and seed (3): 3,000 vertices, 30,000 sampled edges, expected
degrees capped at 300.  Counted on the device with the vertex ids as
generated and relabelled high degree first (the loader's oriented_type 1,
src/dataloader.cpp:96-110); every count equals the oracle's restatement of
the reference loop nest on the generated ids, since none of these patterns'
IEP redundancies leaves a remainder (a relabelling only permutes the
embeddings the loop nest enumerates)."""
import pytest

import graphpi_amd as gp
import oracle_binding as ob

pytestmark = pytest.mark.gpu

STATS = (2999998, 234478088, 3556476596)  # the full configuration's v_cnt, e_cnt, triangles


@pytest.fixture(scope="module")
def reduced():
    host = gp.HostGraph.chung_lu(3000, 30000, 2.1, 300.0, 3)
    want = {}
    for name in ["4clique", "5clique", "tri_edge_tri"]:
        s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, *STATS)
        assert valid
        want[name] = (s, ob.pattern_matching(host, s, threads=16))
        print(f"oracle {name}: {want[name][1]['count']}", flush=True)
    return host, want


@pytest.mark.parametrize("oriented", [0, 1])
def test_reduced_chunglu_all_patterns(reduced, oriented):
    host, want = reduced
    g = host.reorder(oriented) if oriented else host
    dev = gp.DeviceGraph(g)
    for name, (s, w) in want.items():
        got = dev.count(s)
        print(f"oriented {oriented} {name}: {got.count} ({got.stats.kernel_ms:.2f} ms)", flush=True)
        assert (got.raw, got.count) == (w["raw"], w["count"]), (oriented, name)
        if name == "tri_edge_tri":  # also the generic loop nest on the relabelled ids
            gen = gp.Tuning()
            gen.no_special = 1
            assert dev.count(s, gen).raw == w["raw"]
