"""Graph-mode launches (gp_tuning.use_graph = 1) against stream launches.

A call context keeps one instantiated hipGraph (plan upload, counter reset by
a fill kernel, count launch, deferred-parts launch, read-back) and replays it
while the launch shape repeats, updating only the kernel arguments.  Every
replay here is checked against the stream-launch result, with eager counts and
synchronisations between replays and with changes of launch shape (rebuild)
and of batch contents (argument update only).  Counts are integers: bit-exact."""
import pytest

import graphpi_amd as gp
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _scheds(host, tri, sizes):
    out = []
    for k in sizes:
        for adj in gp.motifs(k):
            s, valid = gp.build_schedule(adj, 1, 1, True, host.v_cnt, host.e_cnt, tri)
            assert valid
            out.append(s)
    return out


def test_graph_batch_replays_equal_stream_launches():
    host = gp.HostGraph.rmat(10, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    five = _scheds(host, tri, [5])
    four = _scheds(host, tri, [4])
    tri_s, _ = gp.build_schedule(gp.PATTERNS["triangle"], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    eager5 = [r.raw for r in dev.count_batch(five)[0]]
    eager4 = [r.raw for r in dev.count_batch(four)[0]]
    # the first 4 five-vertex motifs, checked against the oracle too
    for s, raw in zip(five[:4], eager5[:4]):
        assert raw == ob.pattern_matching(host, s, threads=8)["raw"]
    g = gp.Tuning()
    g.use_graph = 1
    # build, replay, replay with other batch contents of the same shape,
    # another shape, and back; an eager count between every two replays
    seq = [five, five, five[::-1], four, five, four, four, five]
    for i, batch in enumerate(seq):
        res, st = dev.count_batch(batch, g)
        want = {id(s): w for s, w in zip(five, eager5)}
        want.update({id(s): w for s, w in zip(four, eager4)})
        got = [r.raw for r in res]
        assert got == [want[id(s)] for s in batch], i
        assert st.kernel_ms > 0
        assert dev.count(tri_s).count == tri
    dev.free()


def test_graph_single_count_with_deferred_parts():
    # a skewed graph and a small split threshold: the graph holds both the
    # task launch and the deferred-parts launch
    host = gp.HostGraph.chung_lu(20000, 200000, 2.1, 2000.0, 3)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    # (no_special keeps House on the generic kernel, whose launches graph mode
    # covers; the specialised House kernel runs as a stream launch)
    eager_tun = gp.Tuning()
    eager_tun.split_min = 16
    eager_tun.no_special = 1
    g = gp.Tuning()
    g.split_min = 16
    g.use_graph = 1
    g.no_special = 1
    for name in ["house", "pentagon", "4clique"]:
        s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        assert valid
        want = dev.count(s, eager_tun)
        assert want.stats.split_tasks > 0, name
        for rep in range(3):
            got = dev.count(s, g)
            assert (got.raw, got.count) == (want.raw, want.count), (name, rep)
            assert got.stats.split_tasks == want.stats.split_tasks
            # eager work between replays
            assert dev.count(s).raw == want.raw
    dev.free()
