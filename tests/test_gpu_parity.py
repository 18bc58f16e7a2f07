"""GPU parity: the HIP counting path against the reference's golden counts and
the oracle, bit for bit (integer counts, tolerance 0).

Every count here goes through the C ABI (gp_count / gp_tc) on the device."""
import itertools
import math

import numpy as np
import pytest

import graphpi_amd as gp
import oracle_binding as ob
from conftest import load_json

pytestmark = pytest.mark.gpu

SETTINGS = [(1, 1, 1), (1, 1, 0), (0, 1, 0), (0, 0, 0), (0, 2, 0)]
SMALL_PATTERNS = ["edge", "wedge", "triangle", "rectangle", "qg3", "house", "pentagon", "4clique", "5clique",
                  "tri_edge_tri", "hourglass", "cycle_6_tri", "6cycle", "5path", "4path"]


def _tc(host):
    return ob.triangle_counting(host)


def _generic(**kw):
    """A tuning that keeps every count on the generic kernel (House under the
    GraphPi schedule would otherwise take its specialised two-hop kernel), so
    that the generic kernel's paths named by kw are the ones compared."""
    t = gp.Tuning()
    t.no_special = 1
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def _sched(host, adj, pm, rt, iep, tri=None):
    tri = _tc(host) if tri is None else tri
    s, valid = gp.build_schedule(adj, pm, rt, iep, max(host.v_cnt, 1), max(host.e_cnt, 1), tri)
    assert valid
    return s


def _random_graph(n, p, seed, hubs=0):
    rng = np.random.default_rng(seed)
    a = rng.random((n, n)) < p
    for h in range(hubs):  # a few high-degree vertices
        a[h, :] = rng.random(n) < 0.7
    a = np.triu(a, 1)
    e = np.argwhere(a)
    return gp.HostGraph.from_edges(e)


@pytest.fixture(scope="module")
def wiki(wiki_vote_path):
    h = gp.HostGraph.load_text(wiki_vote_path)
    return h, gp.DeviceGraph(h)


def test_device_present():
    assert gp.device_count() >= 1


def test_wiki_vote_reference_counts(wiki):
    host, dev = wiki
    recs = [r for r in load_json("reference_counts.json") if r["graph"] == "wiki-vote"]
    assert recs
    assert host.hash() == recs[0]["csr_fnv1a"]
    for r in recs:
        a = gp.adj_from_string(len(r["pattern_adj"]) and int(math.isqrt(len(r["pattern_adj"]))), r["pattern_adj"])
        s, valid = gp.build_schedule(a, r["pm"], r["rt"], r["iep"], r["v_cnt"], r["e_cnt"], r["tri_cnt"])
        assert valid
        got = dev.count(s)
        print(f"wiki-vote {r['pattern']} pm={r['pm']} gpu={got.count} ref={r['count']} "
              f"gpu_ms={got.stats.kernel_ms:.3f} ref_s={r['time_count']:.3f}", flush=True)
        assert got.count == r["count"], r["pattern"]


def test_wiki_vote_triangles_device(wiki):
    host, dev = wiki
    assert dev.triangle_counting() == 608389


def test_rmat13_reference_counts(rmat13_path):
    recs = [r for r in load_json("reference_counts.json") if r["graph"] == "rmat13"]
    assert recs
    host = gp.HostGraph.load_text(rmat13_path)
    assert host.hash() == recs[0]["csr_fnv1a"]
    dev = gp.DeviceGraph(host)
    for r in recs:
        n = int(math.isqrt(len(r["pattern_adj"])))
        a = gp.adj_from_string(n, r["pattern_adj"])
        tri = r["tri_cnt"]
        s, valid = gp.build_schedule(a, r["pm"], r["rt"], r["iep"], r["v_cnt"], r["e_cnt"], tri)
        assert valid
        got = dev.count(s)
        print(f"rmat13 {r['pattern']} ({r['pm']},{r['rt']},{r['iep']}) gpu={got.count} ref={r['count']} "
              f"gpu_ms={got.stats.kernel_ms:.3f} ref_s={r['time_count']:.3f}", flush=True)
        assert got.count == r["count"], (r["pattern"], r["pm"], r["rt"], r["iep"])
    assert dev.triangle_counting() == recs[0]["tri_cnt"]


@pytest.mark.parametrize("seed,n,p,hubs", [(1, 40, 0.25, 0), (2, 70, 0.12, 2), (3, 120, 0.05, 3)])
def test_random_graphs_match_oracle(seed, n, p, hubs):
    host = _random_graph(n, p, seed, hubs)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    for name in SMALL_PATTERNS:
        adj = gp.PATTERNS[name]
        for pm, rt, iep in SETTINGS:
            s, valid = gp.build_schedule(adj, pm, rt, iep, host.v_cnt, host.e_cnt, tri)
            if not valid:
                continue
            want = ob.pattern_matching(host, s, threads=4)
            got = dev.count(s)
            assert (got.raw, got.count) == (want["raw"], want["count"]), (name, pm, rt, iep)


def test_faithful_mode_balg_matches_oracle(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    tun = gp.Tuning()
    tun.collect_balg = 1
    for name in ["triangle", "rectangle", "house", "4clique", "pentagon", "qg3"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        want = ob.pattern_matching(host, s, threads=8)
        got = dev.count(s, tun)
        assert got.count == want["count"], name
        assert got.stats.balg_bytes == want["balg"], name


def test_batched_level_matches_per_candidate(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    tun = _generic(no_batch=1)
    cases = [(name, st) for name in ["triangle", "wedge", "rectangle", "qg3", "house", "4clique", "5clique"]
             for st in [(1, 1, 1), (0, 1, 0), (1, 1, 0)]]
    cases += [("pentagon", (1, 1, 1)), ("cycle_6_tri", (1, 1, 1))]
    for name, (pm, rt, iep) in cases:
        s = _sched(host, gp.PATTERNS[name], pm, rt, iep, tri)
        a, b = dev.count(s, _generic()), dev.count(s, tun)
        print(f"{name} ({pm},{rt},{iep}) batched {a.stats.kernel_ms:.2f} ms per-candidate "
              f"{b.stats.kernel_ms:.2f} ms", flush=True)
        assert a.raw == b.raw, (name, pm, rt, iep)


@pytest.mark.parametrize("name", ["pentagon", "6cycle", "5path", "4path", "house", "4clique"])
def test_pair_level_matches(name):
    host = gp.HostGraph.rmat(11, 8, seed=7)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    no_pair = _generic(no_pair=1)
    no_batch = _generic(no_batch=1)
    lds_map = _generic(pair_map=1)  # the pair level's LDS count map instead of the dense array
    for pm, rt, iep in [(1, 1, 1), (1, 1, 0), (0, 1, 0)]:
        s, valid = gp.build_schedule(gp.PATTERNS[name], pm, rt, iep, host.v_cnt, host.e_cnt, tri)
        if not valid:
            continue
        a, b, c, m = dev.count(s, _generic()), dev.count(s, no_pair), dev.count(s, no_batch), dev.count(s, lds_map)
        print(f"{name} ({pm},{rt},{iep}) pair dense {a.stats.kernel_ms:.2f} ms, LDS map {m.stats.kernel_ms:.2f} ms, "
              f"batched {b.stats.kernel_ms:.2f} ms, per-candidate {c.stats.kernel_ms:.2f} ms", flush=True)
        assert a.raw == b.raw == c.raw == m.raw, (name, pm, rt, iep)
    small = _random_graph(150, 0.06, 9, hubs=4)
    ds = gp.DeviceGraph(small)
    s = _sched(small, gp.PATTERNS[name], 1, 1, 1)
    assert ds.count(s).raw == ob.pattern_matching(small, s, threads=4)["raw"]


def test_task_ranges_add_up(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    for name in ["house", "pentagon", "4clique", "wedge", "triangle"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        whole = dev.count(s)
        for slices in (1, 7, 64):
            part = dev.count_sliced(s, slices)
            assert (part.raw, part.count) == (whole.raw, whole.count), (name, slices)
    # faithful B_alg is additive over ranges too
    s = _sched(host, gp.PATTERNS["house"], 1, 1, 1, tri)
    tun = gp.Tuning()
    tun.collect_balg = 1
    assert dev.count_sliced(s, 5, tun).stats.balg_bytes == dev.count(s, tun).stats.balg_bytes
    with pytest.raises(gp.GPError):
        bad = gp.Tuning()
        bad.task_begin, bad.task_end = 10, 5
        dev.count(s, bad)


def test_split_heavy_tasks_same_counts(rmat13_path):
    # split_min = 1 defers almost every task to the parts launch (strided
    # depth-2 candidates); -1 never splits; 0 is the default threshold
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    tuns = {}
    for v in (1, 40, -1):
        tuns[v] = _generic(split_min=v)
    no_batch = _generic(no_batch=1, split_min=1)
    cases = [(name, (1, 1, 1)) for name in ["triangle", "rectangle", "house", "4clique", "5clique", "pentagon",
                                             "qg3", "cycle_6_tri", "5path"]]
    cases += [(name, (0, 1, 0)) for name in ["house", "pentagon", "4clique"]]
    for name, (pm, rt, iep) in cases:
        s = _sched(host, gp.PATTERNS[name], pm, rt, iep, tri)
        base = dev.count(s, tuns[-1])
        assert base.stats.split_tasks == 0
        for v in (1, 40):
            r = dev.count(s, tuns[v])
            print(f"{name} ({pm},{rt},{iep}) split_min={v}: {r.stats.split_tasks} split tasks, "
                  f"{r.stats.kernel_ms:.2f} ms vs {base.stats.kernel_ms:.2f} ms", flush=True)
            assert r.raw == base.raw, (name, pm, rt, iep, v)
        assert dev.count(s).raw == dev.count(s, _generic()).raw == base.raw
        assert dev.count(s, no_batch).raw == base.raw, (name, "no_batch")
    # with task ranges: the deferred tasks of each range stay in that range
    s = _sched(host, gp.PATTERNS["house"], 1, 1, 1, tri)
    assert dev.count_sliced(s, 9, tuns[1]).raw == dev.count(s, tuns[-1]).raw


def test_no_truncation_same_counts(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    tun = _generic(no_truncate=1)
    for name in ["4clique", "5clique", "house", "tri_edge_tri"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        assert dev.count(s, tun).raw == dev.count(s, _generic()).raw == dev.count(s).raw, name


@pytest.mark.parametrize("wpc,chunk", [(1, 1), (4, 3), (32, 64)])
def test_launch_shapes_same_counts(rmat13_path, wpc, chunk):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    tun = gp.Tuning()
    tun.waves_per_cu = wpc
    tun.chunk = chunk
    gen = _generic(waves_per_cu=wpc, chunk=chunk)
    for name in ["house", "4clique"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        assert dev.count(s, tun).count == dev.count(s, gen).count == dev.count(s).count


def _perm_count(n, k):
    return math.factorial(n) // math.factorial(n - k)


def _ext(n, pairs):
    return sum(1 for perm in itertools.permutations(range(n)) if all(perm[b] < perm[a] for a, b in pairs))


def test_complete_graph_closed_form():
    # On K_n every candidate set is "all unused vertices": the raw sum is
    # C(n, k) * L(pairs checked by the loops) with L the number of orderings
    # obeying them; the reference divides by redundancy = L(front) // L(all)
    # (integer division, src/schedule.cpp:1762).  Where that ratio is exact
    # the result is the true embedding count n!/(n-k)!/|Aut|.
    for n in (5, 8, 12):
        host = gp.HostGraph.complete(n)
        dev = gp.DeviceGraph(host)
        for name in ["triangle", "rectangle", "house", "pentagon", "4clique", "5clique", "hourglass"]:
            adj = gp.PATTERNS[name]
            k = adj.shape[0]
            if k > n:
                continue
            s, _ = gp.build_schedule(adj, 1, 1, 1, host.v_cnt, host.e_cnt, math.comb(n, 3))
            d = s.to_dict()
            pairs = [tuple(p) for p in d["restrict_pair"]]
            front = [p for p in pairs if s.iep_num < 2 or p[1] < k - s.iep_num]
            lf, la = _ext(k, front), _ext(k, pairs)
            assert s.redundancy == lf // la
            got = dev.pattern_matching(s)
            assert got == math.comb(n, k) * lf // s.redundancy, (n, name)
            if lf % la == 0:
                assert got == _perm_count(n, k) // s.multiplicity, (n, name)


def test_complete_bipartite_closed_form():
    a, b = 7, 9
    edges = [(i, a + j) for i in range(a) for j in range(b)]
    host = gp.HostGraph.from_edges(edges)
    dev = gp.DeviceGraph(host)
    s = _sched(host, gp.PATTERNS["rectangle"], 1, 1, 1)
    assert dev.pattern_matching(s) == math.comb(a, 2) * math.comb(b, 2)
    assert dev.triangle_counting() == 0


def test_cycles_and_stars():
    c5 = gp.HostGraph.from_edges([(i, (i + 1) % 5) for i in range(5)])
    d5 = gp.DeviceGraph(c5)
    assert d5.pattern_matching(_sched(c5, gp.PATTERNS["pentagon"], 1, 1, 1)) == 1
    c7 = gp.HostGraph.from_edges([(i, (i + 1) % 7) for i in range(7)])
    d7 = gp.DeviceGraph(c7)
    assert d7.pattern_matching(_sched(c7, gp.PATTERNS["pentagon"], 1, 1, 1)) == 0
    star = gp.HostGraph.from_edges([(0, i) for i in range(1, 40)])
    ds = gp.DeviceGraph(star)
    assert ds.pattern_matching(_sched(star, gp.PATTERNS["wedge"], 1, 1, 1)) == math.comb(39, 2)


def test_empty_and_tiny_graphs():
    empty = gp.HostGraph.from_edges(np.zeros((0, 2), dtype=np.int32))
    assert empty.v_cnt == 0
    de = gp.DeviceGraph(empty)
    s, _ = gp.build_schedule(gp.PATTERNS["triangle"], 0, 1, 0, 1, 1, 0)
    assert de.pattern_matching(s) == 0
    one = gp.HostGraph.from_edges([(3, 9)])
    d1 = gp.DeviceGraph(one)
    assert d1.pattern_matching(s) == 0
    se, _ = gp.build_schedule(gp.PATTERNS["edge"], 0, 1, 0, 2, 2, 0)
    assert d1.pattern_matching(se) == 1


def test_bad_schedule_is_rejected():
    host = gp.HostGraph.complete(5)
    dev = gp.DeviceGraph(host)
    s, _ = gp.build_schedule(gp.PATTERNS["triangle"], 0, 1, 0, 5, 20, 10)
    s.redundancy = 0
    with pytest.raises(gp.GPError):
        dev.count(s)
    s2, _ = gp.build_schedule(gp.PATTERNS["triangle"], 0, 1, 0, 5, 20, 10)
    s2.loop_set[2] = 27
    with pytest.raises(gp.GPError):
        dev.count(s2)


def _motif_schedules(host, sizes, tri, settings=(1, 1, 1)):
    out = []
    for k in sizes:
        for adj in gp.motifs(k):
            s, valid = gp.build_schedule(adj, *settings, host.v_cnt, host.e_cnt, tri)
            if valid:
                out.append(s)
    return out


def test_batch_equals_single_counts(rmat13_path):
    # every connected 3-, 4- and 5-vertex pattern (2 + 6 + 21) in one launch
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    scheds = _motif_schedules(host, (3, 4, 5), tri)
    assert len(scheds) == 29
    res, st = dev.count_batch(scheds)
    single_ms = 0.0
    for s, r in zip(scheds, res):
        one = dev.count(s)
        single_ms += one.stats.kernel_ms
        assert (r.count, r.raw) == (one.count, one.raw)
    print(f"29 patterns: batch {st.kernel_ms:.2f} ms, one at a time {single_ms:.2f} ms, "
          f"{st.split_tasks} split tasks", flush=True)
    # heavy-task deferral across plans, and the per-candidate path
    tun = gp.Tuning()
    tun.split_min = 1
    res2, st2 = dev.count_batch(scheds, tun)
    assert st2.split_tasks > 0
    assert [r.raw for r in res2] == [r.raw for r in res]
    tun.no_batch = 1
    assert [r.raw for r in dev.count_batch(scheds, tun)[0]] == [r.raw for r in res]


def test_batch_matches_oracle_and_balg():
    host = _random_graph(90, 0.1, 5, hubs=3)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    scheds = _motif_schedules(host, (4, 5, 6), tri)
    assert len(scheds) == 6 + 21 + 112
    res, _ = dev.count_batch(scheds)
    for s, r in zip(scheds, res):
        want = ob.pattern_matching(host, s, threads=4)
        assert (r.count, r.raw) == (want["count"], want["raw"])
    # faithful mode: the batch's B_alg is the sum of the single counts'
    tun = gp.Tuning()
    tun.collect_balg = 1
    few = scheds[:12]
    _, st = dev.count_batch(few, tun)
    assert st.balg_bytes == sum(dev.count(s, tun).stats.balg_bytes for s in few)
    # the same schedule twice, and a batch of one
    s = scheds[3]
    twice, _ = dev.count_batch([s, s])
    assert twice[0].raw == twice[1].raw == dev.count(s).raw
    assert dev.count_batch([s])[0][0].raw == twice[0].raw


def test_batch_rejects_bad_input(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    s = _sched(host, gp.PATTERNS["house"], 1, 1, 1)
    tun = gp.Tuning()
    tun.task_begin, tun.task_end = 0, 100
    with pytest.raises(gp.GPError):
        dev.count_batch([s, s], tun)
    with pytest.raises(gp.GPError):
        dev.count_batch([])
    with pytest.raises(gp.GPError):
        dev.count_batch([s] * 257)


def test_autotune_keeps_counts(rmat13_path):
    from graphpi_amd.autotune import autotune
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    for name in ["house", "4clique"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        tun, rep = autotune(dev, s)
        assert rep["sample_ms_tuned"] <= rep["sample_ms_default"]
        assert dev.count(s, tun).raw == dev.count(s).raw, (name, rep)


def test_chung_lu_skewed_graph_matches_oracle():
    # power-law degrees with a few hubs: the skew the deferred-task and
    # probing paths exist for
    host = gp.HostGraph.chung_lu(4000, 40000, 2.1, 600.0, 3)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    assert dev.triangle_counting() == tri
    for name in ["4clique", "5clique", "house", "pentagon", "rectangle", "qg3"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        want = ob.pattern_matching(host, s, threads=8)
        for split in (0, 1):
            tun = gp.Tuning()
            tun.split_min = split
            got = dev.count(s, tun)
            assert (got.raw, got.count) == (want["raw"], want["count"]), (name, split)


def _hub_graph():
    """6,000 vertices, about 35k edges, two hubs of degree about 4,200: large
    enough neighbour sets for the flagged shared search (max degree >= 4,096)
    and for sets beyond the LDS pool."""
    return _random_graph(6000, 0.0015, 11, hubs=2)


def test_shared_search_matches_per_block_search():
    """Batched leaf level: one flagged search of a containing set (default)
    against a search of every block's set (no_shared_search), over every
    connected 5- and 6-vertex motif, and House against the oracle."""
    host = _hub_graph()
    assert host.max_degree >= 4096
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    off = _generic(no_shared_search=1)
    n = shared = 0
    for k in (5, 6):
        for i, adj in enumerate(gp.motifs(k)):
            s = _sched(host, adj, 1, 1, 1, tri)
            shared += gp.kernel_plan(s)["shared_search"]
            a, b = dev.count(s, _generic()), dev.count(s, off)
            assert a.raw == b.raw == dev.count(s).raw, (k, i)
            n += 1
    assert n == 21 + 112 and shared > 0
    s = _sched(host, gp.PATTERNS["house"], 1, 1, 1, tri)
    want = ob.pattern_matching(host, s, threads=8)
    for t in (None, _generic()):
        got = dev.count(s, t)
        assert (got.raw, got.count) == (want["raw"], want["count"])


@pytest.mark.parametrize("graph", ["rmat", "chunglu"])
def test_pair_level_grouped_by_v1(graph):
    """The 5-cycle's pair level with the tasks of a fetch grouped by v1 and
    counts carried between them (default) against both sides streamed per task
    (no_pair_v1), the LDS map (pair_map=1) and the oracle.  Chunk sizes 1, 7
    and 64 start and end fetches mid-row; split_min 16 sends heavy tasks to
    the parts launch."""
    if graph == "rmat":
        host = gp.HostGraph.rmat(12, 8, 0.57, 0.19, 0.19, 2)
    else:
        host = gp.HostGraph.chung_lu(6000, 40000, 2.1, 800.0, 3)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    s = _sched(host, gp.PATTERNS["pentagon"], 1, 1, 1, tri)
    want = ob.pattern_matching(host, s, threads=8)
    base = gp.Tuning()
    base.no_pair_v1 = 1
    ref = dev.count(s, base)
    assert (ref.raw, ref.count) == (want["raw"], want["count"])
    for chunk in (0, 1, 7, 64):
        for split in (0, 16):
            t = gp.Tuning()
            t.chunk = chunk
            t.split_min = split
            got = dev.count(s, t)
            assert (got.raw, got.count) == (want["raw"], want["count"]), (chunk, split)
    m = gp.Tuning()
    m.pair_map = 1
    assert dev.count(s, m).raw == want["raw"]
    # a batch mixing the 5-cycle with other patterns (the dense array is
    # shared by the plans of a launch)
    others = [_sched(host, gp.PATTERNS[n], 1, 1, 1, tri) for n in ("house", "6cycle", "4clique")]
    res, _ = dev.count_batch([s] + others)
    assert res[0].raw == want["raw"]
    for o, r in zip(others, res[1:]):
        assert r.raw == dev.count(o, base).raw


def test_bitmap_staging_matches():
    """Batched level: sets staged as LDS bitmaps (default ratio, and one that
    takes a bitmap whenever it fits) against sorted staging only
    (bitmap_ratio < 0) and the oracle."""
    host = gp.HostGraph.rmat(12, 12, 0.57, 0.19, 0.19, 7)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    off = _generic(bitmap_ratio=-1)
    wide = _generic(bitmap_ratio=1 << 20)
    noshare = _generic(no_shared_search=1, bitmap_ratio=1 << 20)
    for name in ["triangle", "4clique", "5clique", "house", "qg3", "rectangle", "cycle_6_tri", "hourglass"]:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        want = ob.pattern_matching(host, s, threads=8)
        for t in (None, _generic(), off, wide, noshare):
            got = dev.count(s, t)
            assert (got.raw, got.count) == (want["raw"], want["count"]), name


def test_sampled_index_matches():
    """Batched level: block sets too large for the LDS pool get a sampled
    index (every 2^k-th element in LDS, the rest searched in HBM).  A small
    pool (lds_stage) on a graph with hubs forces it for House (whose shared
    search selects the extended kernel); counts equal the default run and
    the oracle."""
    host = _hub_graph()
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    s = _sched(host, gp.PATTERNS["house"], 1, 1, 1, tri)
    assert gp.kernel_plan(s)["shared_search"]
    want = ob.pattern_matching(host, s, threads=8)
    assert dev.count(s).raw == dev.count(s, _generic()).raw == want["raw"]
    for stage in (32, 100):
        small = _generic(lds_stage=stage)
        got = dev.count(s, small)
        assert (got.raw, got.count) == (want["raw"], want["count"]), stage


@pytest.mark.parametrize("graph", ["rmat13", "hubs", "chunglu", "wiki"])
def test_house_twohop_kernel_matches(graph, rmat13_path, wiki_vote_path):
    """House under the GraphPi schedule runs the specialised two-hop kernel
    (per-root counts |N(x) ∩ N(v0)| and edge supports); it must equal the
    generic kernel (no_special), the oracle, and add up over task ranges;
    without restrictions (rt = 0) the unrestricted variant runs."""
    if graph == "rmat13":
        host = gp.HostGraph.load_text(rmat13_path)
    elif graph == "hubs":
        host = _hub_graph()
    elif graph == "chunglu":
        host = gp.HostGraph.chung_lu(6000, 40000, 2.1, 800.0, 3)
    else:
        host = gp.HostGraph.load_text(wiki_vote_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    generic = gp.Tuning()
    generic.no_special = 1
    for rt, shape in [(1, 1), (0, 2)]:
        s = _sched(host, gp.PATTERNS["house"], 1, rt, 1, tri)
        assert gp.kernel_plan(s)["shape"] == shape
        got = dev.count(s)
        assert got.stats.kernel_variant == 2
        want = ob.pattern_matching(host, s, threads=8)
        gen = dev.count(s, generic)
        print(f"{graph} House rt={rt}: {got.count} two-hop {got.stats.kernel_ms:.2f} ms, generic "
              f"{gen.stats.kernel_ms:.2f} ms", flush=True)
        assert (got.raw, got.count) == (want["raw"], want["count"]) == (gen.raw, gen.count), (graph, rt)
        for slices in (3, 17):
            assert dev.count_sliced(s, slices).raw == want["raw"], (graph, rt, slices)
    if graph == "wiki":
        gold = [r["count"] for r in load_json("reference_counts.json")
                if r["graph"] == "wiki-vote" and r["pattern"] == "house" and (r["pm"], r["rt"], r["iep"]) == (1, 1, 1)]
        assert dev.pattern_matching(_sched(host, gp.PATTERNS["house"], 1, 1, 1, tri)) == gold[0]


CFG3_STATS = (2999998, 234478088, 3556476596)  # the 3M-vertex Chung-Lu graph's v_cnt, e_cnt, triangles


@pytest.mark.parametrize("graph", ["rand1", "rand2", "hubs", "wiki"])
def test_tri_edge_tri_closed_form_matches(graph, wiki_vote_path):
    """Two triangles joined by an edge under the GraphPi schedule (the one the
    planner builds for the 3M-vertex Chung-Lu graph's statistics, and for
    wiki-vote's) is counted in closed form from triangles per vertex, edge
    supports and the 4-clique count.  It must equal the oracle's loop nest,
    raw sum included (4 x the count), the generic kernel (no_special), and
    the generic kernel's task ranges added up."""
    if graph == "rand1":
        host = _random_graph(60, 0.2, 31)
    elif graph == "rand2":
        host = _random_graph(150, 0.06, 32, hubs=3)
    elif graph == "hubs":
        host = gp.HostGraph.chung_lu(2000, 16000, 2.1, 200.0, 3)
    else:
        host = gp.HostGraph.load_text(wiki_vote_path)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    stats = CFG3_STATS if graph != "wiki" else (host.v_cnt, host.e_cnt, tri)
    s, valid = gp.build_schedule(gp.PATTERNS["tri_edge_tri"], 1, 1, True, *stats)
    assert valid and gp.kernel_plan(s)["shape"] == 3 and s.redundancy == 4
    got = dev.count(s)
    assert got.stats.kernel_variant == 3
    gen = dev.count(s, _generic())
    print(f"{graph} tri_edge_tri: {got.count} closed form {got.stats.kernel_ms:.2f} ms, generic "
          f"{gen.stats.kernel_ms:.2f} ms", flush=True)
    assert (got.raw, got.count) == (gen.raw, gen.count)
    assert got.raw == (4 * got.count) % (1 << 64)
    assert got.stats.exact_count() == got.count  # no wrap-around at these sizes
    if graph == "wiki":
        gold = [r["count"] for r in load_json("reference_counts.json")
                if r["graph"] == "wiki-vote" and r["pattern"] == "tri_edge_tri"]
        assert got.count == gold[0]
    else:
        want = ob.pattern_matching(host, s, threads=8)
        assert (got.raw, got.count) == (want["raw"], want["count"])
    # task ranges take the generic kernel and still add up to the same raw sum
    assert dev.count_sliced(s, 5).raw == got.raw


@pytest.mark.parametrize("graph", ["dense", "rmat15"])
def test_merge_path_intersections_match(graph):
    """Prefix builds by merge path (tiles staged in LDS) against the
    per-element binary search (merge_ratio < 0) and the oracle.  The dense
    graph with a small LDS pool (64-element tiles) merges most of its
    depth-1 and depth-2 builds; R-MAT scale 15 with the default pool merges
    the hub-against-hub builds only."""
    if graph == "dense":
        host = _random_graph(600, 0.5, 41)
        names = ["4clique", "5clique", "house", "qg3", "rectangle"]
        extra = dict(lds_stage=64)
    else:
        host = gp.HostGraph.rmat(15, 16, 0.57, 0.19, 0.19, 4)
        names = ["4clique", "5clique", "house"]
        extra = {}
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    search = _generic(merge_ratio=-1, **extra)
    merge = _generic(merge_ratio=1 << 20, **extra)
    for name in names:
        s = _sched(host, gp.PATTERNS[name], 1, 1, 1, tri)
        want = ob.pattern_matching(host, s, threads=16)
        a, b = dev.count(s, search), dev.count(s, merge)
        print(f"{graph} {name}: search {a.stats.kernel_ms:.2f} ms, merge {b.stats.kernel_ms:.2f} ms", flush=True)
        assert (a.raw, a.count) == (b.raw, b.count) == (want["raw"], want["count"]), name
        assert dev.count(s).raw == want["raw"]  # the default ratio


def test_motif_batch_rmat_ef16_matches_oracle():
    """The batched configuration's workload at a scale the oracle finishes in
    seconds: all 133 connected 5- and 6-vertex patterns (MotifGenerator
    order) over R-MAT scale 8, edge factor 16, seed 4, in one gp_count_batch
    call; every count against the oracle's restatement of the reference loop
    nest and against gp_count alone (which takes the specialised House
    kernel where the batch takes the generic one)."""
    host = gp.HostGraph.rmat(8, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = _tc(host)
    scheds = _motif_schedules(host, (5, 6), tri)
    assert len(scheds) == 133
    res, st = dev.count_batch(scheds)
    for i, (s, r) in enumerate(zip(scheds, res)):
        want = ob.pattern_matching(host, s, threads=16)
        alone = dev.count(s)
        assert (r.raw, r.count) == (want["raw"], want["count"]) == (alone.raw, alone.count), i
    print(f"133 motifs on R-MAT s8 ef16: batch {st.kernel_ms:.1f} ms", flush=True)
