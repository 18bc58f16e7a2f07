"""Pins for the oracle (oracle/gp_oracle.c) before it is trusted as the checker.

- golden counts written by the reference build (wiki-vote, R-MAT scale 13)
- wiki-vote triangles = 608,389 (reference include/dataloader.h:23)
- an independent brute force (no schedule, restrictions or IEP), itself
  checked against networkx subgraph monomorphisms
- closed forms (K_n, K_{a,b}, cycles) and invariants (IEP on/off, restricted x
  |Aut| = unrestricted, relabelling)
- B_alg on hand-counted tiny cases
"""
import itertools
import math

import numpy as np
import pytest

import graphpi_amd as gp
import oracle_binding as ob
from conftest import load_json


def _random_graph(n, p, seed, hubs=0):
    rng = np.random.default_rng(seed)
    a = rng.random((n, n)) < p
    for h in range(hubs):
        a[h, :] = rng.random(n) < 0.7
    a = np.triu(a, 1)
    return gp.HostGraph.from_edges(np.argwhere(a))


def _sched(host, adj, pm=1, rt=1, iep=1):
    tri = ob.triangle_counting(host)
    s, valid = gp.build_schedule(adj, pm, rt, iep, max(host.v_cnt, 1), max(host.e_cnt, 1), tri)
    assert valid
    return s


def test_wiki_vote_triangles(wiki_vote_path):
    h = gp.HostGraph.load_text(wiki_vote_path)
    assert ob.triangle_counting(h) == 608389


@pytest.mark.parametrize("graph", ["wiki-vote", "rmat13"])
def test_reference_golden_counts(graph, wiki_vote_path, rmat13_path):
    recs = [r for r in load_json("reference_counts.json") if r["graph"] == graph and r["time_count"] < 3.0]
    assert recs
    h = gp.HostGraph.load_text(wiki_vote_path if graph == "wiki-vote" else rmat13_path)
    assert h.hash() == recs[0]["csr_fnv1a"]
    assert ob.triangle_counting(h) == recs[0]["tri_cnt"]
    for r in recs:
        n = math.isqrt(len(r["pattern_adj"]))
        s, valid = gp.build_schedule(gp.adj_from_string(n, r["pattern_adj"]), r["pm"], r["rt"], r["iep"],
                                     r["v_cnt"], r["e_cnt"], r["tri_cnt"])
        assert valid
        assert ob.pattern_matching(h, s, threads=8)["count"] == r["count"], r["pattern"]


def test_brute_force_matches_networkx():
    nx = pytest.importorskip("networkx")
    from networkx.algorithms import isomorphism as iso
    for seed in range(3):
        host = _random_graph(12, 0.4, seed)
        G = nx.Graph()
        G.add_nodes_from(range(host.v_cnt))
        v, e = host.vertex, host.edge
        for u in range(host.v_cnt):
            for w in e[v[u]:v[u + 1]]:
                G.add_edge(u, int(w))
        for name in ["triangle", "rectangle", "house", "4clique", "wedge", "pentagon"]:
            adj = gp.PATTERNS[name]
            P = nx.from_numpy_array(adj)
            mono = sum(1 for _ in iso.GraphMatcher(G, P).subgraph_monomorphisms_iter())
            aut = sum(1 for _ in iso.GraphMatcher(P, P).isomorphisms_iter())
            assert ob.brute_force(host, adj) == mono // aut, (seed, name)


@pytest.mark.parametrize("seed,n,p,hubs", [(1, 30, 0.3, 0), (2, 45, 0.15, 2)])
def test_oracle_matches_brute_force_all_small_motifs(seed, n, p, hubs):
    host = _random_graph(n, p, seed, hubs)
    for size in (3, 4, 5):
        for adj in gp.motifs(size):
            want = ob.brute_force(host, adj)
            for pm, rt, iep in [(1, 1, 1), (0, 0, 0), (0, 1, 0), (1, 1, 0), (0, 2, 0)]:
                s = _sched(host, adj, pm, rt, iep)
                got = ob.pattern_matching(host, s, threads=2)["count"]
                mult = 1 if rt else s.multiplicity
                assert got == want * mult, (size, adj.tolist(), pm, rt, iep)


def _exact(s):
    """True when the IEP redundancy L(front)/L(all) is an integer, i.e. the
    reference's division does not truncate (src/schedule.cpp:1762)."""
    d = s.to_dict()
    k, n = d["iep_num"], d["size"]
    pairs = [tuple(p) for p in d["restrict_pair"]]
    front = [p for p in pairs if k < 2 or p[1] < n - k]
    ext = lambda ps: sum(1 for perm in itertools.permutations(range(n)) if all(perm[b] < perm[a] for a, b in ps))
    return ext(front) % ext(pairs) == 0


def test_oracle_six_vertex_patterns_vs_brute_force():
    host = _random_graph(22, 0.35, 7)
    for name in ["tri_edge_tri", "hourglass", "cycle_6_tri"]:
        adj = gp.PATTERNS[name]
        s = _sched(host, adj)
        got = ob.pattern_matching(host, s, threads=4)["count"]
        if _exact(s):
            assert got == ob.brute_force(host, adj), name
        else:
            # the reference's planner puts restrictions on the IEP depths and
            # truncates the redundancy; its count is then not the embedding
            # count.  The oracle follows the reference; IEP off is exact.
            assert name == "hourglass"
            assert got != ob.brute_force(host, adj)
            assert ob.pattern_matching(host, _sched(host, adj, 1, 1, 0), threads=4)["count"] == ob.brute_force(host, adj)


def test_complete_graph_closed_form():
    for n in (6, 9):
        host = gp.HostGraph.complete(n)
        for name in ["triangle", "rectangle", "house", "pentagon", "4clique", "5clique"]:
            adj = gp.PATTERNS[name]
            k = adj.shape[0]
            s = _sched(host, adj)
            want = math.factorial(n) // math.factorial(n - k) // s.multiplicity
            assert ob.pattern_matching(host, s, threads=2)["count"] == want, (n, name)


def test_complete_bipartite_and_cycles():
    a, b = 6, 8
    host = gp.HostGraph.from_edges([(i, a + j) for i in range(a) for j in range(b)])
    assert ob.pattern_matching(host, _sched(host, gp.PATTERNS["rectangle"]))["count"] == math.comb(a, 2) * math.comb(b, 2)
    assert ob.triangle_counting(host) == 0
    for m in (5, 6, 9):
        c = gp.HostGraph.from_edges([(i, (i + 1) % m) for i in range(m)])
        assert ob.pattern_matching(c, _sched(c, gp.PATTERNS["pentagon"]))["count"] == (1 if m == 5 else 0)


def test_invariants_relabel_and_settings():
    host = _random_graph(40, 0.2, 11, hubs=1)
    v, e = host.vertex, host.edge
    edges = [(u, int(w)) for u in range(host.v_cnt) for w in e[v[u]:v[u + 1]] if u < w]
    perm = np.random.default_rng(0).permutation(host.v_cnt)
    relabelled = gp.HostGraph.from_edges([(int(perm[x]), int(perm[y])) for x, y in edges])
    for name in ["house", "rectangle", "4clique", "qg3"]:
        adj = gp.PATTERNS[name]
        base = ob.pattern_matching(host, _sched(host, adj))["count"]
        assert ob.pattern_matching(relabelled, _sched(relabelled, adj))["count"] == base
        assert ob.pattern_matching(host, _sched(host, adj, 1, 1, 0))["count"] == base
        s0 = _sched(host, adj, 0, 0, 0)
        assert ob.pattern_matching(host, s0)["count"] == base * s0.multiplicity


def test_root_sampling_partitions_the_sum():
    host = _random_graph(50, 0.2, 5, hubs=2)
    s = _sched(host, gp.PATTERNS["house"])
    full = ob.pattern_matching(host, s)["raw"]
    parts = sum(ob.pattern_matching(host, s, root_begin=r, root_stride=3)["raw"] for r in range(3))
    assert parts % (1 << 64) == full


def test_balg_hand_counted():
    # path 0-1-2 plus triangle-free: triangle schedule with restrictions
    # v1 < v0, v2 < v1 on the graph 0-1, 1-2, 0-2 (one triangle)
    host = gp.HostGraph.from_edges([(0, 1), (1, 2), (0, 2)])
    s = _sched(host, gp.PATTERNS["triangle"], 0, 1, 0)
    r = ob.pattern_matching(host, s, threads=1)
    assert r["count"] == 1
    # roots: 3 lookups; depth-1 candidates below v0: vertex 1 has one (0),
    # vertex 2 has two (0, 1) -> 3 lookups, each building N(v0) ∩ N(v1)
    # (2 + 2 elements); 6 lookups x 8 B + 3 x 16 B
    assert r["expanded"] == 3
    assert r["balg"] == 6 * 8 + 3 * 4 * 4
