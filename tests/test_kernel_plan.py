"""gp_kernel_plan: the kernel paths the planner's schedules map to (host only).

Schedules are built for the R-MAT scale-22 graph's statistics (2,009,103
vertices, 65,241,502 directed edges, 454,437,857 triangles) with pm=1, rt=1,
IEP on, so they are the schedules that graph's counts run."""
import pytest

import graphpi_amd as gp

STATS = (2009103, 65241502, 454437857)

# pattern: (leaf, iep_k, root_tasks, batched, pair, pair_grouped_v1, shared_search, blocks, kind1_blocks)
EXPECTED = {
    "edge": (1, 0, 1, 0, 0, 0, 0, 0, 0),
    "triangle": (2, 0, 0, 0, 0, 0, 0, 0, 0),
    "4clique": (3, 0, 0, 1, 0, 0, 0, 1, 1),
    "5clique": (4, 0, 0, 1, 0, 0, 0, 1, 1),
    "house": (3, 2, 0, 1, 0, 0, 1, 3, 2),
    "pentagon": (4, 0, 0, 1, 1, 1, 0, 1, 1),
    "6cycle": (5, 0, 0, 1, 0, 0, 0, 1, 1),
    "tri_edge_tri": (4, 2, 0, 1, 0, 0, 1, 3, 2),
    "cycle_6_tri": (3, 3, 0, 1, 0, 0, 0, 7, 6),
    "hourglass": (4, 2, 0, 1, 0, 0, 0, 3, 3),
    "qg3": (2, 2, 0, 0, 0, 0, 0, 0, 0),
}
FIELDS = ("leaf", "iep_k", "root_tasks", "batched", "pair", "pair_grouped_v1", "shared_search", "blocks",
          "kind1_blocks")


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_kernel_paths(name):
    s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, *STATS)
    assert valid
    info = gp.kernel_plan(s)
    assert tuple(info[f] for f in FIELDS) == EXPECTED[name], info
    # the IEP width and leaf depth agree with the schedule itself
    k = s.iep_num if s.iep_num > 1 else 0
    assert info["iep_k"] == k
    assert info["leaf"] == (s.size - k if k else s.size - 1)


def test_every_motif_has_a_plan():
    """Every connected 3-6-vertex motif maps to a plan; the batched level
    covers each one whose leaf is at depth 3 or more with IEP width <= 4."""
    n = 0
    for size in (3, 4, 5, 6):
        for adj in gp.motifs(size):
            s, valid = gp.build_schedule(adj, 1, 1, True, *STATS)
            assert valid
            info = gp.kernel_plan(s)
            assert info["batched"] == int(info["leaf"] >= 3 and info["iep_k"] <= 4)
            assert not info["pair"] or (info["batched"] and info["iep_k"] == 0)
            assert not info["shared_search"] or info["kind1_blocks"] >= 2
            n += 1
    assert n == 2 + 6 + 21 + 112


def test_specialised_shapes():
    # House under the GraphPi schedule takes the two-hop kernel (shape 1);
    # without restrictions the same loop nest is shape 2; other orders and
    # every other pattern stay on the generic kernel (shape 0)
    def shape(name, pm=1, rt=1, iep=True):
        s, valid = gp.build_schedule(gp.PATTERNS[name], pm, rt, iep, *STATS)
        assert valid
        return gp.kernel_plan(s)["shape"]
    assert shape("house") == 1
    assert shape("house", 0, 1) == 1
    assert shape("house", 1, 0) == 2
    assert shape("house", 1, 1, False) == 0   # no IEP: a plain leaf at depth 4
    assert shape("house", 2, 1) == 0          # the GraphZero model picks another vertex order
    assert shape("tri_edge_tri") == 3          # closed form from triangles, supports and 4-cliques
    assert shape("tri_edge_tri", 1, 1, False) == 0
    for name in sorted(EXPECTED):
        if name not in ("house", "tri_edge_tri"):
            assert shape(name) == 0, name
    # a tampered restriction (v2 < v0 added at depth 2) is not the House shape
    s, _ = gp.build_schedule(gp.PATTERNS["house"], 1, 1, True, *STATS)
    r = s.total_restrict_num
    s.restrict_index[r] = 0
    s.restrict_next[r] = -1
    s.restrict_last[2] = r
    s.total_restrict_num = r + 1
    assert gp.kernel_plan(s)["shape"] == 0
