"""DataLoader semantics (reference src/dataloader.cpp) and the generators."""
import os

import numpy as np
import pytest

import graphpi_amd as gp
from conftest import load_json


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return path


def test_wiki_vote_csr_matches_reference(wiki_vote_path):
    rec = [r for r in load_json("reference_counts.json") if r["graph"] == "wiki-vote"][0]
    h = gp.HostGraph.load_text(wiki_vote_path)
    assert (h.v_cnt, h.e_cnt) == (rec["v_cnt"], rec["e_cnt"])
    assert h.max_degree == rec["max_degree"]
    assert h.hash() == rec["csr_fnv1a"]


def test_rmat13_csr_matches_reference(rmat13_path):
    rec = [r for r in load_json("reference_counts.json") if r["graph"] == "rmat13"][0]
    h = gp.HostGraph.load_text(rmat13_path)
    assert (h.v_cnt, h.e_cnt, h.hash()) == (rec["v_cnt"], rec["e_cnt"], rec["csr_fnv1a"])


def test_first_appearance_remap_selfloops_duplicates(tmp_path):
    # ids 10, 5, 7 appear in that order -> 0, 1, 2; the self-loop line counts
    # toward the header but is dropped; the duplicate is removed
    p = _write(tmp_path / "g.txt", "3 5\n10 5\n5 7\n7 7\n7 10\n5 10\n")
    h = gp.HostGraph.load_text(str(p))
    assert h.v_cnt == 3 and h.e_cnt == 6
    assert list(h.vertex) == [0, 2, 4, 6]
    assert list(h.edge) == [1, 2, 0, 2, 0, 1]
    assert h.max_degree == 3  # degree counted before duplicate removal


@pytest.mark.parametrize("text", ["4 3\n1 2\n2 3\n3 1\n", "3 4\n1 2\n2 3\n3 1\n", "3 2\n1 2\n2 3\n3 1\n",
                                  "3 3\n1 2\n2 x\n3 1\n", "3 3\n1 2\n2 3\n3\n"])
def test_malformed_files_rejected(tmp_path, text):
    p = _write(tmp_path / "bad.txt", text)
    with pytest.raises(gp.GPError) as e:
        gp.HostGraph.load_text(str(p))
    assert e.value.code == -4


def test_missing_file():
    with pytest.raises(gp.GPError):
        gp.HostGraph.load_text("/nonexistent/graph.txt")


def test_degree_reordering(tmp_path):
    p = _write(tmp_path / "g.txt", "4 3\n0 1\n1 2\n1 3\n")
    h = gp.HostGraph.load_text(str(p), oriented_type=1)
    deg = np.diff(h.vertex)
    assert deg[0] == 3  # the hub takes id 0 under high-degree-first
    h2 = gp.HostGraph.load_text(str(p), oriented_type=2)
    assert np.diff(h2.vertex)[-1] == 3


def test_binary_round_trip(tmp_path):
    h = gp.HostGraph.rmat(10, 8, seed=3)
    p = str(tmp_path / "g.bin")
    h.write_binary(p)
    b = gp.HostGraph.load_binary(p)
    assert (b.v_cnt, b.e_cnt, b.hash()) == (h.v_cnt, h.e_cnt, h.hash())


def test_text_round_trip_is_isomorphic(tmp_path):
    h = gp.HostGraph.rmat(10, 8, seed=4)
    p = str(tmp_path / "g.txt")
    h.write_text(p)
    r = gp.HostGraph.load_text(p)
    assert (r.v_cnt, r.e_cnt) == (h.v_cnt, h.e_cnt)
    assert sorted(np.diff(r.vertex)) == sorted(np.diff(h.vertex))


def test_generator_deterministic_and_simple():
    a = gp.HostGraph.rmat(12, 8, seed=5)
    b = gp.HostGraph.rmat(12, 8, seed=5)
    c = gp.HostGraph.rmat(12, 8, seed=6)
    assert a.hash() == b.hash() != c.hash()
    v, e = a.vertex, a.edge
    for u in range(a.v_cnt):
        row = e[v[u]:v[u + 1]]
        assert np.all(np.diff(row) > 0) and u not in row
    cl = gp.HostGraph.chung_lu(20000, 60000, 2.1, 500.0, seed=3)
    assert cl.v_cnt > 0 and cl.e_cnt > 0 and cl.max_degree <= 2000


def test_complete_graph():
    h = gp.HostGraph.complete(6)
    assert h.v_cnt == 6 and h.e_cnt == 30
    assert all(np.diff(h.vertex) == 5)


def _bin(path, v_cnt, e_cnt, max_deg, vertex, edge):
    words = [v_cnt, e_cnt, max_deg] + list(vertex) + list(edge)
    np.asarray(words, dtype=np.uint32).tofile(path)
    return str(path)


@pytest.mark.parametrize("vertex,edge,why", [
    ([0, 2, 1, 3], [1, 2, 0], "offsets go backwards"),
    ([0, 1, 2, 4], [1, 0, 0, 7], "neighbour id >= v_cnt"),
    ([0, 2, 3, 4], [2, 1, 0, 0], "unsorted list"),
    ([0, 2, 3, 4], [1, 1, 0, 0], "duplicate neighbour"),
    ([0, 2, 3, 4], [0, 1, 0, 0], "self-loop"),
    ([1, 2, 3, 4], [1, 0, 0, 0], "vertex[0] != 0"),
])
def test_malformed_binary_rejected(tmp_path, vertex, edge, why):
    p = _bin(tmp_path / "bad.bin", 3, len(edge), 2, vertex, edge)
    with pytest.raises(gp.GPError) as e:
        gp.HostGraph.load_binary(p)
    assert e.value.code == -4, why


def test_binary_vertex_count_beyond_int32_rejected(tmp_path):
    p = str(tmp_path / "huge.bin")
    np.asarray([0x80000000, 0, 0, 0], dtype=np.uint32).tofile(p)
    with pytest.raises(gp.GPError) as e:
        gp.HostGraph.load_binary(p)
    assert e.value.code == -4


@pytest.mark.parametrize("header", ["3000000000 1", "3 3000000000"])
def test_text_header_beyond_csr_range_rejected(tmp_path, header):
    p = _write(tmp_path / "huge.txt", header + "\n1 2\n")
    with pytest.raises(gp.GPError) as e:
        gp.HostGraph.load_text(str(p))
    assert e.value.code == -4


@pytest.mark.parametrize("oriented", [1, 2])
def test_reorder_equals_loading_with_orientation(tmp_path, oriented):
    # the in-memory relabelling is the text loader's oriented_type step
    # (src/dataloader.cpp:96-110): reorder(load(f, 0)) == load(f, oriented)
    for seed in (3, 4):
        h = gp.HostGraph.rmat(11, 8, seed=seed)
        p = str(tmp_path / f"g{seed}.txt")
        h.write_text(p)
        plain = gp.HostGraph.load_text(p)
        want = gp.HostGraph.load_text(p, oriented_type=oriented)
        got = plain.reorder(oriented)
        assert (got.v_cnt, got.e_cnt, got.hash()) == (want.v_cnt, want.e_cnt, want.hash())
        deg = np.diff(got.vertex.astype(np.int64))
        assert np.all(np.diff(deg) <= 0) if oriented == 1 else np.all(np.diff(deg) >= 0)
    assert plain.reorder(0).hash() == plain.hash()
    with pytest.raises(gp.GPError):
        plain.reorder(3)
