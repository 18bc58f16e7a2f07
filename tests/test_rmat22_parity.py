"""The plain drop-in's headline graph, R-MAT scale 22, edge factor 8,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2 (2,009,103 vertices, 32,620,751
undirected edges): device counts against the oracle's restatement of the
reference loop nest (src/graph.cpp:199-245), bit for bit.

- triangles: the whole graph (src/graph.cpp:112-139);
- 4-clique: the whole count, several minutes of host OpenMP;
- House and the 5-cycle: their full counts take the host hours to days, so
  they are checked on slices spread over the id range.  A device task range
  [vertex[a], vertex[b]) holds the edge slots whose row is in [a, b).  For
  House that row is v0, the oracle's root range.  The 5-cycle's plan reads a
  slot reversed (its row is v1, so that a fetch's tasks share v1), which the
  oracle matches with its depth-1 filter; with no_pair_v1 the row is v0
  again and the oracle's root range applies.
"""
import os
import time

import pytest

import graphpi_amd as gp
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _threads():
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


@pytest.fixture(scope="module")
def rmat22():
    t0 = time.time()
    host = gp.HostGraph.rmat(22, 8, 0.57, 0.19, 0.19, 2)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    print(f"R-MAT s22: {host.v_cnt} vertices, {host.e_cnt // 2} edges, generated and uploaded in "
          f"{time.time() - t0:.1f} s; host threads {_threads()}", flush=True)
    return host, dev, tri


def _sched(host, name, tri):
    s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    assert valid
    return s


def _slots(host, a, b):
    v = host.vertex
    return int(v[a]), int(v[b])


def test_rmat22_shape_and_triangles(rmat22):
    host, dev, tri = rmat22
    assert (host.v_cnt, host.e_cnt) == (2009103, 65241502)
    t0 = time.time()
    want = ob.triangle_counting(host, threads=_threads())
    print(f"triangles {tri} (oracle {time.time() - t0:.1f} s)", flush=True)
    assert tri == want == 454437857


def test_rmat22_4clique_whole_count(rmat22):
    host, dev, tri = rmat22
    s = _sched(host, "4clique", tri)
    sliced = gp.Tuning()
    sliced.slice_ms = 50                     # launches that stop fetching after 50 ms, each resuming the last
    got = dev.count(s, sliced)
    one = gp.Tuning()
    one.slice_ms = -1
    single = dev.count(s, one)               # one launch
    default = dev.count(s)                   # 2 s slices: one launch for this count
    print(f"4-clique device {got.count} in {got.stats.kernel_ms:.1f} ms over {got.stats.launches} launches "
          f"(one launch {single.stats.kernel_ms:.1f} ms)", flush=True)
    assert got.stats.launches > 1 and single.stats.launches == 1
    assert (got.raw, got.count) == (single.raw, single.count) == (default.raw, default.count)
    # the oracle's whole count as 16 interleaved root slices (roots i, i + 16,
    # ...), whose raw sums add up; each slice reports when it is done
    t0 = time.time()
    raw = 0
    for i in range(16):
        part = ob.pattern_matching(host, s, threads=_threads(), root_begin=i, root_stride=16)
        raw = (raw + part["raw"]) % (1 << 64)
        print(f"  oracle 4-clique slice {i + 1}/16: {time.time() - t0:.1f} s", flush=True)
    print(f"4-clique oracle raw {raw} in {time.time() - t0:.1f} s on {_threads()} threads", flush=True)
    assert got.raw == raw and got.count == 12898716441 == raw // (s.redundancy or 1)


HOUSE_ROOTS = [(40, 41), (4000, 4001), (40000, 40016), (400000, 400256), (1000000, 1016384)]


@pytest.mark.parametrize("a,b", HOUSE_ROOTS)
def test_rmat22_house_root_slices(rmat22, a, b):
    host, dev, tri = rmat22
    s = _sched(host, "house", tri)
    t = gp.Tuning()
    t.task_begin, t.task_end = _slots(host, a, b)
    got = dev.count(s, t)
    t0 = time.time()
    want = ob.pattern_matching(host, s, threads=_threads(), root_begin=a, root_end=b)
    print(f"House roots [{a}, {b}): device raw {got.raw} ({got.stats.kernel_ms:.1f} ms), oracle "
          f"{time.time() - t0:.1f} s", flush=True)
    assert got.raw == want["raw"]


PENTAGON_V1 = [(40000, 40001), (400000, 400064), (1000000, 1000020), (1500000, 1500100), (1900000, 1900500)]


@pytest.mark.parametrize("a,b", PENTAGON_V1)
def test_rmat22_pentagon_v1_slices(rmat22, a, b):
    host, dev, tri = rmat22
    s = _sched(host, "pentagon", tri)
    assert gp.kernel_plan(s)["pair_grouped_v1"] == 1
    t = gp.Tuning()
    t.task_begin, t.task_end = _slots(host, a, b)
    got = dev.count(s, t)
    t0 = time.time()
    want = ob.pattern_matching(host, s, threads=_threads(), v1_begin=a, v1_end=b)
    print(f"5-cycle v1 in [{a}, {b}): device raw {got.raw} ({got.stats.kernel_ms:.1f} ms), oracle "
          f"{time.time() - t0:.1f} s", flush=True)
    assert got.raw == want["raw"]


PENTAGON_V0 = [(1000000, 1000030), (1500000, 1500200), (2000000, 2002000)]


@pytest.mark.parametrize("a,b", PENTAGON_V0)
def test_rmat22_pentagon_root_slices_without_grouping(rmat22, a, b):
    host, dev, tri = rmat22
    s = _sched(host, "pentagon", tri)
    t = gp.Tuning()
    t.no_pair_v1 = 1
    t.task_begin, t.task_end = _slots(host, a, b)
    got = dev.count(s, t)
    t0 = time.time()
    want = ob.pattern_matching(host, s, threads=_threads(), root_begin=a, root_end=b)
    print(f"5-cycle (no grouping) roots [{a}, {b}): device raw {got.raw} ({got.stats.kernel_ms:.1f} ms), "
          f"oracle {time.time() - t0:.1f} s", flush=True)
    assert got.raw == want["raw"]
