"""gp_count_noalloc, the allocation-free variant of gp_count (every buffer in
one static device array), through the C ABI: equal to gp_count and the oracle
on two calls in a row, the second with a smaller input; with the wave count
capped by the array's size; with the 2^16-entry deferred list overflowing."""
import ctypes

import pytest

import graphpi_amd as gp
import oracle_binding as ob

pytestmark = pytest.mark.gpu

PATTERNS = ["triangle", "house", "4clique", "pentagon", "5clique", "tri_edge_tri", "rectangle"]


def _noalloc(dev, sched, tuning=None):
    raw = ctypes.c_uint64()
    cnt = ctypes.c_int64()
    st = gp.Stats()
    rc = gp.lib().gp_count_noalloc(dev._h, ctypes.byref(sched), ctypes.byref(tuning) if tuning else None,
                                   ctypes.byref(raw), ctypes.byref(cnt), ctypes.byref(st))
    assert rc == 0, gp.lib().gp_last_error().decode()
    return raw.value, cnt.value, st


def _check_graph(host, oracle=True):
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    for name in PATTERNS:
        s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        assert valid
        want = dev.count(s)
        raw, cnt, st = _noalloc(dev, s)
        print(f"{host.v_cnt} vertices {name}: {cnt} (noalloc {st.kernel_ms:.2f} ms on {st.waves} waves, "
              f"gp_count {want.stats.kernel_ms:.2f} ms)", flush=True)
        assert (raw, cnt) == (want.raw, want.count), name
        assert st.launches == 1 and st.tasks_total == dev.count_tasks(s)
        if oracle:
            w = ob.pattern_matching(host, s, threads=8)
            assert (raw, cnt) == (w["raw"], w["count"]), name
    return dev, tri


def test_noalloc_equals_gp_count_two_calls(rmat13_path):
    # first call: R-MAT scale 13 (edge factor 12); second, smaller: scale 9,
    # also checked against the oracle
    _check_graph(gp.HostGraph.load_text(rmat13_path), oracle=False)
    _check_graph(gp.HostGraph.rmat(9, 8, seed=3))


def test_noalloc_caps_waves_to_the_static_array():
    # R-MAT scale 16 (edge factor 16): the default 24 waves per CU would need
    # more per-wave scratch than the 128 MiB array holds
    host = gp.HostGraph.rmat(16, 16, seed=4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    for name in ["house", "4clique", "5clique"]:
        s, _ = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        raw, cnt, st = _noalloc(dev, s)
        want = dev.count(s)
        print(f"s16 {name}: noalloc {st.waves} waves {st.kernel_ms:.1f} ms, gp_count {want.stats.waves} waves "
              f"{want.stats.kernel_ms:.1f} ms", flush=True)
        assert (raw, cnt) == (want.raw, want.count), name
    assert st.waves < want.stats.waves


def test_noalloc_deferred_overflow_and_ranges(rmat13_path):
    host = gp.HostGraph.load_text(rmat13_path)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    s, _ = gp.build_schedule(gp.PATTERNS["house"], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    t = gp.Tuning()
    t.split_min = 1  # every task with more than one depth-2 candidate is deferred
    raw, cnt, st = _noalloc(dev, s, t)
    assert st.split_tasks == 1 << 16 and st.inplace_tasks > 0
    assert raw == dev.count(s).raw
    # a task range, as gp_count takes it
    r = gp.Tuning()
    r.task_begin, r.task_end = 1000, 50000
    assert _noalloc(dev, s, r)[0] == dev.count(s, r).raw
