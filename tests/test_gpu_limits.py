"""Device paths at the API's limits, against the oracle (bit for bit):
7- and 8-vertex patterns (IEP width up to 6, loop depths up to 7), the
deferred-task list overflowing into in-place runs, counts split into bounded
launches, the host-side time limit, and the C++ drop-in itself on the device."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import graphpi_amd as gp
import oracle_binding as ob
from conftest import load_json

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _random_graph(n, p, seed, hubs=0):
    rng = np.random.default_rng(seed)
    a = rng.random((n, n)) < p
    for h in range(hubs):
        a[h, :] = rng.random(n) < 0.7
    a = np.triu(a, 1)
    return gp.HostGraph.from_edges(np.argwhere(a))


def _star(k):
    return gp.adj_from_edges(k, [(0, i) for i in range(1, k)])


def _cycle(k):
    return gp.adj_from_edges(k, [(i, (i + 1) % k) for i in range(k)])


BIG = {
    "clique_7_minus": gp.PATTERNS["clique_7_minus"],
    "star7": _star(7),                       # IEP over 6 trailing vertices: Bell(6) = 203 groups
    "cycle7": _cycle(7),
    "cycle8": _cycle(8),                     # leaf at depth 7
    "wheel7": gp.adj_from_edges(7, [(0, i) for i in range(1, 7)] + [(i, i % 6 + 1) for i in range(1, 7)]),
    "two_squares8": gp.adj_from_edges(8, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 4), (4, 5), (5, 6), (6, 7),
                                           (7, 4)]),
    "path8": gp.adj_from_edges(8, [(i, i + 1) for i in range(7)]),
    "clique8_minus": gp.adj_from_edges(8, [(j, i) for i in range(8) for j in range(i) if (i + j) % 5]),
}

CASES = [
    ("dense40", ["clique_7_minus", "star7", "cycle7", "wheel7", "two_squares8", "clique8_minus"],
     [(1, 1, 1), (1, 1, 0)]),
    ("dense40", ["cycle8", "path8"], [(1, 1, 1)]),
    ("rand60", ["star7", "cycle7", "two_squares8", "wheel7", "clique8_minus"], [(1, 1, 1), (1, 1, 0)]),
    ("rmat10", ["star7", "clique_7_minus", "wheel7", "clique8_minus"], [(1, 1, 1)]),
]


def _graph(name):
    if name == "dense40":
        return _random_graph(40, 0.55, 23)
    if name == "rand60":
        return _random_graph(60, 0.3, 21, hubs=2)
    return gp.HostGraph.rmat(10, 8, seed=9)


@pytest.mark.parametrize("gname,names,settings", CASES)
def test_seven_and_eight_vertex_patterns(gname, names, settings):
    host = _graph(gname)
    dev = gp.DeviceGraph(host)
    tri = ob.triangle_counting(host)
    scheds = []
    for name in names:
        for st in settings:
            s, valid = gp.build_schedule(BIG[name], *st, host.v_cnt, host.e_cnt, tri)
            assert valid
            want = ob.pattern_matching(host, s, threads=8)
            got = dev.count(s)
            print(f"{gname} {name} {st}: {got.count} ({got.stats.kernel_ms:.2f} ms)", flush=True)
            assert (got.raw, got.count) == (want["raw"], want["count"]), (gname, name, st)
            scheds.append((s, want))
    # the same schedules in one batch
    res, _ = dev.count_batch([s for s, _ in scheds])
    for (s, want), r in zip(scheds, res):
        assert (r.raw, r.count) == (want["raw"], want["count"])


def test_iep_wider_than_six_is_rejected():
    # a star with 7 leaves needs IEP over 7 trailing vertices (Bell(7) = 877
    # groups), beyond the flattened schedule's tables
    with pytest.raises(gp.GPError) as e:
        gp.build_schedule(_star(8), 1, 1, True, 100, 1000, 10)
    assert e.value.code == -3


def test_deferred_list_overflow_runs_in_place():
    """split_min = 1 defers every House task with more than one depth-2
    candidate; R-MAT s17 (edge factor 16) has about 1.86M of them, more than
    the 2^20 entries of the deferred list, so the rest run in place."""
    host = gp.HostGraph.rmat(17, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    s, _ = gp.build_schedule(gp.PATTERNS["house"], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    over = gp.Tuning()
    over.split_min = 1
    over.slice_ms = -1
    over.no_special = 1  # the generic kernel (House's own kernel defers nothing)
    r = dev.count(s, over)
    print(f"deferred {r.stats.split_tasks}, in place {r.stats.inplace_tasks}, {r.stats.kernel_ms:.1f} ms", flush=True)
    assert r.stats.split_tasks == 1 << 20 and r.stats.inplace_tasks > 100000
    never = gp.Tuning()
    never.split_min = -1
    never.slice_ms = -1
    never.no_special = 1
    base = dev.count(s, never)
    assert base.stats.split_tasks == 0
    assert r.raw == base.raw == dev.count(s).raw


def test_bounded_launches_equal_one_launch():
    host = gp.HostGraph.rmat(16, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    assert dev.count_tasks(gp.build_schedule(gp.PATTERNS["house"], 1, 1, True, host.v_cnt, host.e_cnt, tri)[0]) \
        > 1 << 20
    one = gp.Tuning()
    one.slice_ms = -1
    for name, ms, special in [("house", 1, 0), ("house", 1, 1), ("4clique", 1, 0), ("pentagon", 1, 0),
                              ("rectangle", 1, 0)]:
        # launches of about ms each: the generic kernel stops fetching at
        # that deadline and the next launch resumes where it stopped; the
        # House kernel (special = 0) takes task ranges sized for ms
        tiny = gp.Tuning()
        tiny.slice_ms = ms
        tiny.no_special = special
        s, _ = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        a, b, c = dev.count(s, tiny), dev.count(s, one), dev.count(s)
        assert a.stats.launches <= 1300
        print(f"{name} (no_special {special}): {a.stats.launches} launches at slice_ms {ms} "
              f"({a.stats.kernel_ms:.1f} ms), one launch {b.stats.kernel_ms:.1f} ms, "
              f"{c.stats.launches} by default", flush=True)
        assert b.stats.launches == 1
        # (a wave checks the deadline between fetches; the 5-cycle's pair
        # level fetches 512 tasks sharing v1, which on this graph is every
        # wave's only fetch, so it is one launch whatever the slice)
        if b.stats.kernel_ms > 8 * ms and not gp.kernel_plan(s)["pair_grouped_v1"]:
            assert a.stats.launches > 2
        assert a.stats.tasks_total == b.stats.tasks_total == dev.count_tasks(s)
        assert a.stats.tasks == b.stats.tasks == dev.count_tasks(s) or special == 0
        assert (a.raw, a.count) == (b.raw, b.count) == (c.raw, c.count), name


def test_time_limit_stops_between_launches():
    host = gp.HostGraph.rmat(16, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    s, _ = gp.build_schedule(gp.PATTERNS["house"], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    t = gp.Tuning()
    t.slice_ms = 1
    t.time_limit_s = 1e-6  # passed by the time the first launch returns
    with pytest.raises(gp.GPTimeout) as e:
        dev.count(s, t)
    part = e.value.partial
    assert e.value.code == gp.GP_ETIME and part.count is None
    assert part.stats.finished == 0 and part.stats.launches == 1
    assert 0 < part.stats.tasks < part.stats.tasks_total
    # the partial sum is the count of the range that ran
    r = gp.Tuning()
    r.task_end = int(part.stats.tasks)
    r.slice_ms = -1
    assert dev.count(s, r).raw == part.raw
    # a limit that is not reached changes nothing
    t.time_limit_s = 3600.0
    assert dev.count(s, t).raw == dev.count(s).raw


def test_cpp_drop_in_counts_on_the_device(tmp_path, wiki_vote_path):
    """A driver shaped like the reference's tianhe/baseline_test.cpp, compiled
    against include/ and linked with the library: DataLoader, Pattern,
    Schedule, then g->pattern_matching(s, 24), which runs on the device."""
    src = tmp_path / "baseline.cpp"
    src.write_text(r'''
#include "common.h"
#include "dataloader.h"
#include "graph.h"
#include "pattern.h"
#include "schedule.h"
#include <cstdio>
int main(int argc, char** argv) {
    Graph* g;
    DataLoader D;
    DataType type;
    GetDataType(type, "Wiki-Vote");
    if (!D.load_data(g, type, argv[1])) return 1;
    const char* pats[2] = {"0111010011100011100001100", "0111101111011110"};  // House, 4-clique
    const int sizes[2] = {5, 4};
    for (int i = 0; i < 2; ++i) {
        Pattern p(sizes[i], pats[i]);
        bool valid;
        Schedule s(p, valid, 1, 1, true, g->v_cnt, g->e_cnt, g->tri_cnt);
        if (!valid) return 2;
        double t0 = get_wall_time();
        long long ans = g->pattern_matching(s, 24);
        std::printf("ans %lld\ntime %.6f\nlaunches %d\n", ans, get_wall_time() - t0, g->last_stats.launches);
    }
    delete g;
    return 0;
}
''')
    exe = tmp_path / "baseline"
    r = subprocess.run(["g++", "-std=c++11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        gp.LIB_PATH, "-Wl,-rpath," + os.path.dirname(gp.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), wiki_vote_path], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    ans = [int(l.split()[1]) for l in out.stdout.splitlines() if l.startswith("ans ")]
    gold = {r["pattern"]: r["count"] for r in load_json("reference_counts.json")
            if r["graph"] == "wiki-vote" and (r["pm"], r["rt"], r["iep"]) == (1, 1, 1)}
    assert ans == [gold["house"], gold["4clique"]]


def test_progress_batches_resume_exactly():
    """gp_count_batch_progress: time-sliced counts of several schedules that
    resume from where the last call stopped.  Each count, summed over the
    slices, equals gp_count's; a slice's progress is a prefix of the tasks
    (the raw sum of [0, done) equals a task-range count of the same range)."""
    host = gp.HostGraph.rmat(16, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    names = ["triangle", "4clique", "house", "pentagon", "5clique", "rectangle"]
    scheds = [gp.build_schedule(gp.PATTERNS[n], 1, 1, True, host.v_cnt, host.e_cnt, tri)[0] for n in names]
    want = [dev.count(s) for s in scheds]
    totals = [dev.count_tasks(s) for s in scheds]
    t = gp.Tuning()
    t.slice_ms = 2
    t.time_limit_s = 0.005
    begin, raws, calls = [0] * len(scheds), [0] * len(scheds), 0
    while True:
        done, raw, st, finished, _ = dev.count_batch_progress(scheds, begin, t)
        calls += 1
        assert all(b <= d <= tot for b, d, tot in zip(begin, done, totals))
        if calls == 1:
            assert not finished and st.finished == 0
            # the first slice's progress is a prefix: a task-range count of it agrees
            for i, s in enumerate(scheds):
                if 0 < done[i] < totals[i]:
                    r = gp.Tuning()
                    r.task_end = done[i]
                    r.slice_ms = -1
                    r.no_special = 1
                    assert dev.count(s, r).raw == raw[i], names[i]
        raws = [(x + y) % (1 << 64) for x, y in zip(raws, raw)]
        begin = done
        if finished:
            break
        assert calls < 2000
    print(f"{calls} progress calls", flush=True)
    assert begin == totals
    for i, s in enumerate(scheds):
        assert raws[i] == want[i].raw, names[i]
    # no time limit: one call finishes everything
    done, raw, st, finished, aborted = dev.count_batch_progress(scheds)
    assert finished and done == totals and raw == [w.raw for w in want] and not any(aborted)
    # a schedule past its abort time is abandoned without touching the
    # others: a 6-vertex motif whose count takes minutes here, stopped after
    # 0.2 s, between counts that finish
    heavy, _ = gp.build_schedule(gp.motifs(6)[77], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    mix = [scheds[1], heavy, scheds[4]]
    done, raw, st, finished, aborted = dev.count_batch_progress(mix, abort_s=[0.0, 0.2, 0.0])
    assert not finished and aborted == [False, True, False]
    assert done[0] == totals[1] and done[2] == totals[4] and done[1] < dev.count_tasks(heavy)
    assert (raw[0], raw[2]) == (want[1].raw, want[4].raw)
    # the same through a per-task limit: a limit of 0.1 ms abandons the
    # motif (whose hub tasks run far longer); the schedules it leaves are
    # exact, and a limit longer than the other counts take whole leaves them
    t = gp.Tuning()
    t.task_abort_s = 1e-4
    done, raw, st, finished, aborted = dev.count_batch_progress(mix, tuning=t)
    assert not finished and aborted[1]
    for k, i in ((0, 1), (2, 4)):
        if not aborted[k]:
            assert done[k] == totals[i] and raw[k] == want[i].raw
    t.task_abort_s = 5.0
    done, raw, st, finished, aborted = dev.count_batch_progress([scheds[1], scheds[4]], tuning=t)
    assert finished and not any(aborted) and raw == [want[1].raw, want[4].raw]


def test_hard_limit_abandons_a_long_launch():
    """gp_tuning.abort_s: a launch still running at the limit abandons its
    work wherever its waves are (inside intersections and long lists too),
    and gp_count returns GP_ETIME right after, with the result of the
    launches that completed before it (a prefix of the tasks)."""
    import time
    host = gp.HostGraph.rmat(16, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    # a 6-vertex motif whose full count here takes minutes, as one launch,
    # with single tasks that run longer than the limit: nothing completes
    heavy, _ = gp.build_schedule(gp.motifs(6)[77], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    # the 4-clique on R-MAT scale 22 (about 330 ms, no task longer than about
    # a millisecond) in launches of about 5 ms: those before the limit stand
    big = gp.HostGraph.rmat(22, 8, 0.57, 0.19, 0.19, 2)
    bdev = gp.DeviceGraph(big)
    btri = bdev.triangle_counting()
    four, _ = gp.build_schedule(gp.PATTERNS["4clique"], 1, 1, True, big.v_cnt, big.e_cnt, btri)
    for d, s, slice_ms, limit in ((dev, heavy, -1, 0.3), (bdev, four, 5, 0.15)):
        t = gp.Tuning()
        t.slice_ms = slice_ms
        t.abort_s = limit
        t0 = time.perf_counter()
        with pytest.raises(gp.GPTimeout) as e:
            d.count(s, t)
        wall = time.perf_counter() - t0
        part = e.value.partial
        print(f"slice_ms {slice_ms}: stopped after {wall:.3f} s, {part.stats.launches} launches "
              f"({part.stats.kernel_ms:.1f} ms), {part.stats.tasks} of {part.stats.tasks_total} tasks", flush=True)
        # the waves leave within milliseconds of the limit
        assert wall < limit + 0.25
        if slice_ms < 0:
            assert part.stats.tasks == 0 and part.raw == 0
        else:
            assert 0 < part.stats.tasks < part.stats.tasks_total
            r = gp.Tuning()
            r.task_end = int(part.stats.tasks)
            r.slice_ms = -1
            assert d.count(s, r).raw == part.raw
    # a limit that is not reached changes nothing
    t = gp.Tuning()
    t.abort_s = 600.0
    s4, _ = gp.build_schedule(gp.PATTERNS["4clique"], 1, 1, True, host.v_cnt, host.e_cnt, tri)
    assert dev.count(s4, t).raw == dev.count(s4).raw


def test_unreached_hard_stop_changes_nothing_on_every_path():
    """A hard stop that is never reached runs the kernel instantiation with
    the clock checks (one per extended path: plain, flagged shared search,
    pair level grouped by v1, merge-path builds) and must give the same raw
    sums as the instantiation without them."""
    host = gp.HostGraph.rmat(16, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()

    def sched(name):
        return gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)[0]

    house, pent, four = sched("house"), sched("pentagon"), sched("4clique")
    assert gp.kernel_plan(house)["shared_search"] and gp.kernel_plan(pent)["pair_grouped_v1"]
    cases = [("4clique", four, {}), ("house generic", house, {"no_special": 1}),
             ("pentagon", pent, {}), ("4clique merge path", four, {"merge_ratio": 8})]
    for name, s, fields in cases:
        base = gp.Tuning()
        for k, v in fields.items():
            setattr(base, k, v)
        want = dev.count(s, base).raw
        hard = gp.Tuning()
        for k, v in fields.items():
            setattr(hard, k, v)
        hard.abort_s = 600.0
        hard.task_abort_s = 600.0
        got = dev.count(s, hard).raw
        print(f"{name}: {got} (no hard stop {want})", flush=True)
        assert got == want, name
