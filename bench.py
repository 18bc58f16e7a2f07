"""Benchmark: GraphPi pattern counting on one MI355X.

Graph (default): R-MAT scale 22, edge factor 8, (a,b,c,d)=(0.57,0.19,0.19,0.05),
seed 2, symmetrised and deduplicated (2,009,103 vertices, 32.6M undirected
edges, int32 CSR), schedules from the planner (performance_modeling_type=1,
restricts_type=1, IEP on).  This graph's configuration names 4-clique, 5-cycle
and House.

One step = one full count of the timed pattern (the 4-clique) over the
resident CSR, launched as consecutive task ranges (the raw sums add up
exactly); `value` is its kernel time.  House and the 5-cycle are counted once
each outside the step loop, with one progress line per task range, and are
reported in `per_pattern` with their kernel time, host OpenMP baseline and
ratio.  The CSR is uploaded and the schedules are built before any timed
region.  Prints one JSON line.

`roofline` is physical: HBM-side bytes of the dominant kernel, read from the
rocprofv3 FETCH_SIZE / WRITE_SIZE runs committed under profiles/ (scaled by
the copy calibration stated there), over the kernel time measured live here.
B_alg, the reference loop nest's algorithmic bytes, is reported beside it as
a work measure only.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config rmat22|rmat13|chunglu3m]
                    [--patterns a,b] [--once a,b] [--no-cpu] [--no-balg] [--no-batched] [--no-served]
    python bench.py --profile house [--task-range A:B]   # one count, for rocprofv3
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "rmat22": {
        "workload": "R-MAT scale 22, edge factor 8, (0.57,0.19,0.19,0.05), seed 2; pm=1 rt=1 IEP on",
        "gen": ("rmat", 22, 8, 2),
        "patterns": ["4clique"],
        "once": ["house", "pentagon"],
        "named": ["4clique", "pentagon", "house"],
    },
    "chunglu3m": {
        "workload": "Chung-Lu 3,000,000 vertices, Pareto exponent 2.1, at least 117M distinct edges, max expected "
                    "degree 30000, seed 3; pm=1 rt=1 IEP on",
        "gen": ("chunglu", 3000000, 117000000, 3),
        "oriented": 1,  # the loader's high-degree-first relabelling (src/dataloader.cpp:96-110)
        "patterns": ["4clique"],
        "once": ["5clique", "tri_edge_tri"],
        "named": ["4clique", "5clique", "tri_edge_tri"],
    },
    "rmat13": {
        "workload": "R-MAT scale 13, edge factor 12, (0.57,0.19,0.19,0.05), seed 1; pm=1 rt=1 IEP on",
        "gen": ("rmat", 13, 12, 1),
        "patterns": ["triangle", "rectangle", "house"],
        "once": [],
        "named": ["triangle", "rectangle", "house"],
    },
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Threads the CPU baseline uses (every core this process may run on) and
    what the host is."""
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count() or 1
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": max(1, aff), "nproc": os.cpu_count(), "affinity": aff, "model": model}


def cpu_baseline(host, scheds, names, budget_s=15.0, seed=7):
    """The oracle's restatement of the reference loop nest (OpenMP, dynamic
    chunk 1 over roots, src/graph.cpp:199-245) on every core this process may
    use, over roots in a seeded random order for budget_s seconds per pattern.
    A root still running at the deadline is abandoned and counts for nothing.
    The full count's time is estimated as the mean thread time per finished
    root x all roots / threads; abandoned roots are the expensive ones, so the
    estimate is a lower bound on the host time (and the GPU ratio a lower
    bound too) unless every root finished, when it is the measured time."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob
    cpu = host_cpu()
    threads = cpu["threads"]
    out = {}
    order = np.random.default_rng(seed).permutation(host.v_cnt).astype(np.int32)
    for name in names:
        t0 = time.perf_counter()
        r = ob.pattern_matching_timed(host, scheds[name], order, budget_s, threads=threads)
        wall = time.perf_counter() - t0
        done = r["roots_done"]
        if done == host.v_cnt:
            est, kind = wall, "measured"
        elif done == 0:
            # not one root finished: there is nothing to scale
            est, kind = None, f"unmeasured: no root finished within {budget_s:g} s"
        else:
            est = r["busy_s"] / max(1, done) * host.v_cnt / threads
            kind = "lower-bound estimate"
        out[name] = {"value": None if est is None else round(est, 3), "kind": kind, "roots_done": done,
                     "roots": host.v_cnt, "sample_s": round(wall, 2)}
        log(f"cpu {name}: {done} of {host.v_cnt} random roots in {wall:.1f} s on {threads} threads -> "
            f"{'-' if est is None else f'{est:.1f}'} s ({kind})")
    return cpu, out


def batched_bench(gp, scale=12):
    """Every connected 5- and 6-vertex pattern (21 + 112, MotifGenerator order)
    over one resident R-MAT graph (edge factor 16, seed 4): one gp_count_batch
    call against the same schedules counted one gp_count call at a time."""
    host = gp.HostGraph.rmat(scale, 16, 0.57, 0.19, 0.19, 4)
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    scheds = []
    for k in (5, 6):
        for adj in gp.motifs(k):
            sch, valid = gp.build_schedule(adj, 1, 1, True, host.v_cnt, host.e_cnt, tri)
            assert valid
            scheds.append(sch)
    dev.count_batch(scheds[:4])  # warm-up
    t0 = time.perf_counter()
    loop_ms, singles = 0.0, []
    for i, sch in enumerate(scheds):
        r = dev.count(sch)
        loop_ms += r.stats.kernel_ms
        singles.append(r.raw)
        if r.stats.kernel_ms > 5000.0 or (i + 1) % 40 == 0:
            log(f"batched: pattern {i + 1} of {len(scheds)}: {r.stats.kernel_ms:.0f} ms kernel "
                f"({time.perf_counter() - t0:.1f} s so far)")
    loop_wall = time.perf_counter() - t0
    log(f"batched: one at a time {loop_wall:.2f} s ({loop_ms:.0f} ms kernel)")
    t0 = time.perf_counter()
    res, st = dev.count_batch(scheds)
    batch_wall = time.perf_counter() - t0
    exact = all(r.raw == x for r, x in zip(res, singles))
    log(f"batched: one launch {batch_wall:.2f} s ({st.kernel_ms:.0f} ms kernel), identical={exact}")
    assert exact
    # the same batch as hipGraph launches: built on the first call, replayed
    # (kernel arguments updated) on the second
    gt = gp.Tuning()
    gt.use_graph = 1
    dev.count_batch(scheds, gt)
    t0 = time.perf_counter()
    gres, gst = dev.count_batch(scheds, gt)
    graph_wall = time.perf_counter() - t0
    gexact = all(r.raw == x for r, x in zip(gres, singles))
    log(f"batched: graph replay {graph_wall:.2f} s ({gst.kernel_ms:.0f} ms), identical={gexact}")
    assert gexact
    n = len(scheds)
    return {"workload": f"R-MAT scale {scale}, edge factor 16, seed 4; all {n} connected 5- and 6-vertex patterns",
            "v_cnt": host.v_cnt, "undirected_edges": host.e_cnt // 2,
            "patterns_per_s_batched": round(n / batch_wall, 3), "patterns_per_s_one_at_a_time": round(n / loop_wall, 3),
            "batched_s": round(batch_wall, 4), "one_at_a_time_s": round(loop_wall, 4),
            "batched_graph_replay_s": round(graph_wall, 4), "patterns_per_s_graph_replay": round(n / graph_wall, 3),
            "speedup": round(loop_wall / batch_wall, 3), "exact": exact}


def served_bench(gp, scales=(16, 20, 22, 24), rates=(50, 200, 800), n_req=60, seed=12, deadline_s=1.0,
                 quantum_s=0.05, use_graph=0, sizes=(3, 7), slots=1):
    """Counting requests against resident graphs (R-MAT edge factor 16,
    seeds 5.., one per scale), random connected 3-7-vertex patterns (uniform
    size; spanning tree + Bernoulli(0.3) extra edges) on a uniformly chosen
    graph, Poisson arrivals at each rate; every request has a deadline of
    deadline_s after its arrival.

    The loop: CountServer (slots concurrent calls, batches of up to 32 new
    requests, time slices of quantum_s, continuing requests earliest
    deadline first, admission by footprint within 85 % of the free HBM).  Serial:
    the same requests one at a time through plain gp_count with the same
    deadline as its time limit, first come first served at the same
    arrivals.  Each request's service time alone is measured once; the
    serial latencies follow from those (a request whose deadline has passed
    when its turn comes, or would pass during its count, times out).  p50 and
    p99 are over answered requests; timed-out requests are reported by class
    (pattern size, graph), never dropped.  Throughput counts answered
    requests.  Every answered result equals the count made alone."""
    import gc
    import numpy as np
    from graphpi_amd.serve import CountServer, Request, percentile, poisson_arrivals, random_pattern, run_load
    gc.collect()  # device objects left from earlier legs are freed now, not during the load
    t0 = time.perf_counter()
    graphs = {}
    for i, sc in enumerate(scales):
        graphs[f"s{sc}"] = gp.HostGraph.rmat(sc, 16, 0.57, 0.19, 0.19, 5 + i)
        log(f"served: R-MAT s{sc} ef16 seed {5 + i}: {graphs[f's{sc}'].v_cnt} vertices, "
            f"{graphs[f's{sc}'].e_cnt // 2} edges ({time.perf_counter() - t0:.1f} s)")
    names = sorted(graphs, key=lambda x: int(x[1:]))
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_req):
        g = names[int(rng.integers(0, len(names)))]
        specs.append((g, random_pattern(rng, int(rng.integers(sizes[0], sizes[1] + 1)))))
    out = {"workload": "R-MAT scales " + "/".join(str(x) for x in scales) +
                       f", edge factor 16, seeds 5-{4 + len(scales)}; random connected {sizes[0]}-{sizes[1]}-vertex "
                       f"patterns; Poisson arrivals" +
                       (f"; deadline {deadline_s:g} s after arrival" if deadline_s else ""),
           "requests_per_rate": n_req, "deadline_s": deadline_s, "quantum_s": quantum_s, "slots": slots,
           "rates": {}}
    tun = gp.Tuning()
    tun.use_graph = use_graph
    with CountServer(graphs, slots=slots, max_batch=32, tuning=tun, quantum_s=quantum_s,
                     deadline_s=deadline_s) as srv:
        out["resident_graph_bytes"] = sum(d.nbytes() for d in srv._dev.values())
        out["admission_budget_bytes"] = srv._budget
        log(f"served: {len(graphs)} graphs resident ({out['resident_graph_bytes'] / 2**30:.2f} GiB), admission "
            f"budget {srv._budget / 2**30:.1f} GiB ({time.perf_counter() - t0:.1f} s)")
        scheds = [srv.schedule_for(g, adj, (1, 1, 1)) for g, adj in specs]
        # each request alone, once: plain gp_count with the deadline as its time limit
        service, alone = [], []
        for (g, adj), s in zip(specs, scheds):
            t = gp.Tuning()
            if deadline_s:
                t.time_limit_s = deadline_s
                t.abort_s = deadline_s
                t.slice_ms = max(1, int(quantum_s * 1000))
            a = time.perf_counter()
            try:
                r = srv._dev[g].count(s, t)
            except gp.GPTimeout:
                r = None
            service.append(time.perf_counter() - a)
            alone.append(r)
            log(f"served: alone {len(alone)}/{n_req}: {adj.shape[0]}-vertex pattern on {g}: "
                f"{service[-1] * 1e3:.1f} ms{'' if r is not None else ' (deadline passed)'}")
        log(f"served: {sum(r is not None for r in alone)} of {n_req} requests finish alone"
            f"{f' within {deadline_s:g} s' if deadline_s else ''} ({sum(service):.1f} s of service in all)")
        classes = {}
        for (g, adj), r in zip(specs, alone):
            c = classes.setdefault(f"{adj.shape[0]}v {g}", {"requests": 0, "alone_finished": 0})
            c["requests"] += 1
            c["alone_finished"] += r is not None
        for rate in rates:
            rrng = np.random.default_rng(seed + rate)
            reqs = [Request(i, g, adj) for i, (g, adj) in enumerate(specs)]
            offs = poisson_arrivals(rrng, float(rate), n_req)
            wall = run_load(srv, reqs, offs)
            ans = [r for r in reqs if r.error is None]
            exact = all(a is None or (r.count, r.raw) == (a.count, a.raw) for r, a in zip(ans, [alone[r.rid] for r in ans]))
            lat = [r.latency * 1e3 for r in ans]
            # the serial queue at the same arrivals
            clock, slat, s_ok = 0.0, [], 0
            for off, sv, a in zip(offs, service, alone):
                start = max(off, clock)
                budget = off + deadline_s - start if deadline_s else float("inf")
                if budget <= 0:
                    clock = start
                elif a is not None and sv <= budget:
                    clock = start + sv
                    slat.append((clock - off) * 1e3)
                    s_ok += 1
                else:
                    clock = start + min(sv, budget)
            per_class = {}
            for r in reqs:
                c = per_class.setdefault(f"{r.adj.shape[0]}v {r.graph}", {"answered": 0, "timed_out": 0})
                c["answered" if r.error is None else "timed_out"] += 1
            rec = {"answered": len(ans), "timed_out": sum(r.timed_out for r in reqs),
                   "p50_ms": round(percentile(lat, 50), 3) if lat else None,
                   "p99_ms": round(percentile(lat, 99), 3) if lat else None,
                   "throughput_qps": round(len(ans) / wall, 2),
                   "serial_answered": s_ok,
                   "serial_p50_ms": round(percentile(slat, 50), 3) if slat else None,
                   "serial_p99_ms": round(percentile(slat, 99), 3) if slat else None,
                   "serial_throughput_qps": round(s_ok / clock, 2) if clock > 0 else None,
                   "slices_max": max(r.slices for r in reqs), "exact": exact, "classes": per_class}
            log(f"served {rate} q/s: " + json.dumps({k: v for k, v in rec.items() if k != "classes"}))
            assert exact
            out["rates"][str(rate)] = rec
        out["classes_alone"] = classes
    return out


def load_traffic():
    """HBM-side bytes per (config, pattern) from the committed rocprofv3 runs."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="rmat22", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of host OpenMP sampling per pattern for the CPU baseline (default 15)")
    ap.add_argument("--no-balg", action="store_true")
    ap.add_argument("--no-relabelled", action="store_true",
                    help="skip timing the timed patterns on the graph relabelled high degree first")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-pattern batched measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--autotune", action="store_true",
                    help="pick launch settings per pattern on a task sample before the timed region")
    ap.add_argument("--batched-scale", type=int, default=12,
                    help="R-MAT scale of the many-pattern batched measurement (default 12)")
    ap.add_argument("--served-scales", default="16,20,22,24",
                    help="R-MAT scales of the served graphs (default 16,20,22,24)")
    ap.add_argument("--served-requests", type=int, default=60, help="requests per arrival rate (default 60)")
    ap.add_argument("--served-deadline", type=float, default=1.0,
                    help="seconds after arrival by which a served request must be answered (default 1)")
    ap.add_argument("--patterns", default=None, help="comma-separated timed patterns (exploration only)")
    ap.add_argument("--once", default=None,
                    help="comma-separated patterns counted once outside the steps ('' for none)")
    ap.add_argument("--scale", type=int, default=None, help="override the R-MAT scale (exploration only)")
    ap.add_argument("--oriented", type=int, default=None, choices=(0, 1, 2),
                    help="relabel the generated graph by degree as the loader's oriented_type does "
                         "(1 = high degree first; default: the configuration's choice)")
    ap.add_argument("--sample-frac", type=float, default=0.0, metavar="F",
                    help="exploration only: count 64 task ranges spread over the task space, covering a fraction F "
                         "of the tasks, and print the full count's estimated kernel time (no count is reported)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (reference point for HBM byte counters)")
    ap.add_argument("--profile", default=None, metavar="PATTERN",
                    help="only count PATTERN once as one launch (for rocprofv3 runs)")
    ap.add_argument("--task-range", default=None, metavar="A:B",
                    help="with --profile: count only tasks [A, B)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 4, "bytes_written": n * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (the counting path has no multi-GPU mode here); "
              f"got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import graphpi_amd as gp
    cfg = dict(CONFIGS[args.config])
    kind, scale, ef, seed = cfg["gen"]
    if args.patterns:
        cfg["patterns"] = args.patterns.split(",")
    if args.once is not None:
        cfg["once"] = [x for x in args.once.split(",") if x]
    if args.scale:
        scale = args.scale
        cfg["workload"] = cfg["workload"].replace("scale %d" % cfg["gen"][1], "scale %d" % scale)
    t0 = time.perf_counter()
    if kind == "chunglu":
        host = gp.HostGraph.chung_lu(scale, ef, 2.1, 30000.0, seed)
    else:
        host = gp.HostGraph.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    oriented = cfg.get("oriented", 0) if args.oriented is None else args.oriented
    if oriented:
        host = host.reorder(oriented)
        cfg["workload"] += f"; relabelled {'high' if oriented == 1 else 'low'} degree first (oriented_type {oriented})"
    log(f"graph: {host.v_cnt} vertices, {host.e_cnt // 2} undirected edges, max degree {host.max_degree} "
        f"({time.perf_counter() - t0:.1f} s)")
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    log(f"triangles {tri}")
    all_names = list(dict.fromkeys(cfg["patterns"] + cfg["once"] + ([args.profile] if args.profile else [])))
    scheds = {}
    for name in all_names:
        s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        assert valid
        scheds[name] = s
    totals = {name: dev.count_tasks(scheds[name]) for name in all_names}

    if args.profile:
        # one launch (no internal slicing) over the given range: the
        # dispatches a profiler attributes to the pattern
        tun = gp.Tuning()
        tun.slice_ms = -1
        if args.task_range:
            a, b = args.task_range.split(":")
            tun.task_begin, tun.task_end = int(a), int(b)
        r = dev.count(scheds[args.profile], tun)
        print(json.dumps({"mode": "profile", "pattern": args.profile, "workload": cfg["workload"],
                          "task_begin": tun.task_begin, "task_end": tun.task_end or totals[args.profile],
                          "raw": r.raw, "kernel_ms": round(r.stats.kernel_ms, 3),
                          "kernel_variant": r.stats.kernel_variant, "plan": gp.kernel_plan(scheds[args.profile])}),
              flush=True)
        return

    if args.sample_frac > 0:
        est = {}
        for name in cfg["patterns"] + [n for n in cfg["once"] if n not in cfg["patterns"]]:
            total, parts = totals[name], 64
            width = max(1, int(total * args.sample_frac / parts))
            tun, ms, sampled = gp.Tuning(), 0.0, 0
            for i in range(parts):
                tun.task_begin = total * i // parts
                tun.task_end = min(total, tun.task_begin + width)
                r = dev.count(scheds[name], tun)
                ms += r.stats.kernel_ms
                sampled += tun.task_end - tun.task_begin
                log(f"  {name} sample {i + 1}/{parts}: tasks {tun.task_begin}..{tun.task_end}: "
                    f"{r.stats.kernel_ms:.1f} ms")
            est[name] = {"sampled_tasks": sampled, "total_tasks": total, "sample_kernel_ms": round(ms, 3),
                         "estimated_full_s": round(ms * total / sampled / 1e3, 1)}
            log(f"  {name}: estimated {est[name]['estimated_full_s']} s for the full count")
        print(json.dumps({"mode": "estimate (sampled task ranges; not a count)", "workload": cfg["workload"],
                          "v_cnt": host.v_cnt, "undirected_edges": host.e_cnt // 2, "estimates": est}), flush=True)
        return
    tunings, tune_reports = {}, {}
    if args.autotune:
        from graphpi_amd.autotune import autotune
        for name in cfg["patterns"]:
            tunings[name], tune_reports[name] = autotune(dev, scheds[name], log=log)

    def count_pattern(name, target_ms=5000.0, tuning=None, quiet=False):
        """One full count of `name` as consecutive launches, with a progress
        line per launch.  The generic kernel runs with a deadline of
        target_ms per launch: each gp_count call makes one launch, which
        stops fetching at the deadline and counts a prefix of the remaining
        tasks (the call's time limit then returns it), and the next call
        resumes there.  House's two-hop kernel takes a root per fetch and no
        deadline: a first range of 1/4 of its tasks, then ranges at most four
        times the previous one and sized for about target_ms at the previous
        range's rate.  The raw sums add up to the whole count."""
        s = scheds[name]
        total = totals[name]
        tun = gp.Tuning()
        if tuning is not None:
            ctypes.memmove(ctypes.byref(tun), ctypes.byref(tuning), ctypes.sizeof(gp.Tuning))
        shape = 0 if (tun.collect_balg or tun.no_special) else gp.kernel_plan(s)["shape"]
        t0 = time.perf_counter()
        if shape == 3:
            # counted in closed form (triangles, edge supports, 4-cliques):
            # one call for the whole graph, a task range would take the loop nest
            r = dev.count(s, tun)
            if not quiet:
                log(f"    {name}: closed form, {r.stats.kernel_ms:.1f} ms in {r.stats.launches} launches")
            return {"count": r.count, "raw": r.raw, "kernel_ms": r.stats.kernel_ms, "launches": r.stats.launches,
                    "wall_s": time.perf_counter() - t0, "balg": 0, "exact": r.stats.exact_count()}
        begin, raw, ms, launches, balg_sum = 0, 0, 0.0, 0, 0
        size = total if total <= (1 << 20) else max(1, total // 4)
        while begin < total:
            tun.task_begin = begin
            if shape:
                tun.task_end = min(total, begin + size)
                tun.slice_ms = -1
                r = dev.count(s, tun)
                done = tun.task_end
            else:
                tun.task_end = total
                tun.slice_ms = int(target_ms)
                tun.time_limit_s = 1e-9  # one launch per call
                try:
                    r = dev.count(s, tun)
                except gp.GPTimeout as e:
                    r = e.partial
                done = begin + r.stats.tasks
            raw = (raw + r.raw) % (1 << 64)
            ms += r.stats.kernel_ms
            balg_sum += r.stats.balg_bytes
            launches += r.stats.launches
            k = r.stats.kernel_ms
            if not quiet:
                log(f"    {name}: tasks {begin}..{done} of {total}: {k:.1f} ms (tail: half {r.stats.finish_p50_ms:.1f}, "
                    f"all {r.stats.finish_max_ms:.1f} ms; {r.stats.split_tasks} tasks split)")
            if shape:
                rate = max(k, 0.01) / max(1, done - begin)   # ms per task in the last range
                size = max(1, min(4 * (done - begin), int(target_ms / rate)))
                if total - done <= size + size // 4:
                    size = total - done
            begin = done
        signed = raw - (1 << 64) if raw >= (1 << 63) else raw
        red = s.redundancy or 1
        return {"count": signed // red if signed >= 0 else -((-signed) // red), "raw": raw, "kernel_ms": ms,
                "launches": launches, "wall_s": time.perf_counter() - t0, "balg": balg_sum}

    def step():
        out = {}
        for name in cfg["patterns"]:
            log(f"  counting {name} ...")
            out[name] = count_pattern(name, tuning=tunings.get(name))
            log(f"  {name}: {out[name]['count']} in {out[name]['kernel_ms']:.1f} ms kernel time "
                f"({out[name]['launches']} launches)")
        return out

    for _ in range(args.warmup):
        step()
    kernel_ms = {n: 0.0 for n in cfg["patterns"]}
    counts = {}
    launches = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        for n in r:
            kernel_ms[n] += r[n]["kernel_ms"]
            counts[n] = r[n]["count"]
            launches[n] = r[n]["launches"]
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    for n in kernel_ms:
        kernel_ms[n] /= args.steps
    log("timed " + " ".join(f"{n}: {kernel_ms[n]:.2f} ms" for n in kernel_ms))

    # the other named patterns: one count each, outside the steps
    once = {}
    for n in cfg["once"]:
        if n in kernel_ms:
            continue
        log(f"  counting {n} once ...")
        # longer launches than the timed patterns': each launch ends with a
        # tail (the fetches in flight at its deadline), so fewer is cheaper
        once[n] = count_pattern(n, target_ms=15000.0)
        log(f"  {n}: {once[n]['count']} in {once[n]['kernel_ms']:.1f} ms kernel time ({once[n]['launches']} launches)"
            + (f"; without wrap-around {once[n]['exact']}" if once[n].get("exact") is not None else ""))
        kernel_ms[n] = once[n]["kernel_ms"]
        counts[n] = once[n]["count"]
        launches[n] = once[n]["launches"]

    # the timed patterns again on the same graph relabelled high degree first
    # (the loader's oriented_type 1, src/dataloader.cpp:96-110), outside the
    # steps.  A pattern count is a graph invariant, so the count must not
    # change; only the work does (shorter lists below each vertex).
    relabelled = None
    if not oriented and not args.no_relabelled:
        t0r = time.perf_counter()
        rhost = host.reorder(1)
        rdev = gp.DeviceGraph(rhost)
        rtri = rdev.triangle_counting()
        relabelled = {"workload": "the same graph relabelled high degree first (oriented_type 1); timed patterns: "
                                  "best of 3 counts, the others once", "per_pattern": {}}
        for n in cfg["patterns"] + [x for x in cfg["once"] if x not in cfg["patterns"]]:
            rs, valid = gp.build_schedule(gp.PATTERNS[n], 1, 1, True, rhost.v_cnt, rhost.e_cnt, rtri)
            assert valid
            rt = gp.Tuning()
            rt.slice_ms = 2000 if n in cfg["patterns"] else 15000
            runs = [rdev.count(rs, rt) for _ in range(3 if n in cfg["patterns"] else 1)]
            assert all(r.count == counts[n] for r in runs), (n, [r.count for r in runs], counts[n])
            best = min(r.stats.kernel_ms for r in runs)
            relabelled["per_pattern"][n] = {"count": runs[0].count, "kernel_ms": round(best, 3),
                                            "launches": runs[0].stats.launches,
                                            "speedup_vs_named_order": round(kernel_ms[n] / best, 3)}
            log(f"  {n} relabelled high degree first: {runs[0].count} in {best:.1f} ms kernel time "
                f"({time.perf_counter() - t0r:.1f} s with the relabelling and upload)")
        del rdev

    # B_alg: the reference-faithful kernel (no truncation, no batching) tallies
    # 4 B x (|A|+|B|) per intersection + 8 B per neighbour-list lookup.  It
    # first runs on 16 task ranges spread over the task space (1/256 of the
    # tasks); when the whole faithful count is then predicted to take under
    # 30 s it runs in full (exact), otherwise the sample is scaled by the task
    # fraction (estimated).
    balg, balg_note = {}, {}
    if not args.no_balg:
        tun = gp.Tuning()
        tun.collect_balg = 1
        for n in cfg["patterns"]:
            total = totals[n]
            sz, k = scheds[n].size, scheds[n].iep_num
            root_mode = (sz - k if k > 1 else sz - 1) == 1   # tasks are roots, not edge slots
            parts, sampled, b, t_ms = 16, 0, 0, 0.0
            width = max(1, total // (parts * 16))
            for i in range(parts):
                tun.task_begin = total * i // parts
                tun.task_end = min(total, tun.task_begin + width)
                rf = dev.count(scheds[n], tun)
                b += rf.stats.balg_bytes - (8 * host.v_cnt if i == 0 and not root_mode else 0)
                sampled += tun.task_end - tun.task_begin
                t_ms += rf.stats.kernel_ms
            log(f"  B_alg sample {n}: {b} bytes over {sampled} of {total} tasks in {t_ms:.0f} ms")
            if t_ms * total / max(1, sampled) < 30000.0:
                rf = count_pattern(n, tuning=tun, quiet=True)
                assert rf["count"] == counts[n], (n, rf["count"], counts[n])
                balg[n] = rf["balg"]
                balg_note[n] = "exact"
                log(f"  B_alg {n}: {balg[n]} bytes (faithful kernel {rf['kernel_ms']:.0f} ms)")
            else:
                balg[n] = int(b * total / sampled) + (0 if root_mode else 8 * host.v_cnt)
                balg_note[n] = f"estimated from {parts} task ranges covering {sampled} of {total} tasks"

    # physical roofline per pattern: profiled HBM-side bytes over the kernel
    # time measured here for the same task range
    traffic = load_traffic()
    per_pattern = {}
    for n in kernel_ms:
        rec = {"count": counts[n], "kernel_ms": round(kernel_ms[n], 3), "launches": launches[n],
               "timed": n in cfg["patterns"], "kernel_shape": gp.kernel_plan(scheds[n])["shape"]}
        if n in once and once[n].get("exact") is not None and once[n]["exact"] != counts[n]:
            # past 2^63 the reference's long long sum wraps; `count` is what it returns
            rec["count_without_wraparound"] = str(once[n]["exact"])
        if n in tune_reports:
            rec["autotune"] = tune_reports[n]
        if n in balg:
            rec["balg_bytes"] = balg[n]
            rec["balg_GBps"] = round(balg[n] / (kernel_ms[n] * 1e-3) / 1e9, 1)
            rec["balg"] = balg_note[n]
        t = traffic.get(f"{args.config}:{scale}:{n}")
        if t:
            tb, te = t.get("task_begin", 0), t.get("task_end", 0) or totals[n]
            if tb == 0 and te >= totals[n]:
                ms = kernel_ms[n]
            else:  # the profiled slice, timed here as one launch
                tun = gp.Tuning()
                tun.slice_ms = -1
                tun.task_begin, tun.task_end = tb, te
                ms = dev.count(scheds[n], tun).stats.kernel_ms
            gbps = t["hbm_bytes"] / (ms * 1e-3) / 1e9
            rec["hbm"] = {"bytes": t["hbm_bytes"], "tasks": [tb, te], "kernel_ms": round(ms, 3),
                          "GBps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4),
                          "l2_hit_pct": t.get("l2_hit_pct"), "occupancy_pct": t.get("occupancy_pct"),
                          "valu_util_pct": t.get("valu_util_pct"), "valu_busy_pct": t.get("valu_busy_pct"),
                          "source": t["source"]}
        per_pattern[n] = rec
    dom = max(cfg["patterns"], key=lambda n: kernel_ms[n])
    roofline = None
    h = per_pattern[dom].get("hbm")
    if h:
        assert h["frac"] <= 1.0, ("measured HBM traffic above the peak: stale profile?", h)
        roofline = {"bound": "hbm", "achieved": h["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": h["frac"],
                    "traffic": h["bytes"], "kernel": f"gp_count_kernel[{dom}]",
                    "achieved_basis": "rocprofv3 FETCH_SIZE + WRITE_SIZE bytes (copy-calibrated) / live kernel time",
                    "l2_hit_pct": h["l2_hit_pct"], "occupancy_pct": h["occupancy_pct"],
                    "valu_util_pct": h["valu_util_pct"], "source": h["source"]}
        if dom in balg:
            roofline["balg_GBps_work_measure"] = per_pattern[dom]["balg_GBps"]

    batched = None if args.no_batched else batched_bench(gp, scale=args.batched_scale)
    served = served_light = None
    if not args.no_served:
        served = served_bench(gp, scales=tuple(int(x) for x in args.served_scales.split(",")),
                              n_req=args.served_requests, deadline_s=args.served_deadline)
        # light requests on small graphs, no deadline: where batching new
        # requests into shared launches pays (every request finishes)
        served_light = served_bench(gp, scales=(8, 9, 10, 11), n_req=120, deadline_s=None, quantum_s=0.05,
                                    sizes=(3, 6), slots=2)

    cpu = None
    if not args.no_cpu:
        info, per = cpu_baseline(host, scheds, list(kernel_ms), budget_s=args.cpu_budget)
        for n, c in per.items():
            per_pattern[n]["cpu_s"] = c["value"]
            per_pattern[n]["cpu_kind"] = c["kind"]
            per_pattern[n]["gpu_speedup"] = (None if c["value"] is None
                                             else round(c["value"] / (kernel_ms[n] * 1e-3), 1))
        tot = sum(per[n]["value"] or 0.0 for n in cfg["patterns"])
        cpu = {"value": round(tot, 3), "unit": "s", "cores": info["threads"], "kind": "port",
               "host": info,
               "sample": "; ".join(f"{n}: {per[n]['roots_done']} of {per[n]['roots']} random roots finished in "
                                   f"{per[n]['sample_s']} s -> {per[n]['value']} s ({per[n]['kind']})"
                                   for n in cfg["patterns"]),
               "speedup_vs_gpu_count_time": round(tot / (sum(kernel_ms[n] for n in cfg["patterns"]) * 1e-3), 1)}

    value = sum(kernel_ms[n] for n in cfg["patterns"]) * 1e-3
    out = {
        "metric": "pattern count time (s) + HBM GB/s on 1 MI355X vs host OpenMP; batched p99 ms",
        "value": round(value, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "v_cnt": host.v_cnt, "undirected_edges": host.e_cnt // 2,
                   "triangles": tri, "patterns_timed": cfg["patterns"], "patterns_once": cfg["once"],
                   "patterns_named": cfg["named"]},
        "per_pattern": per_pattern,
        "relabelled": relabelled,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "served_light": served_light,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
