"""Benchmark: GraphPi pattern counting on one MI355X.

Workload (default): R-MAT scale 22, edge factor 8, (a,b,c,d)=(0.57,0.19,0.19,0.05),
seed 2, symmetrised and deduplicated (2,009,103 vertices, 32.6M undirected
edges, int32 CSR); count 4-clique, 5-cycle and House with the planner's
schedule (performance_modeling_type=1, restricts_type=1, IEP on).

One step = counting all three patterns over the resident CSR (one pass of the
hot path).  The CSR is uploaded and the schedules are built before the timed
region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2|cfg1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "cfg2": {
        "workload": "R-MAT scale 22, edge factor 8, (0.57,0.19,0.19,0.05), seed 2; 4-clique, 5-cycle, House; "
                    "pm=1 rt=1 IEP on",
        "gen": ("rmat", 22, 8, 2),
        "patterns": ["4clique", "pentagon", "house"],
    },
    "cfg1": {
        "workload": "R-MAT scale 13, edge factor 12, (0.57,0.19,0.19,0.05), seed 1; Triangle, Rectangle, House; "
                    "pm=1 rt=1 IEP on",
        "gen": ("rmat", 13, 12, 1),
        "patterns": ["triangle", "rectangle", "house"],
    },
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(host, scheds, names, budget_s=20.0):
    """The oracle's restatement of the reference loop nest (OpenMP, dynamic
    chunk-1 over roots) on every host core available, over a sample of roots
    v = r, r+S, r+2S, ...; the full-workload time is estimated as the sample
    time x S."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob
    threads = cpu_threads()
    per = budget_s / max(1, len(names))
    total_est = 0.0
    samples = []
    v = host.v_cnt
    for name in names:
        s = scheds[name]
        stride = max(1, v // 64)
        t_sample = 0.0
        for _ in range(4):
            t0 = time.perf_counter()
            ob.pattern_matching(host, s, threads=threads, root_begin=stride // 2, root_end=-1, root_stride=stride)
            t_sample = time.perf_counter() - t0
            if t_sample >= per / 4 or stride == 1:
                break
            stride = max(1, int(stride * max(t_sample, 1e-3) / per))
        est = t_sample * stride
        total_est += est
        samples.append(f"{name}: roots v = {stride // 2} + {stride}*i ({(v + stride - 1) // stride} roots), "
                       f"{t_sample:.2f} s -> {est:.1f} s")
        log(f"cpu sample {samples[-1]}")
    return {"value": total_est, "unit": "s", "cores": threads, "kind": "port",
            "sample": "; ".join(samples) + "; full time estimated as sample time x stride"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-balg", action="store_true")
    ap.add_argument("--patterns", default=None, help="comma-separated subset (exploration only)")
    ap.add_argument("--scale", type=int, default=None, help="override the R-MAT scale (exploration only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (the counting path has no multi-GPU mode here); "
              f"got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import graphpi_amd as gp
    cfg = dict(CONFIGS[args.config])
    kind, scale, ef, seed = cfg["gen"]
    if args.patterns:
        cfg["patterns"] = args.patterns.split(",")
    if args.scale:
        scale = args.scale
        cfg["workload"] = cfg["workload"].replace("scale %d" % cfg["gen"][1], "scale %d" % scale)
    t0 = time.perf_counter()
    host = gp.HostGraph.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    log(f"graph: {host.v_cnt} vertices, {host.e_cnt // 2} undirected edges, max degree {host.max_degree} "
        f"({time.perf_counter() - t0:.1f} s)")
    dev = gp.DeviceGraph(host)
    tri = dev.triangle_counting()
    log(f"triangles {tri}")
    scheds = {}
    for name in cfg["patterns"]:
        s, valid = gp.build_schedule(gp.PATTERNS[name], 1, 1, True, host.v_cnt, host.e_cnt, tri)
        assert valid
        scheds[name] = s

    def step():
        out = {}
        for name in cfg["patterns"]:
            log(f"  counting {name} ...")
            out[name] = dev.count(scheds[name])
            st = out[name].stats
            log(f"  {name}: {out[name].count} in {st.kernel_ms:.2f} ms (waves done: half {st.finish_p50_ms:.1f} ms, "
                f"90% {st.finish_p90_ms:.1f} ms, all {st.finish_max_ms:.1f} ms)")
        return out

    for _ in range(args.warmup):
        r = step()
        log("warmup " + " ".join(f"{n}={r[n].count} ({r[n].stats.kernel_ms:.1f} ms)" for n in r))
    kernel_ms = {n: 0.0 for n in cfg["patterns"]}
    counts = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        for n in r:
            kernel_ms[n] += r[n].stats.kernel_ms
            counts[n] = r[n].count
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    for n in kernel_ms:
        kernel_ms[n] /= args.steps
    log("timed " + " ".join(f"{n}: {kernel_ms[n]:.2f} ms" for n in kernel_ms))

    per_pattern = {}
    balg = {}
    if not args.no_balg:
        tun = gp.Tuning()
        tun.collect_balg = 1
        for n in cfg["patterns"]:
            rf = dev.count(scheds[n], tun)
            assert rf.count == counts[n], (n, rf.count, counts[n])
            balg[n] = rf.stats.balg_bytes
    for n in cfg["patterns"]:
        per_pattern[n] = {"count": counts[n], "kernel_ms": round(kernel_ms[n], 3)}
        if n in balg:
            per_pattern[n]["balg_bytes"] = balg[n]
            per_pattern[n]["balg_GBps"] = round(balg[n] / (kernel_ms[n] * 1e-3) / 1e9, 1)
    dom = max(kernel_ms, key=kernel_ms.get)
    roofline = None
    if dom in balg:
        achieved = balg[dom] / (kernel_ms[dom] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": f"gp_count_kernel[{dom}]"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(host, scheds, cfg["patterns"])
        cpu["speedup_vs_gpu_count_time"] = round(cpu["value"] / (sum(kernel_ms.values()) * 1e-3), 1)

    value = sum(kernel_ms.values()) * 1e-3
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
                   "triangles": tri, "patterns": cfg["patterns"]},
        "per_pattern": per_pattern,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
