"""Launch autotuner for one (graph, schedule) pair.

The counting kernel has a few knobs that trade against each other per graph
and pattern: persistent waves per CU, tasks per work fetch (root/edge chunk),
the depth-2 candidate count above which a task is deferred and split into
parts, the LDS pool for staged sets (which decides whether a set is
searched in LDS or by binary search in HBM), the bias between probing a
staged set into a long list and streaming the list, and the size threshold
for staging a set as an LDS bitmap instead of a sorted table.  autotune() measures candidate
settings on a fixed sample of the task space (evenly spread task ranges, so
hub-heavy and light regions are both represented) and keeps the fastest,
one knob at a time (coordinate descent from the defaults).  The result must
then also beat the defaults on a held-out sample (the ranges shifted by half
their spacing), or the defaults are kept.  Every candidate must return the
same raw sum on the sample; a mismatch raises, since settings never change a
count.  Results can be cached per graph and schedule.
"""
import ctypes
import hashlib
import json
import os

from . import Tuning

KNOBS = {
    "waves_per_cu": [16, 24, 32],
    "chunk": [0, 1, 4, 16, 64, 256, 512],
    "split_min": [0, 128, 512, 4096, -1],
    "lds_stage": [0, 128, 256],
    "probe_pct": [0, 25, 50, 200, 400],
    "bitmap_ratio": [0, -1, 4, 64],
}


def _copy(t):
    c = Tuning()
    ctypes.memmove(ctypes.byref(c), ctypes.byref(t), ctypes.sizeof(Tuning))
    return c


def _sample_ranges(total, parts, frac, phase=0.0):
    """`parts` ranges covering `frac` of the tasks, spread evenly; `phase`
    (0..1) shifts every range by that share of the spacing, so a second
    sample with another phase holds other tasks."""
    width = max(1, int(total * frac / parts))
    out = []
    for i in range(parts):
        b = total * i // parts + int(total * phase / parts)
        e = min(total, b + width)
        if e > b:
            out.append((b, e))
    return out


def _key(dev, sched):
    h = hashlib.sha1()
    h.update(str((dev.v_cnt, dev.e_cnt)).encode())
    h.update(bytes(sched))
    return h.hexdigest()


def measure(dev, sched, tuning, ranges, repeats=1):
    """Kernel milliseconds of the sample under `tuning`, and its raw sum."""
    best, raw = None, 0
    for _ in range(repeats):
        ms, r = 0.0, 0
        for b, e in ranges:
            t = _copy(tuning)
            t.task_begin, t.task_end = b, e
            res = dev.count(sched, t)
            ms += res.stats.kernel_ms
            r = (r + res.raw) % (1 << 64)
        best = ms if best is None else min(best, ms)
        raw = r
    return best, raw


def autotune(dev, sched, frac=1.0 / 32, parts=8, cache_path=None, log=None):
    """Returns (Tuning, report).  frac: share of the tasks measured per
    candidate, spread over `parts` ranges."""
    key = _key(dev, sched)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    if key in cache:
        t = Tuning()
        for k, v in cache[key]["knobs"].items():
            setattr(t, k, v)
        return t, dict(cache[key], cached=True)
    total = dev.count_tasks(sched)
    ranges = _sample_ranges(total, parts, frac) if total > 4096 else [(0, total)]
    cur = Tuning()
    base_ms, base_raw = measure(dev, sched, cur, ranges, repeats=2)
    best_ms = base_ms
    tried = 1
    for knob, values in KNOBS.items():
        for v in values:
            if getattr(cur, knob) == v:
                continue
            cand = _copy(cur)
            setattr(cand, knob, v)
            ms, raw = measure(dev, sched, cand, ranges, repeats=2)
            tried += 1
            if raw != base_raw:
                raise RuntimeError(f"tuning {knob}={v} changed the raw sum ({raw} != {base_raw})")
            if ms < best_ms * 0.97:  # keep only clear wins
                best_ms, cur = ms, cand
        if log:
            log(f"autotune {knob}: {getattr(cur, knob)} ({best_ms:.2f} ms on the sample)")
    # a held-out sample (other tasks) must confirm the win, or the defaults stay
    held = None
    if total > 4096 and best_ms < base_ms:
        hold = _sample_ranges(total, parts, frac, phase=0.5)
        d_ms, d_raw = measure(dev, sched, Tuning(), hold, repeats=2)
        t_ms, t_raw = measure(dev, sched, cur, hold, repeats=2)
        if t_raw != d_raw:
            raise RuntimeError(f"tuned settings changed the raw sum on the held-out sample ({t_raw} != {d_raw})")
        held = {"default_ms": round(d_ms, 3), "tuned_ms": round(t_ms, 3)}
        if not t_ms < d_ms * 0.97:
            cur = Tuning()
            best_ms = base_ms
        if log:
            log(f"autotune held-out sample: default {d_ms:.2f} ms, tuned {t_ms:.2f} ms -> "
                f"{'tuned' if best_ms < base_ms else 'defaults'}")
    knobs = {k: getattr(cur, k) for k in KNOBS}
    report = {"knobs": knobs, "sample_ms_default": round(base_ms, 3), "sample_ms_tuned": round(best_ms, 3),
              "candidates": tried, "sample_tasks": sum(e - b for b, e in ranges), "tasks": total,
              "held_out": held}
    if cache_path:
        cache[key] = report
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return cur, dict(report, cached=False)
