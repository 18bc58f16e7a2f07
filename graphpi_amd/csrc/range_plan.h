// Task-range plan of a long count on a kernel that cannot stop itself (the
// House kernel, a root per fetch): which consecutive task ranges gp_count
// launches, and when a time limit stops it.  (The generic kernel stops
// fetching at a deadline instead and needs no plan.)  Pure host code (no HIP), so the
// sizing and deadline rules can be exercised on the CPU with made-up kernel
// times.
//
// Rules: a count of at most 2^20 tasks, or one with slice_ms < 0, is one
// launch.  Otherwise the first launch takes 1/256 of the tasks (1/first_div);
// each later
// range is sized from the previous range's kernel time per task for about
// slice_ms, at most four times the previous range and at least 1/1024 of the
// tasks (a launch's fixed cost must not shrink the ranges to nothing when it
// alone exceeds slice_ms, so a count makes at most about 1,300 launches);
// the rest goes into one launch when it is at most 1.25 ranges.  The growth
// cap keeps a cheap first range from sizing the rest as one launch when the
// expensive tasks lie further on: on a graph relabelled high degree first,
// the first tasks are the hubs' edges, which have almost no candidates below
// them, so the first range's rate says nothing about the rest.  After each
// launch but the last, the wall time since the start is compared with the
// time limit (the reference's Graph::max_running_time, include/graph.h:12,
// checked in its commented-out test at src/graph.cpp:229-237).
#pragma once
#include <algorithm>
#include <cstdint>

struct GpRangePlan {
    uint64_t begin = 0, end = 0; // task range of the whole count
    uint64_t cur = 0;            // first task not yet launched
    uint64_t size = 0;           // tasks of the next launch
    double slice_ms = 2000.0;
    double time_limit_s = 0.0;   // <= 0: none
    double t_start = 0.0;
    uint64_t floor_size = 1;     // smallest range after the first
    int launches = 0;
    bool timed_out = false;

    // first_div: the first launch takes 1/first_div of the tasks (256 by
    // default; a kernel whose unit of work is a whole root passes a smaller
    // divisor, since a short first range there holds a few heavy roots and
    // leaves most waves idle)
    GpRangePlan(uint64_t b, uint64_t e, int32_t slice_ms_req, double time_limit, double now, uint64_t first_div = 256)
        : begin(b), end(std::max(b, e)), cur(b), time_limit_s(time_limit), t_start(now) {
        const uint64_t n = end - begin;
        if (slice_ms_req > 0) slice_ms = (double)slice_ms_req;
        size = (slice_ms_req < 0 || n <= (1ull << 20)) ? n : std::max<uint64_t>(1, n / std::max<uint64_t>(first_div, 1));
        floor_size = std::max<uint64_t>(1, n / 1024);
    }
    bool single() const { return size == end - begin && launches == 0; }
    // the next range to launch; false when every task has been launched or
    // the time limit stopped the count
    bool next(uint64_t* b, uint64_t* e) const {
        if (timed_out || (cur >= end && launches > 0)) return false;
        *b = cur;
        *e = std::min(end, cur + size);
        return true;
    }
    // record the launch of [cur, e) that took kernel_ms; now = wall seconds
    void finished(uint64_t e, double kernel_ms, double now) {
        const uint64_t did = e > cur ? e - cur : 0;
        cur = std::max(cur, e);
        ++launches;
        if (cur >= end) return;
        if (time_limit_s > 0 && now - t_start > time_limit_s) {
            timed_out = true;
            return;
        }
        const double rate = std::max(kernel_ms, 0.01) / (double)std::max<uint64_t>(did, 1);
        const uint64_t left = end - cur;
        const double fit = slice_ms / rate;
        size = std::max<uint64_t>(floor_size, std::min<uint64_t>(4 * std::max<uint64_t>(did, 1),
                                                                 fit >= 1.8e19 ? ~0ull : (uint64_t)fit));
        if (left <= size + size / 4) size = left;
    }
};
