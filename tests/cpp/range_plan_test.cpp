// CPU check of gp_count's task-range plan (graphpi_amd/csrc/range_plan.h)
// with made-up kernel times: ranges are contiguous and cover the count
// exactly, launches stay near the slice target, and a time limit stops the
// plan between launches.  Prints one line per scenario; exits non-zero on a
// failed check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "range_plan.h"

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);            \
            ++fails;                                                    \
        }                                                               \
    } while (0)

// cost(t) = ms per task at task t; the clock advances by each launch's time
template <class Cost>
static GpRangePlan run(uint64_t b, uint64_t e, int slice, double limit, Cost cost, std::vector<double>* ms_out,
                       uint64_t* covered) {
    double now = 100.0;
    GpRangePlan p(b, e, slice, limit, now);
    uint64_t x = 0, y = 0, expect = b;
    *covered = 0;
    while (p.next(&x, &y)) {
        CHECK(x == expect);
        CHECK(y > x || (b == e));
        double ms = 0;
        for (uint64_t t = x; t < y; t += std::max<uint64_t>(1, (y - x) / 1000))
            ms += cost(t) * (double)std::min<uint64_t>(std::max<uint64_t>(1, (y - x) / 1000), y - t);
        ms_out->push_back(ms);
        now += ms / 1000.0;
        *covered += y - x;
        expect = y;
        p.finished(y, ms, now);
    }
    return p;
}

int main() {
    std::vector<double> ms;
    uint64_t cov = 0;
    // small counts: one launch
    {
        GpRangePlan p = run(0, 1000, 0, 0, [](uint64_t) { return 1.0; }, &ms, &cov);
        CHECK(p.launches == 1 && cov == 1000 && !p.timed_out);
        std::printf("small: %d launches\n", p.launches);
    }
    // slice_ms < 0: one launch whatever the size
    ms.clear();
    {
        GpRangePlan p = run(5, 50000000, -1, 0, [](uint64_t) { return 1e-3; }, &ms, &cov);
        CHECK(p.launches == 1 && cov == 50000000 - 5);
    }
    // uniform cost, 65M tasks at 5e-6 ms (325 ms in all): ranges of 1, 4, 16
    // and 64 / 256 of the tasks, then the rest
    ms.clear();
    {
        GpRangePlan p = run(0, 65000000, 0, 0, [](uint64_t) { return 5e-6; }, &ms, &cov);
        CHECK(cov == 65000000 && p.launches == 5);
        std::printf("uniform cheap: %d launches\n", p.launches);
    }
    // a cheap head and an expensive rest (a degree-ordered graph: the hubs'
    // edge slots come first and have few candidates), 234M tasks, 1e-6 ms
    // per task for the first 1/256 and 5e-4 ms after it (117 s in all): no
    // launch is sized from the head's rate alone
    ms.clear();
    {
        const uint64_t n = 234000000;
        GpRangePlan p = run(0, n, 0, 0, [&](uint64_t t) { return t < n / 256 ? 1e-6 : 5e-4; }, &ms, &cov);
        CHECK(cov == n);
        double mx = 0;
        for (double m : ms) mx = std::max(mx, m);
        CHECK(mx < 2.0 * 2000.0 * 1.26);
        std::printf("cheap head: %d launches, longest %.0f ms\n", p.launches, mx);
    }
    // uniform cost, 200 s in all: launches of about the 2 s slice after the first
    ms.clear();
    {
        const double rate = 200000.0 / 50000000.0;
        GpRangePlan p = run(0, 50000000, 0, 0, [&](uint64_t) { return rate; }, &ms, &cov);
        CHECK(cov == 50000000);
        double mx = 0;
        for (size_t i = 1; i < ms.size(); ++i) mx = std::max(mx, ms[i]);
        CHECK(mx <= 2.0 * 2000.0 * 1.26);
        std::printf("uniform heavy: %d launches, longest %.0f ms\n", p.launches, mx);
    }
    // cost rising 100x along the task order: each range is sized at the
    // previous range's rate, so a launch overshoots by at most the rise
    // within it (here under 10x the slice), never the whole count
    ms.clear();
    {
        const uint64_t n = 40000000;
        GpRangePlan p = run(1000, 1000 + n, 1000, 0,
                            [&](uint64_t t) { return 1e-5 * std::pow(100.0, (double)(t - 1000) / n); }, &ms, &cov);
        CHECK(cov == n);
        double mx = 0;
        for (double m : ms) mx = std::max(mx, m);
        CHECK(mx < 10.0 * 1000.0 && p.launches >= 4);
        std::printf("rising: %d launches, longest %.0f ms\n", p.launches, mx);
    }
    // a fixed cost per launch above the slice target: ranges keep a floor of
    // 1/1024 of the tasks instead of shrinking to single tasks
    ms.clear();
    {
        double now = 0;
        GpRangePlan p(0, 2000000, 1, 0, now);
        uint64_t x = 0, y = 0, covered = 0;
        while (p.next(&x, &y)) {
            const double k = 2.0 + 1e-7 * (double)(y - x); // 2 ms whatever the range
            covered += y - x;
            now += k / 1000.0;
            p.finished(y, k, now);
            CHECK(p.launches < 5000);
            if (p.launches >= 5000) break;
        }
        CHECK(covered == 2000000 && p.launches <= 1300);
        std::printf("fixed cost: %d launches\n", p.launches);
    }
    // time limit: stops between launches, with part of the tasks counted
    ms.clear();
    {
        const double rate = 100000.0 / 30000000.0; // 100 s in all
        GpRangePlan p = run(0, 30000000, 0, 5.0, [&](uint64_t) { return rate; }, &ms, &cov);
        CHECK(p.timed_out);
        CHECK(cov < 30000000 && cov == p.cur);
        double total = 0;
        for (double m : ms) total += m;
        CHECK(total / 1000.0 > 5.0 && total / 1000.0 < 5.0 + 2 * 2.0 * 1.26);
        std::printf("deadline: stopped after %d launches, %.1f s, %llu tasks\n", p.launches, total / 1000.0,
                    (unsigned long long)cov);
    }
    // a time limit that is never reached changes nothing
    ms.clear();
    {
        GpRangePlan p = run(0, 30000000, 0, 1e9, [](uint64_t) { return 1e-5; }, &ms, &cov);
        CHECK(!p.timed_out && cov == 30000000);
    }
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
