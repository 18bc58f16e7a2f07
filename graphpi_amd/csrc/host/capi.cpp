// Host half of the C ABI: error reporting, planner and motif entry points.
#include <cstdio>
#include <cstring>
#include <string>

#include "../gp_internal.h"
#include "motif_generator.h"
#include "pattern.h"
#include "schedule.h"

namespace {
thread_local std::string g_last_error;
}

int gp_fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

extern "C" const char* gp_last_error(void) { return g_last_error.c_str(); }

extern "C" int gp_schedule_build(int32_t size, const int32_t* adj, int32_t performance_modeling_type,
                                 int32_t restricts_type, int32_t use_in_exclusion_optimize,
                                 int32_t v_cnt, uint32_t e_cnt, int64_t tri_cnt, gp_sched_flat* out,
                                 int32_t* valid) {
    if (!adj || !out) return gp_fail(GP_EINVAL, "null argument");
    if (size < 2 || size > GP_MAX_SIZE) return gp_fail(GP_ERANGE, "pattern size must be 2..8");
    if (performance_modeling_type < 0 || performance_modeling_type > 3 || restricts_type < 0 ||
        restricts_type > 2)
        return gp_fail(GP_EINVAL, "unknown performance_modeling_type or restricts_type");
    if (performance_modeling_type != 0 && tri_cnt == -1)
        return gp_fail(GP_EINVAL, "performance modeling needs the triangle count (tri_cnt == -1)");
    if (performance_modeling_type != 0 && (v_cnt <= 0 || e_cnt == 0))
        return gp_fail(GP_EINVAL, "performance modeling needs v_cnt > 0 and e_cnt > 0");
    Pattern p(size);
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            if (adj[i * size + j]) {
                if (i == j) return gp_fail(GP_EINVAL, "pattern has a self-loop");
                p.add_edge(i, j);
            }
    // The Schedule constructor (like the reference's, src/schedule.cpp:268-284)
    // aborts when restriction generation meets a vertex with no earlier
    // neighbour in the given order; report such input as invalid instead.
    if (performance_modeling_type == 0) {
        for (int i = 1; i < size; ++i) {
            bool linked = false;
            for (int j = 0; j < i && !linked; ++j) linked = adj[i * size + j] != 0;
            if (!linked) {
                if (valid) *valid = 0;
                std::memset(out, 0, sizeof(*out));
                out->size = size;
                return GP_OK;
            }
        }
    }
    bool ok = false;
    Schedule s(p, ok, performance_modeling_type, restricts_type, use_in_exclusion_optimize != 0, v_cnt,
               e_cnt, tri_cnt);
    if (valid) *valid = ok ? 1 : 0;
    if (!ok) {
        std::memset(out, 0, sizeof(*out));
        out->size = size;
        return GP_OK;
    }
    int rc = s.flatten(out);
    if (rc != GP_OK) return gp_fail(rc, "schedule exceeds the flattened-table limits");
    return GP_OK;
}

extern "C" int gp_motifs(int32_t n, int32_t* adj_out, int32_t capacity, int32_t* count) {
    if (!count) return gp_fail(GP_EINVAL, "null argument");
    if (n < 1 || n > 7) return gp_fail(GP_ERANGE, "motif size must be 1..7");
    MotifGenerator mg(n);
    std::vector<Pattern> v = mg.generate();
    *count = (int32_t)v.size();
    if (!adj_out) return GP_OK;
    if (capacity < (int32_t)v.size()) return gp_fail(GP_ERANGE, "capacity too small");
    for (size_t i = 0; i < v.size(); ++i)
        for (int k = 0; k < n * n; ++k) adj_out[i * n * n + k] = v[i].get_adj_mat_ptr()[k];
    return GP_OK;
}
