// Test-infrastructure only: a driver that links the UNMODIFIED reference
// GraphPi sources (compiled from where they lie by oracle/Makefile) and prints
// golden values as one JSON line prefixed by "JSON ".  Nothing in the product
// links or runs this; only tests/golden/make_goldens.py and the CPU-baseline
// calibration do, in the build container.
//
// Modes
//   count <graph> <size> <adj> <pm> <rt> <iep> <threads>
//       tianhe/baseline_test.cpp:13-35 with tri_cnt computed by
//       Graph::triangle_counting_mt (src/graph.cpp:112) before the Schedule.
//   sched <size> <adj> <pm> <rt> <iep> <v_cnt> <e_cnt> <tri_cnt>
//       Schedule constructor only (src/schedule.cpp:9-253), dumped through
//       its public accessors (include/schedule.h:23-37,56-58).
//   motif <n>        MotifGenerator(n).generate() (src/motif_generator.cpp:39)
//   tc <graph> <threads>   triangle_counting_mt and triangle_counting
#include "graph.h"
#include "dataloader.h"
#include "pattern.h"
#include "schedule.h"
#include "common.h"
#include "motif_generator.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static std::string adj_string(const int* adj, int n) {
    std::string s;
    for (int i = 0; i < n * n; ++i) s.push_back(adj[i] ? '1' : '0');
    return s;
}

static std::string ivec(const std::vector<int>& v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) s += ",";
        s += std::to_string(v[i]);
    }
    return s + "]";
}

static std::string sched_json(const Schedule& s) {
    int n = s.get_size();
    std::string out = "{";
    out += "\"size\":" + std::to_string(n);
    out += ",\"adj\":\"" + adj_string(s.get_adj_mat_ptr(), n) + "\"";
    out += ",\"restrict_pair\":[";
    for (size_t i = 0; i < s.restrict_pair.size(); ++i) {
        if (i) out += ",";
        out += "[" + std::to_string(s.restrict_pair[i].first) + "," +
               std::to_string(s.restrict_pair[i].second) + "]";
    }
    out += "]";
    out += ",\"iep_num\":" + std::to_string(s.get_in_exclusion_optimize_num());
    out += ",\"iep_groups\":[";
    for (size_t g = 0; g < s.in_exclusion_optimize_group.size(); ++g) {
        if (g) out += ",";
        out += "[";
        for (size_t b = 0; b < s.in_exclusion_optimize_group[g].size(); ++b) {
            if (b) out += ",";
            out += ivec(s.in_exclusion_optimize_group[g][b]);
        }
        out += "]";
    }
    out += "]";
    out += ",\"iep_vals\":" + ivec(s.in_exclusion_optimize_val);
    out += ",\"redundancy\":" + std::to_string(s.get_in_exclusion_optimize_redundancy());
    int P = s.get_total_prefix_num();
    out += ",\"total_prefix_num\":" + std::to_string(P);
    std::vector<int> father, next, last, loop, rlast;
    for (int p = 0; p < P; ++p) {
        father.push_back(s.get_father_prefix_id(p));
        next.push_back(s.get_next(p));
    }
    for (int d = 0; d < n; ++d) {
        last.push_back(s.get_last(d));
        loop.push_back(s.get_loop_set_prefix_id(d));
        rlast.push_back(s.get_restrict_last(d));
    }
    out += ",\"father\":" + ivec(father) + ",\"next\":" + ivec(next) +
           ",\"last\":" + ivec(last) + ",\"loop_set\":" + ivec(loop);
    int R = s.get_total_restrict_num();
    std::vector<int> rnext, ridx;
    for (int r = 0; r < R; ++r) {
        rnext.push_back(s.get_restrict_next(r));
        ridx.push_back(s.get_restrict_index(r));
    }
    out += ",\"total_restrict_num\":" + std::to_string(R) + ",\"restrict_last\":" + ivec(rlast) +
           ",\"restrict_next\":" + ivec(rnext) + ",\"restrict_index\":" + ivec(ridx);
    out += ",\"multiplicity\":" + std::to_string(s.get_multiplicity());
    out += "}";
    return out;
}

static unsigned long long fnv1a(const void* p, size_t n, unsigned long long h) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

static Graph* load(const char* path) {
    Graph* g = nullptr;
    DataLoader D;
    if (!D.load_data(g, DataType::complete8, path)) {
        fprintf(stderr, "load failed: %s\n", path);
        exit(2);
    }
    return g;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: see header\n");
        return 1;
    }
    std::string mode = argv[1];
    if (mode == "count" && argc == 9) {
        Graph* g = load(argv[2]);
        int size = atoi(argv[3]);
        int pm = atoi(argv[5]), rt = atoi(argv[6]), iep = atoi(argv[7]);
        int threads = atoi(argv[8]);
        double t0 = get_wall_time();
        g->tri_cnt = g->triangle_counting_mt(threads);
        double t1 = get_wall_time();
        Pattern p(size, argv[4]);
        bool valid = false;
        Schedule s(p, valid, pm, rt, iep != 0, g->v_cnt, g->e_cnt, g->tri_cnt);
        double t2 = get_wall_time();
        long long ans = valid ? g->pattern_matching(s, threads) : -1;
        double t3 = get_wall_time();
        unsigned long long h = 1469598103934665603ULL;
        h = fnv1a(g->vertex, sizeof(unsigned) * (size_t)(g->v_cnt + 1), h);
        h = fnv1a(g->edge, sizeof(int) * (size_t)g->e_cnt, h);
        printf("JSON {\"count\":%lld,\"valid\":%d,\"v_cnt\":%d,\"e_cnt\":%u,\"tri_cnt\":%lld,"
               "\"max_degree\":%d,\"csr_fnv1a\":\"%016llx\",\"time_tc\":%.6f,\"time_sched\":%.6f,"
               "\"time_count\":%.6f,\"threads\":%d,\"schedule\":%s}\n",
               ans, (int)valid, g->v_cnt, g->e_cnt, g->tri_cnt, g->max_degree, h, t1 - t0,
               t2 - t1, t3 - t2, threads, valid ? sched_json(s).c_str() : "null");
        delete g;
        return 0;
    }
    if (mode == "sched" && argc == 10) {
        int size = atoi(argv[2]);
        Pattern p(size, argv[3]);
        bool valid = false;
        Schedule s(p, valid, atoi(argv[4]), atoi(argv[5]), atoi(argv[6]) != 0, atoi(argv[7]),
                   (unsigned)strtoul(argv[8], nullptr, 10), atoll(argv[9]));
        printf("JSON {\"valid\":%d,\"schedule\":%s}\n", (int)valid,
               valid ? sched_json(s).c_str() : "null");
        return 0;
    }
    if (mode == "motif" && argc == 3) {
        int n = atoi(argv[2]);
        MotifGenerator mg(n);
        std::vector<Pattern> v = mg.generate();
        std::string out = "[";
        for (size_t i = 0; i < v.size(); ++i) {
            if (i) out += ",";
            out += "\"" + adj_string(v[i].get_adj_mat_ptr(), n) + "\"";
        }
        printf("JSON {\"size\":%d,\"patterns\":%s}\n", n, (out + "]").c_str());
        return 0;
    }
    if (mode == "tc" && argc == 4) {
        Graph* g = load(argv[2]);
        int threads = atoi(argv[3]);
        double t0 = get_wall_time();
        long long mt = g->triangle_counting_mt(threads);
        double t1 = get_wall_time();
        long long st = g->triangle_counting();
        printf("JSON {\"tc_mt\":%lld,\"tc\":%lld,\"time_mt\":%.6f}\n", mt, st, t1 - t0);
        delete g;
        return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 1;
}
