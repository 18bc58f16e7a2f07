// Graph: the drop-in surface of the reference's include/graph.h.  Counting
// goes to the device through the C ABI (gp_count / gp_tc); a failure there is
// reported and aborts, as the reference does on internal errors — there is no
// CPU fallback.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "common.h"
#include "graph.h"

namespace {
[[noreturn]] void die(const char* what, int rc) {
    std::fprintf(stderr, "graphpi_amd: %s failed (%d): %s\n", what, rc, gp_last_error());
    std::fflush(stderr);
    std::abort();
}
} // namespace

Graph::~Graph() {
    release_device();
    delete[] edge;
    delete[] vertex;
}

int Graph::upload() {
    if (dev_) return GP_OK;
    return gp_graph_upload(edge, vertex, v_cnt, e_cnt, &dev_);
}

void Graph::release_device() {
    if (dev_) gp_graph_free(dev_);
    dev_ = nullptr;
}

// Host helpers kept from the reference interface (src/graph.cpp:17-95): the
// number of common neighbours, and of common neighbours below v2.
int Graph::intersection_size(int v1, int v2) {
    unsigned l1, r1, l2, r2;
    get_edge_index(v1, l1, r1);
    get_edge_index(v2, l2, r2);
    int n = 0;
    while (l1 < r1 && l2 < r2) {
        if (edge[l1] < edge[l2])
            ++l1;
        else if (edge[l2] < edge[l1])
            ++l2;
        else {
            ++n;
            ++l1;
            ++l2;
        }
    }
    return n;
}

int Graph::intersection_size_clique(int v1, int v2) {
    unsigned l1, r1, l2, r2;
    get_edge_index(v1, l1, r1);
    get_edge_index(v2, l2, r2);
    int n = 0;
    while (l1 < r1 && l2 < r2 && edge[l1] < v2 && edge[l2] < v2) {
        if (edge[l1] < edge[l2])
            ++l1;
        else if (edge[l2] < edge[l1])
            ++l2;
        else {
            ++n;
            ++l1;
            ++l2;
        }
    }
    return n;
}

long long Graph::triangle_counting() { return triangle_counting_mt(1); }

long long Graph::triangle_counting_mt(int) {
    int rc = upload();
    if (rc != GP_OK) die("gp_graph_upload", rc);
    uint64_t t = 0;
    rc = gp_tc(dev_, &t, &last_stats);
    if (rc != GP_OK) die("gp_tc", rc);
    return (long long)t;
}

long long Graph::pattern_matching(const Schedule& schedule, int, bool) {
    gp_sched_flat flat;
    int rc = schedule.flatten(&flat);
    if (rc != GP_OK) die("Schedule::flatten", rc);
    rc = upload();
    if (rc != GP_OK) die("gp_graph_upload", rc);
    int64_t count = 0;
    // max_running_time bounds the count as the reference's commented-out
    // check did (src/graph.cpp:229-237): checked between bounded launches,
    // and on expiry the process stops with "TIMEOUT!" like its assert(0)
    gp_tuning t = tuning;
    if (t.time_limit_s <= 0 && max_running_time > 0) t.time_limit_s = max_running_time;
    rc = gp_count(dev_, &flat, &t, nullptr, &count, &last_stats);
    if (rc == GP_ETIME) {
        std::printf("TIMEOUT!\n");
        std::fflush(stdout);
    }
    if (rc != GP_OK) die("gp_count", rc);
    return (long long)count;
}

long long Graph::unfold_pattern_matching(const Schedule& schedule, int thread_count, bool clique) {
    return pattern_matching(schedule, thread_count, clique);
}

long long Graph::pattern_matching_mpi(const Schedule& schedule, int thread_count, bool clique) {
    return pattern_matching(schedule, thread_count, clique);
}

// ---- common.h ----

double get_wall_time() {
    struct timeval t;
    if (gettimeofday(&t, nullptr)) return 0;
    return (double)t.tv_sec + (double)t.tv_usec * 0.000001;
}

void PatternType_printer(PatternType type) {
    static const char* names[] = {"Rectangle", "QG3",         "Pentagon",      "House",
                                  "Hourglass", "Cycle_6_Tri", "Clique_7_Minus"};
    if (type != QG3 && type >= Rectangle && type <= Clique_7_Minus) std::printf("%s\n", names[type]);
}

bool is_equal_adj_mat(const int* a, const int* b, int size) {
    for (int i = 0; i < size * size; ++i)
        if (a[i] != b[i]) return false;
    return true;
}

void GetDataType(DataType& type, const std::string str) {
    type = Invalid;
    if (str == "Patents") type = Patents;
    if (str == "Orkut") type = Orkut;
    if (str == "complete8") type = complete8;
    if (str == "LiveJournal") type = LiveJournal;
    if (str == "MiCo") type = MiCo;
    if (str == "Twitter") type = Twitter;
    if (str == "CiteSeer") type = CiteSeer;
    if (str == "Wiki-Vote") type = Wiki_Vote;
}
