// graph.h — undirected graph in CSR form with the counting entry points.
// Same public fields and methods as the reference's include/graph.h:7-61, so
// `g->pattern_matching(schedule, n)` is a drop-in.  The counting methods run
// on the MI355X through the C ABI in graphpi_amd.h; there is no CPU path.
// The first counting call uploads the CSR to the current HIP device and keeps
// it resident until the Graph is destroyed (or release_device() is called).
#pragma once
#include <assert.h>

#include "graphpi_amd.h"
#include "schedule.h"

class Graph {
public:
    int v_cnt;          // number of vertices
    unsigned int e_cnt; // number of directed edges (2 x undirected)
    long long tri_cnt;  // number of triangles, -1 when unknown
    double max_running_time = 60 * 60 * 24; // seconds; a count still running after this aborts with "TIMEOUT!"

    int* edge;             // neighbours, sorted ascending per vertex
    unsigned int* vertex;  // N(v) = edge[vertex[v] .. vertex[v+1])

    Graph() : v_cnt(0), e_cnt(0), tri_cnt(-1), edge(nullptr), vertex(nullptr), max_degree(0) {}
    ~Graph();

    // host helpers on two adjacency lists (reference src/graph.cpp:17-95)
    int intersection_size(int v1, int v2);
    int intersection_size_clique(int v1, int v2);

    // triangles, each counted once (device)
    long long triangle_counting();
    long long triangle_counting_mt(int thread_count);

    // count embeddings of schedule's pattern (device).  thread_count and
    // clique are accepted for interface parity and do not change the result.
    long long pattern_matching(const Schedule& schedule, int thread_count, bool clique = false);
    // same result; kept for code written against the generated-code variant
    long long unfold_pattern_matching(const Schedule& schedule, int thread_count, bool clique = false);
    // single-GPU build: identical to pattern_matching (one device, no ranks)
    long long pattern_matching_mpi(const Schedule& schedule, int thread_count, bool clique = false);

    int max_degree;

    // ---- additions ----
    int max_intersection_size = 0; // second-largest degree
    // upload explicitly (optional); returns a gp_* status
    int upload();
    void release_device();
    const gp_graph* device() const { return dev_; }
    // statistics of the most recent counting call
    gp_stats last_stats{};
    // tuning applied to counting calls
    gp_tuning tuning{};

private:
    gp_graph* dev_ = nullptr;
    Graph(const Graph&);
    Graph& operator=(const Graph&);
    void get_edge_index(int v, unsigned int& l, unsigned int& r) const {
        l = vertex[v];
        r = vertex[v + 1];
    }
};
