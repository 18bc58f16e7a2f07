/*
 * gp_oracle.h — TEST INFRASTRUCTURE ONLY.  A CPU restatement of GraphPi's
 * counting loop nest used as the parity checker and as the CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product (graphpi_amd) never links or calls it.
 *
 * Parity pins for each function are listed in gp_oracle.c.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H
#include <stdint.h>
#include "../include/graphpi_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Graph::pattern_matching restated (src/graph.cpp:199-398, src/vertex_set.cpp).
 * Roots v = root_begin, root_begin+root_stride, ... < root_end (root_end < 0
 * means v_cnt).  raw_sum: unsigned two's-complement sum before the division
 * by redundancy.  balg: algorithmic bytes of the loop nest (may be NULL).
 * expanded: get_edge_index calls (may be NULL).  Returns raw_sum / redundancy
 * as int64 (meaningful for the full root range). */
int64_t gpo_pattern_matching(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                             const gp_sched_flat* s, int threads, int64_t root_begin,
                             int64_t root_end, int64_t root_stride, uint64_t* raw_sum,
                             uint64_t* balg, uint64_t* expanded);

/* The same loop nest with a depth-1 filter, for slicing a count the way the
 * device slices it: only embeddings whose depth-1 vertex lies in
 * [v1_begin, v1_end) are counted (v1_end < 0 means v_cnt).  The reference has
 * no such filter; with the full range this is gpo_pattern_matching.  For a
 * pattern whose leaf is at depth 1 the filter is never consulted. */
int64_t gpo_pattern_matching_v1(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                                const gp_sched_flat* s, int threads, int64_t root_begin,
                                int64_t root_end, int64_t root_stride, int64_t v1_begin,
                                int64_t v1_end, uint64_t* raw_sum, uint64_t* balg,
                                uint64_t* expanded);

/* Timed sampling for a CPU baseline: the loop nest over the given roots (a
 * sample of root ids), OpenMP dynamic chunk 1 as above, until time_limit_s
 * has passed; a root still running then is abandoned and counts for
 * nothing.  raw_sum: the sum over the roots that finished; roots_done: how
 * many; busy_s: thread-seconds spent on them.  Returns roots_done. */
int64_t gpo_pattern_matching_timed(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                                   const gp_sched_flat* s, int threads, const int32_t* roots,
                                   int64_t n_roots, double time_limit_s, uint64_t* raw_sum,
                                   int64_t* roots_done, double* busy_s);

/* Graph::triangle_counting_mt restated (src/graph.cpp:112-139, :66-95). */
int64_t gpo_triangle_counting(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                              int threads);

/* Independent brute force: number of injective maps of the pattern into the
 * graph that send pattern edges to graph edges (non-induced), divided by the
 * number of automorphisms of the pattern.  Uses no schedule, restriction or
 * IEP machinery; for small graphs only. */
int64_t gpo_brute_force(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt, int32_t size,
                        const int32_t* adj);

#ifdef __cplusplus
}
#endif
#endif
