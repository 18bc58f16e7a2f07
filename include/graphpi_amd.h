/*
 * graphpi_amd.h — C ABI of the MI355X counting engine.
 *
 * Plain C types only (no C++ or torch types), so cgo / JNI / ctypes / N-API
 * bindings can declare these directly.  Every function returns an int status:
 * GP_OK (0) on success, a positive hipError_t value for HIP failures, or a
 * negative GP_E* code for API misuse.  gp_last_error() returns the text of the
 * calling thread's most recent failure.  No exception crosses this boundary.
 *
 * Reference interfaces replaced (thu-pacman/GraphPi, paths relative to the
 * reference repository):
 *   - Graph::pattern_matching(const Schedule&, int, bool)   include/graph.h:39,
 *     src/graph.cpp:199-398 (the OpenMP loop nest)          -> gp_count
 *   - VertexSet set algebra  src/vertex_set.cpp:56-209       -> kernels behind gp_count
 *   - Graph::triangle_counting_mt  src/graph.cpp:112-139     -> gp_tc
 *   - Graph CSR fields  include/graph.h:9-15                 -> gp_graph_upload / gp_graph_free
 *   - Schedule runtime tables  include/schedule.h:23-37,56-58 -> gp_sched_flat
 *   - Schedule constructor  src/schedule.cpp:9-253           -> gp_schedule_build (host only)
 *   - DataLoader::load_data  src/dataloader.cpp:8-190        -> gp_csr_load (host only)
 *   - MotifGenerator::generate  src/motif_generator.cpp:39   -> gp_motifs (host only)
 */
#ifndef GRAPHPI_AMD_H
#define GRAPHPI_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_API __attribute__((visibility("default")))

/* ---- status codes ---- */
#define GP_OK 0
#define GP_EINVAL (-1)     /* bad argument */
#define GP_ENOMEM (-2)     /* allocation failed or over the configured budget */
#define GP_ERANGE (-3)     /* size limit exceeded (pattern too large, ...) */
#define GP_EIO (-4)        /* file missing or malformed */
#define GP_EPATTERN (-5)   /* pattern has a disconnected prefix (is_pattern_valid == false) */
#define GP_ENODEV (-6)     /* no HIP device */
#define GP_ETIME (-7)      /* gp_tuning.time_limit_s ran out before the count finished */

/* ---- limits ---- */
#define GP_MAX_SIZE 8                       /* pattern vertices */
#define GP_MAX_PREFIX 28                    /* size*(size-1)/2 */
#define GP_MAX_RESTRICT 28
#define GP_MAX_IEP_K 6                      /* trailing independent set counted by IEP */
#define GP_MAX_IEP_GROUPS 203               /* Bell(6) */

/*
 * Flattened Schedule: a POD copy of the tables Graph::pattern_matching reads
 * (include/schedule.h:23-37,56-58; built by src/schedule.cpp:329-398,1114-1210,
 * 1746-1764).  Unused tail entries are -1 / 0.
 *   father[p]          get_father_prefix_id(p), -1 = the neighbour list itself
 *   last[d], next[p]   prefixes whose last element is depth d (linked list)
 *   loop_set[d]        get_loop_set_prefix_id(d)
 *   restrict_*         restriction lists keyed by depth: vertex at depth d must
 *                      be < embedding[restrict_index[r]] for r in list(d)
 *   iep_num            in_exclusion_optimize_num (0 = off)
 *   iep_block[g][i]    block id (0-based, in order of first appearance) of
 *                      trailing vertex i in partition g; iep_val[g] its weight
 *   redundancy         get_in_exclusion_optimize_redundancy()
 */
typedef struct gp_sched_flat {
    int32_t size;
    int32_t total_prefix_num;
    int32_t father[GP_MAX_PREFIX];
    int32_t next[GP_MAX_PREFIX];
    int32_t last[GP_MAX_SIZE];
    int32_t loop_set[GP_MAX_SIZE];
    int32_t total_restrict_num;
    int32_t restrict_last[GP_MAX_SIZE];
    int32_t restrict_next[GP_MAX_RESTRICT];
    int32_t restrict_index[GP_MAX_RESTRICT];
    int32_t iep_num;
    int32_t iep_group_num;
    int8_t iep_block[GP_MAX_IEP_GROUPS][GP_MAX_SIZE];
    int64_t iep_val[GP_MAX_IEP_GROUPS];
    int64_t redundancy;
    int32_t adj[GP_MAX_SIZE * GP_MAX_SIZE]; /* schedule-ordered adjacency matrix */
    int32_t restrict_pair_num;             /* Schedule::restrict_pair after pruning */
    int32_t restrict_pair[GP_MAX_RESTRICT][2];
    int32_t multiplicity;                  /* |Aut(pattern)| */
} gp_sched_flat;

/* Counting options.  Zero-initialise and set what you need. */
typedef struct gp_tuning {
    int32_t waves_per_cu;   /* persistent waves per CU; 0 = default (24) */
    int32_t chunk;          /* tasks per work fetch; 0 = 16, or fewer when there are under
                               32 fetches per wave */
    int32_t lds_stage;      /* set elements the LDS hash pool holds (pool = the power of two
                               >= 2 x lds_stage, at most 8192 entries); 0 = default
                               (512 for the batched leaf level, 1024 for the pair level) */
    int32_t no_truncate;    /* 1 = disable restriction-bound truncation */
    int32_t collect_balg;   /* 1 = run the reference-faithful loop nest and count B_alg */
    int32_t no_batch;       /* 1 = evaluate the last loop level one candidate at a time */
    int32_t no_pair;        /* 1 = never evaluate the last two loop levels together */
    int32_t split_min;      /* a task whose depth-2 loop has more candidates than this is
                               deferred and run as up to 64 parts in a second launch;
                               0 = default (1024; 256 when the last two loop levels run
                               as a pair), < 0 = never split */
    /* Count only tasks [task_begin, task_end) (task_end <= 0: to the end).
     * Tasks are directed CSR edge slots, or root vertices when the pattern's
     * leaf is at depth 1 (gp_count_tasks gives the total).  Raw sums of
     * disjoint ranges add up to the raw sum of the whole count. */
    int64_t task_begin;
    int64_t task_end;
    int32_t pair_map;       /* last two levels as a pair: 0 = dense per-wave count array when
                               it fits (64 GiB over all waves), else the LDS map; 1 = LDS map */
    int32_t probe_pct;      /* batched level: a long candidate list is probed instead of streamed
                               when probing's modelled cost x probe_pct / 100 is lower;
                               0 = default (100) */
    int32_t use_graph;      /* 1 = run the call's device work (plan upload, counter reset, the
                               count and deferred-parts launches, result read-back) as one
                               hipGraph, kept per call context and replayed with updated
                               kernel arguments while the launch shape repeats; 0 = stream
                               launches.  Counts are identical either way.  Covers the generic
                               kernel; the House kernel (no_special = 0) runs as stream launches */
    int32_t no_shared_search; /* 1 = batched leaf level: search every block's set on its own
                                 instead of one flagged search of a containing set */
    int32_t no_pair_v1;     /* 1 = pair level: stream both sides per task instead of grouping
                               the tasks of a fetch by v1 */
    int32_t bitmap_ratio;   /* batched level: a set goes into an LDS bitmap when the bitmap's
                               words are at most this many times its sorted table's entries;
                               0 = default (16), < 0 = never */
    int32_t slice_ms;       /* gp_count runs a long count as consecutive launches of about this
                               many milliseconds: on the generic kernel each launch stops
                               fetching tasks at that deadline (counted from its first wave's
                               start) and the next resumes at the first task not counted; the
                               House kernel takes root ranges sized from the previous range's
                               rate.  A launch overruns by the fetches in flight and its
                               deferred parts.  0 = default (2000), < 0 = one launch */
    int32_t no_special;     /* 1 = never use a schedule-specialised kernel (House under the
                               GraphPi schedule is counted per root with two-hop counts and
                               edge supports by default); counts are identical either way */
    double time_limit_s;    /* gp_count: when > 0, stop between launches once this much wall
                               time has passed and return GP_ETIME (the reference's
                               Graph::max_running_time, include/graph.h:12); every launch
                               still runs to its end.  0 = no limit */
    int32_t merge_ratio;    /* > 0: prefix-set builds merge two lists (merge path, tiles staged in
                               LDS) instead of searching the longer one per element of the shorter,
                               when the longer does not fit the LDS pool, the shorter has at least
                               256 elements and the longer is at most this many times as long.
                               Single-schedule counts whose plan takes no other extended path;
                               its own kernel instantiation.  0 = off (the default: on the R-MAT
                               and Chung-Lu configurations restriction bounds cut hub lists short
                               and merging measured no faster).  Counts are identical either way */
    double abort_s;         /* gp_count, gp_count_batch_progress: when > 0, a launch still running
                               this long after the call started abandons its work (wherever its
                               waves are), and the call returns GP_ETIME with the result of the
                               launches that completed before it (a prefix of the tasks); the clock
                               is read down to the intersection loops, so the call returns within
                               about 15 ms of the limit.  For callers that must answer by a
                               deadline even when one task would run past it (a serving loop).
                               Such a gp_count takes the generic kernel, in the instantiation with
                               these checks (counts without a hard stop compile them out) */
    double task_abort_s;    /* gp_count_batch_progress: when > 0, a schedule one of whose work
                               fetches (a few tasks, sized for about a millisecond from the wave's
                               last fetch) has run this long is abandoned as at its abort time
                               (aborted[i] = 1; its launches before stand).  Bounds how long a
                               launch runs past its slice, and lets a caller find the counts whose
                               tasks are long and give them launches of their own */
} gp_tuning;

/* Per-call statistics (all optional outputs). */
typedef struct gp_stats {
    double kernel_ms;       /* hipEvent time of the counting kernel(s) */
    uint64_t balg_bytes;    /* algorithmic bytes (collect_balg=1 only) */
    uint64_t tasks;         /* depth-1 tasks processed */
    int32_t waves;          /* persistent waves launched */
    int32_t kernel_variant; /* which kernel ran */
    double finish_p50_ms;   /* time after kernel start by which half the waves had finished */
    double finish_p90_ms;   /* ... 90% of the waves */
    double finish_max_ms;   /* ... every wave (load-balance tail) */
    uint64_t split_tasks;   /* tasks deferred and run in parts */
    uint64_t inplace_tasks; /* heavy tasks past the deferred list's 2^20 entries, run in place */
    int32_t launches;       /* count launches (task ranges) the call made */
    int32_t finished;       /* 1 when every task was counted; 0 after GP_ETIME */
    uint64_t tasks_total;   /* tasks the call was asked to count */
    int32_t exact_count_valid; /* 1 when the count is also known without wrap-around (a closed form
                                  computed it in 128 bits): exact_count_hi:lo, a signed 128-bit
                                  integer.  The returned count follows the reference's long long
                                  arithmetic, which wraps past 2^63 */
    uint64_t exact_count_lo;
    uint64_t exact_count_hi;
} gp_stats;

typedef struct gp_graph gp_graph; /* device-resident CSR (opaque) */

/* ---- device side ---- */

/* Copy a host CSR (include/graph.h:9-15 layout: vertex[v_cnt+1], edge[e_cnt],
 * e_cnt directed, each list sorted ascending) to the current HIP device.
 * The CSR is checked first (one pass over it): vertex[0] == 0, offsets
 * non-decreasing up to vertex[v_cnt] == e_cnt, ids in [0, v_cnt), each list
 * strictly ascending with no self-loop; anything else is GP_EINVAL. */
GP_API int gp_graph_upload(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                           uint32_t e_cnt, gp_graph** out);
GP_API int gp_graph_free(gp_graph* g);
/* Bytes of HBM the handle holds. */
GP_API int gp_graph_bytes(const gp_graph* g, uint64_t* bytes);

/* Count one flattened schedule over an uploaded graph.  *raw_sum receives the
 * unsigned 64-bit two's-complement sum of the per-root counts (before the
 * division by sched->redundancy); *count receives (int64)raw_sum / redundancy,
 * which is what Graph::pattern_matching returns (src/graph.cpp:244).  Either
 * pointer may be NULL, as may tuning and stats.  Synchronous on the caller's
 * thread; uses its own HIP stream.  A long count runs as several bounded
 * launches over consecutive task ranges (gp_tuning.slice_ms); the raw sums of
 * the ranges add up.  With gp_tuning.time_limit_s set, the call returns
 * GP_ETIME once the limit has passed between two launches: *raw_sum then
 * holds the partial sum of the ranges counted, stats->tasks how many tasks
 * they covered, and *count is left unchanged. */
GP_API int gp_count(const gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning,
                    uint64_t* raw_sum, int64_t* count, gp_stats* stats);

/* gp_count without allocating device memory, for callers that cannot call
 * hipMalloc (code under stream capture, tight serving loops, harnesses that
 * forbid it).  No counterpart in the reference.  Every buffer the count needs
 * besides the uploaded graph and the caller's outputs lives in one static
 * device array of 128 MiB per device, held for the life of the process; the
 * count runs as one launch over the tuning's task range (gp_count with
 * slice_ms < 0) on the null stream, on the generic kernel, with fewer
 * persistent waves when the per-wave prefix sets would not fit, at most 2^16
 * deferred tasks (the rest run in place) and the last two loop levels on the
 * LDS count map in the ordinary task order.  Results equal gp_count's.  A schedule whose per-wave sets
 * alone exceed the array is rejected with GP_ENOMEM.  Calls on one device
 * are serialised inside the library (they share the array). */
GP_API int gp_count_noalloc(const gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning,
                            uint64_t* raw_sum, int64_t* count, gp_stats* stats);

/* Count n_scheds schedules over one uploaded graph with persistent launches
 * shared by the schedules: the schedules are grouped by the LDS their kernel
 * plan needs (at most four groups), and within a group the work items of
 * every schedule share one work counter in one launch (plus one launch for the
 * parts of deferred heavy tasks), so schedules of uneven cost balance across
 * the device.  raw_sums[i] / counts[i] equal what gp_count returns
 * for scheds[i]; stats (optional) covers the whole batch.  tuning applies to
 * every schedule; a task range is rejected (GP_EINVAL) when n_scheds > 1.
 * At most 256 schedules per call.  No reference counterpart: it serves the
 * many-pattern drivers (MotifGenerator output, all_schedule_test.cpp). */
GP_API int gp_count_batch(const gp_graph* g, const gp_sched_flat* scheds, int32_t n_scheds,
                          const gp_tuning* tuning, uint64_t* raw_sums, int64_t* counts, gp_stats* stats);

/* gp_count_batch over part of each schedule's tasks, resumable.  Schedule i
 * counts its tasks from begin[i] (NULL: from 0) in launches over the
 * schedules' remaining tasks laid end to end, each of which stops fetching
 * after gp_tuning.slice_ms (or what is left of gp_tuning.time_limit_s) and
 * so counts a prefix of them; the time limit is checked between launches.
 * abort_s[i] (NULL: gp_tuning.abort_s for every schedule; <= 0: none) is
 * schedule i's hard stop, in seconds after the call starts: a launch still
 * on its tasks then abandons them (wherever its waves are) and that launch's
 * part of schedule i is discarded; aborted[i] (optional) receives 1 and the
 * schedule takes no further launches, while the others carry on.  On return
 * done[i] is the first task of schedule i not counted (gp_count_tasks when
 * its count is complete) and raw_sums[i] the u64 raw sum of its tasks
 * [begin[i], done[i]).  Raw sums of consecutive task ranges add up, so a
 * caller resumes a count by passing done[] back as begin[] and adding the
 * raw sums; the count is then (int64)raw / redundancy, as gp_count returns
 * it.  Returns GP_OK when every schedule's count is complete, GP_ETIME
 * otherwise (done[] and raw_sums[] are valid either way).  gp_tuning task
 * ranges and collect_balg are rejected (GP_EINVAL).  No reference
 * counterpart: it lets a serving loop give long counts time slices and
 * deadlines (graphpi_amd/serve.py). */
GP_API int gp_count_batch_progress(const gp_graph* g, const gp_sched_flat* scheds, int32_t n_scheds,
                                   const gp_tuning* tuning, const uint64_t* begin, const double* abort_s,
                                   uint64_t* done, uint64_t* raw_sums, int32_t* aborted, gp_stats* stats);

/* Device memory (bytes, besides the uploaded graph) that gp_count_batch with
 * the same arguments would use: per-wave scratch, plans, counters, stamps
 * and the deferred-task list.  Lets a server admit requests by footprint.
 * No reference counterpart (the reference allocates per thread on the host,
 * src/vertex_set.cpp:10-48). */
GP_API int gp_count_footprint(const gp_graph* g, const gp_sched_flat* scheds, int32_t n_scheds,
                              const gp_tuning* tuning, uint64_t* bytes);

/* Number of tasks gp_count splits the schedule's count into (see
 * gp_tuning.task_begin). */
GP_API int gp_count_tasks(const gp_graph* g, const gp_sched_flat* sched, uint64_t* tasks);

/* Which kernel paths a schedule takes.  No counterpart in the reference: it
 * describes how this library evaluates the loop nest of
 * Graph::pattern_matching (src/graph.cpp:199-398) for the schedule, for
 * callers who tune or inspect launches.  Host only; needs no device. */
typedef struct gp_plan_info {
    int32_t leaf;            /* depth evaluated without a loop */
    int32_t iep_k;           /* IEP width (0 = plain leaf) */
    int32_t root_tasks;      /* 1 = tasks are root vertices, 0 = directed edge slots */
    int32_t batched;         /* the last loop level runs as one batch */
    int32_t pair;            /* the last two loop levels run together */
    int32_t pair_grouped_v1; /* the pair level groups a fetch's tasks by v1 (single-schedule launches) */
    int32_t shared_search;   /* one flagged search answers every kind-1 block */
    int32_t blocks;          /* leaf blocks of the batched level (1 for a plain leaf) */
    int32_t kind1_blocks;    /* blocks intersected with each candidate's list */
    int32_t scratch_slots;   /* prefix sets materialised per wave */
    int32_t shape;           /* schedule-specialised kernel a single count takes: 0 = none (the
                                generic kernel), 1 = House with v1 < v0 (two-hop counts per root),
                                2 = the same House loop nest without that restriction, 3 = two
                                triangles joined by an edge (closed form over edge supports) */
} gp_plan_info;
GP_API int gp_kernel_plan(const gp_sched_flat* sched, gp_plan_info* out);

/* Triangle count with each triangle counted once (Graph::triangle_counting_mt,
 * src/graph.cpp:112-139). */
GP_API int gp_tc(const gp_graph* g, uint64_t* triangles, gp_stats* stats);

/* Number of HIP devices visible (0 when none). */
GP_API int gp_device_count(int32_t* n);
/* Free and total bytes of the current HIP device's memory (for admission by
 * footprint; no reference counterpart).  GP_ENODEV without a device. */
GP_API int gp_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);

GP_API const char* gp_last_error(void);

/* ---- host side (no device needed) ---- */

/* Build a Schedule exactly as Schedule(pattern, valid, pm, rt, iep, v_cnt,
 * e_cnt, tri_cnt) does (src/schedule.cpp:9-253) and flatten it.  adj is the
 * row-major size x size 0/1 adjacency of the pattern.  *valid mirrors
 * is_pattern_valid; the call still returns GP_OK when the pattern is invalid. */
GP_API int gp_schedule_build(int32_t size, const int32_t* adj, int32_t performance_modeling_type,
                             int32_t restricts_type, int32_t use_in_exclusion_optimize,
                             int32_t v_cnt, uint32_t e_cnt, int64_t tri_cnt, gp_sched_flat* out,
                             int32_t* valid);

/* Host CSR owned by the library (free with gp_csr_free). */
typedef struct gp_csr {
    int32_t v_cnt;
    uint32_t e_cnt;      /* directed edge count */
    int32_t max_degree;
    int32_t max_intersection_size; /* second-largest degree */
    int64_t tri_cnt;     /* -1 when unknown */
    int32_t* edge;
    uint32_t* vertex;
} gp_csr;

/* Load a GraphPi text graph ("v_cnt e_cnt" then one undirected edge per line)
 * with DataLoader::general_load_data semantics (src/dataloader.cpp:20-163):
 * ids remapped by first appearance, self-loops dropped, duplicates removed,
 * header counts checked.  oriented_type: 0 keep, 1 high degree first, 2 low
 * degree first.  binary=1 reads the u32 CSR format of
 * DataLoader::twitter_load_data (src/dataloader.cpp:165-190). */
GP_API int gp_csr_load(const char* path, int32_t binary, int32_t oriented_type, gp_csr** out);
/* Relabel by degree as the text loader's oriented_type does
 * (src/dataloader.cpp:96-110): vertices ranked by degree with std::sort on
 * (id, degree) pairs, 1 = high degree first, 2 = low degree first, 0 = copy.
 * reorder(load_text(f, 0), t) equals load_text(f, t) for a file f without
 * duplicate lines (the reference's degrees count duplicates). */
GP_API int gp_csr_reorder(const gp_csr* in, int32_t oriented_type, gp_csr** out);
/* Complete graph K_n (DataLoader::load_complete, src/dataloader.cpp:192-262). */
GP_API int gp_csr_complete(int32_t n, gp_csr** out);
/* Build from an undirected edge list with the same semantics as writing it as
 * a GraphPi text file and loading it (self-loops dropped, duplicates removed,
 * first-appearance remap).  edges holds 2*n_edges int32 endpoints. */
GP_API int gp_csr_from_edges(const int32_t* edges, uint64_t n_edges, gp_csr** out);
/* Synthetic graphs.  gen = 0: R-MAT(scale, edge_factor, a, b, c, seed);
 * gen = 1: Chung-Lu(n = 1<<scale is NOT used; n = p0, avg undirected edges =
 * p1, Pareto exponent = a, max expected degree = b, seed).  See the generator
 * source for the exact sampling procedure. */
GP_API int gp_csr_generate(int32_t gen, int64_t p0, int64_t p1, double a, double b, double c,
                           uint64_t seed, gp_csr** out);
/* Write a GraphPi text file (unique undirected edges u<v, ids as in the CSR). */
GP_API int gp_csr_write_text(const gp_csr* g, const char* path);
/* Write the binary u32 CSR [v_cnt, e_cnt, max_deg, vertex[v_cnt+1], edge[e_cnt]]. */
GP_API int gp_csr_write_binary(const gp_csr* g, const char* path);
/* FNV-1a 64 over vertex[] then edge[] (identifies a CSR bit for bit). */
GP_API int gp_csr_hash(const gp_csr* g, uint64_t* hash);
GP_API int gp_csr_free(gp_csr* g);

/* All connected patterns on n vertices in MotifGenerator order
 * (src/motif_generator.cpp:5-44).  adj_out receives count*n*n int32 entries;
 * pass NULL to query *count first. */
GP_API int gp_motifs(int32_t n, int32_t* adj_out, int32_t capacity, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* GRAPHPI_AMD_H */
