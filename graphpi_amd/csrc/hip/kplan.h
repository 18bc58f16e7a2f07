// Kernel plan: the flattened schedule recompiled into the small tables the
// counting kernel walks.  Built on the host per gp_count call.
#pragma once
#include <cstdint>

#include "graphpi_amd.h"

#define GP_MAX_UB_MASKS 4
#define GP_MAX_BLOCKS 63

struct KPlan {
    int32_t n;           // pattern size
    int32_t leaf;        // depth whose count is evaluated without a loop
    int32_t k;           // IEP width (0 = plain leaf)
    int32_t nprefix;
    int32_t nslots;      // materialised prefix sets per wave
    int32_t fused;       // plain leaf counted while building its set at leaf-1
    int32_t fused_prefix;
    int32_t root_mode;   // leaf == 1: one task per root vertex
    int8_t father[GP_MAX_PREFIX];
    int8_t slot[GP_MAX_PREFIX];      // scratch slot, -1 when not materialised
    int8_t depth_of[GP_MAX_PREFIX];  // depth at which the prefix is built
    uint8_t ub_n[GP_MAX_PREFIX];     // 0 = keep every element
    uint8_t ub_mask[GP_MAX_PREFIX][GP_MAX_UB_MASKS];
    uint8_t rmask[GP_MAX_SIZE];      // restriction bound at each depth: min emb[j], j in mask
    int8_t loop_set[GP_MAX_SIZE];
    int8_t nbuild[GP_MAX_SIZE];
    int8_t build[GP_MAX_SIZE][GP_MAX_PREFIX];
    int32_t nblocks;                  // distinct IEP blocks (subsets of the k trailing vertices)
    uint8_t block_mask[GP_MAX_BLOCKS];
    int32_t ngroups;
    uint8_t group_nb[GP_MAX_IEP_GROUPS];
    uint8_t group_blk[GP_MAX_IEP_GROUPS][GP_MAX_IEP_K];
    int64_t group_val[GP_MAX_IEP_GROUPS];
    // batched last loop level: the candidates of depth leaf-1 are processed
    // together; each leaf block is one of
    //   kind 0: |G \ emb|            G fixed for the whole batch
    //   kind 1: |G ∩ N(v) \ emb|     G fixed, N(v) the candidate's list
    //   kind 2: |N(v) \ emb|
    // where G is the intersection of the block's source prefixes (a scratch
    // slot holds it when there are two or more).  A plain leaf is one block
    // with the restriction bound rmask[leaf]; an IEP leaf uses the groups.
    int32_t batched;
    int32_t nbslots;
    uint8_t bl_kind[GP_MAX_BLOCKS];
    uint8_t bl_nsrc[GP_MAX_BLOCKS];
    int8_t bl_src[GP_MAX_BLOCKS][GP_MAX_IEP_K];
    int8_t bl_slot[GP_MAX_BLOCKS];
    // kind-1 blocks sharing one search: bl_base is a kind-1 block whose set
    // contains every other kind-1 block's set (each of its sources is, or is
    // an ancestor of, a source of every other kind-1 block), or -1.  When the
    // base set is staged in LDS, entry i holds base[i] << 8 | flags, with bit
    // bl_bit[b] set when base[i] is also in block b's set; one search then
    // answers every kind-1 block.  At most 8 kind-1 blocks.
    int32_t bl_base;
    int32_t nk1;
    int8_t bl_bit[GP_MAX_BLOCKS];
    // bit j: block b's set lies inside N(v_j), so the placed vertex v_j is
    // never in it (no self-loops) and needs no membership test
    uint8_t bl_vmask[GP_MAX_BLOCKS];
    // bit d: the loop set at depth d was already cut at exactly the bound of
    // rmask[d] when it was built, so entering depth d needs no lower_bound
    uint32_t cut_done;
    // pair level (plain leaf): depths leaf-2 and leaf-1 evaluated together
    // when the leaf set is f ∩ N(v) with f = g ∩ N(a) built at leaf-2, and
    // the leaf-1 candidates and both bounds do not depend on a.  Then
    //   sum_a sum_{v in C} |G(a) ∩ N(v) ∩ X| - sum_{a in C} |G(a) ∩ X|
    // with G(a) = g ∩ N(a), X = [0, bound(leaf)) minus placed vertices.
    int32_t pair;
    int32_t pair_g;   // prefix g, -1 = every vertex
    // pair level grouped by v1 (leaf = 4, no g): v1 < v0 at depth 1, the
    // depth-3 loop runs over N(v1) below v0.  Tasks are then read from the
    // edge slot reversed (v1 = the slot's row), so that the tasks of one
    // fetch share v1 with v0 ascending; the counts of N(v3) for v3 in
    // N(v1) ∩ [0, v0) grow incrementally across those tasks and only the
    // depth-2 side is streamed per task.
    int32_t pair_v1;
    // set by the launcher: a task whose depth-2 loop has more candidates is
    // deferred and run in parts (INT32_MAX: never)
    int32_t split_min;
    // set by the launcher: the plan's tasks are abandoned once this many
    // wall-clock ticks have passed since the call's first launch started
    // (0: never)
    int64_t abort_ticks;
    // set by the launcher: the plan is abandoned the same way when a fetch of
    // its tasks (a part, in the parts launch) has run this many ticks
    // (0: never)
    int64_t task_ticks;
};

// Loop-nest shapes with a schedule-specialised kernel.
#define GP_SHAPE_GENERIC 0
#define GP_SHAPE_HOUSE 1     // House in the GraphPi order: v1 < v0 in N(v0), v2 in N(v0), IEP over
                             // v3 in N(v0) ∩ N(v1) and v4 in N(v1) ∩ N(v2)
#define GP_SHAPE_HOUSE_ALL 2 // the same loop nest without the restriction v1 < v0
#define GP_SHAPE_TRI_EDGE_TRI 3 // two triangles joined by an edge in the GraphPi order: loops over the
                                // 3-path v0-v1-v2-v3 with v3 < v0, IEP over v4 in N(v2) ∩ N(v3) and
                                // v5 in N(v0) ∩ N(v1); redundancy 4 (raw sum = 4 x the count)
int gp_sched_shape(const gp_sched_flat* s);

// Returns GP_OK or an error code (with gp_fail message) for unsupported input.
int gp_make_kplan(const gp_sched_flat* s, bool faithful, bool allow_batch, bool allow_pair, KPlan* out);
