// Builds the kernel plan from a flattened schedule.
//
// Truncation by restriction bounds.  A prefix set built at depth dp is read by
//   - the loop at depth d, which stops at bound(d) = min emb[j] over the
//     restriction indices j keyed at d (src/graph.cpp:367-376),
//   - the plain leaf, which counts elements below bound(n-1) (:350-359),
//   - the IEP leaf, which reads every element (:306-344),
//   - its child prefixes (father of), through their own readers.
// So only elements below max over readers of (that reader's bound) are ever
// used, and the set may be cut there when it is built.  Bounds that mention a
// vertex not yet placed at depth dp (index j > dp) are replaced by that
// vertex's own bound: emb[j] < min emb[i] over i in rmask[j], because emb[j]
// passed the loop bound at depth j.  A reader with no bound, or a bound that
// cannot be expressed with placed vertices, keeps the whole set.  Every count
// stays identical: each reader still sees all elements below its own bound.
#include "../hip/kplan.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../gp_internal.h"

namespace {

// rewrite a bound mask so it only mentions vertices placed at depth <= dp;
// returns 0 when that is impossible (unbounded)
uint32_t restrict_mask(uint32_t m, int dp, const uint8_t* rmask) {
    for (int guard = 0; guard < 64; ++guard) {
        uint32_t high = m & ~((2u << dp) - 1);
        if (!high) return m;
        int j = 31 - __builtin_clz(high);
        if (rmask[j] == 0) return 0;
        m = (m & ~(1u << j)) | rmask[j];
    }
    return 0;
}

} // namespace

int gp_make_kplan(const gp_sched_flat* s, bool faithful, bool allow_batch, bool allow_pair, KPlan* kp) {
    std::memset(kp, 0, sizeof(*kp));
    const int n = s->size;
    if (n < 2 || n > GP_MAX_SIZE) return gp_fail(GP_ERANGE, "pattern size must be 2..8");
    const int P = s->total_prefix_num;
    if (P < 0 || P > GP_MAX_PREFIX) return gp_fail(GP_ERANGE, "too many prefixes");
    const int k = s->iep_num > 1 ? s->iep_num : 0;
    if (k > GP_MAX_IEP_K || k >= n) return gp_fail(GP_ERANGE, "IEP width out of range");
    if (s->redundancy == 0) return gp_fail(GP_EINVAL, "redundancy is zero");
    kp->n = n;
    kp->k = k;
    kp->leaf = k ? n - k : n - 1;
    kp->nprefix = P;
    // The batched level tracks at most 15 leaf blocks (IEP width <= 4).  A
    // leaf at depth 2 stays per-edge: batching it would make a hub root one
    // indivisible task.
    kp->batched = allow_batch && !faithful && kp->leaf >= 3 && k <= 4;
    kp->root_mode = kp->leaf == 1;
    for (int p = 0; p < GP_MAX_PREFIX; ++p) {
        kp->father[p] = -1;
        kp->slot[p] = -1;
        kp->depth_of[p] = -1;
    }
    for (int p = 0; p < P; ++p) {
        if (s->father[p] < -1 || s->father[p] >= P) return gp_fail(GP_EINVAL, "bad father prefix id");
        kp->father[p] = (int8_t)s->father[p];
    }
    for (int d = 0; d < n; ++d) {
        kp->loop_set[d] = (int8_t)s->loop_set[d];
        if (d >= 1 && (s->loop_set[d] < 0 || s->loop_set[d] >= P)) return gp_fail(GP_EINVAL, "bad loop set id");
        int cnt = 0;
        for (int p = s->last[d]; p != -1; p = s->next[p]) {
            if (p < 0 || p >= P || cnt >= GP_MAX_PREFIX) return gp_fail(GP_EINVAL, "bad prefix list");
            kp->build[d][cnt++] = (int8_t)p;
            kp->depth_of[p] = (int8_t)d;
        }
        kp->nbuild[d] = (int8_t)cnt;
        uint32_t m = 0;
        int guard = 0;
        for (int r = s->restrict_last[d]; r != -1; r = s->restrict_next[r]) {
            if (r < 0 || r >= GP_MAX_RESTRICT || ++guard > GP_MAX_RESTRICT) return gp_fail(GP_EINVAL, "bad restriction list");
            int idx = s->restrict_index[r];
            if (idx < 0 || idx >= d) return gp_fail(GP_EINVAL, "restriction index must precede its depth");
            m |= 1u << idx;
        }
        kp->rmask[d] = (uint8_t)m;
    }
    for (int p = 0; p < P; ++p)
        if (kp->depth_of[p] < 0 || kp->depth_of[p] >= kp->leaf)
            return gp_fail(GP_EINVAL, "prefix built outside the loop depths");

    // ---- truncation bounds ----
    std::vector<std::vector<uint32_t> > readers(P);
    std::vector<char> unbounded(P, faithful ? 1 : 0);
    auto add_reader = [&](int p, uint32_t mask) {
        if (mask == 0)
            unbounded[p] = 1;
        else
            readers[p].push_back(mask);
    };
    if (!faithful) {
        for (int d = 1; d < kp->leaf; ++d) add_reader(kp->loop_set[d], kp->rmask[d]);
        if (k == 0)
            add_reader(kp->loop_set[kp->leaf], kp->rmask[kp->leaf]);
        else
            for (int i = 0; i < k; ++i) unbounded[kp->loop_set[kp->leaf + i]] = 1;
        // children before fathers: deeper prefixes first
        std::vector<int> order(P);
        for (int p = 0; p < P; ++p) order[p] = p;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return kp->depth_of[a] > kp->depth_of[b]; });
        for (int c : order) {
            // rewrite c's masks for its own depth, then hand them to its father
            std::vector<uint32_t> mine;
            for (uint32_t m : readers[c]) {
                uint32_t r = restrict_mask(m, kp->depth_of[c], kp->rmask);
                if (!r) {
                    unbounded[c] = 1;
                    break;
                }
                mine.push_back(r);
            }
            if (unbounded[c]) mine.clear();
            // drop masks whose bound can never exceed another's (supersets)
            std::vector<uint32_t> minimal;
            for (uint32_t m : mine) {
                bool dominated = false;
                for (uint32_t o : mine)
                    if (o != m && (o & m) == o) dominated = true;
                if (!dominated && std::find(minimal.begin(), minimal.end(), m) == minimal.end()) minimal.push_back(m);
            }
            if (minimal.size() > GP_MAX_UB_MASKS) unbounded[c] = 1;
            readers[c] = minimal;
            int f = kp->father[c];
            if (f >= 0) {
                if (unbounded[c])
                    unbounded[f] = 1;
                else
                    for (uint32_t m : readers[c]) readers[f].push_back(m);
            }
        }
        for (int p = 0; p < P; ++p) {
            if (unbounded[p] || readers[p].empty()) continue;
            kp->ub_n[p] = (uint8_t)readers[p].size();
            for (size_t i = 0; i < readers[p].size(); ++i) kp->ub_mask[p][i] = (uint8_t)readers[p][i];
        }
    }

    kp->fused = 0;
    kp->fused_prefix = -1;
    kp->split_min = INT32_MAX; // the launcher decides

    // ---- scratch slots ----
    // in batched mode the sets of depth leaf-1 are never built: the leaf
    // blocks read their fathers and the candidate's neighbour list instead
    int slots = 0;
    for (int p = 0; p < P; ++p)
        if (kp->father[p] >= 0 && !(kp->batched && kp->depth_of[p] == kp->leaf - 1)) kp->slot[p] = (int8_t)slots++;
    kp->nslots = slots;

    // ---- IEP tables ----
    if (k) {
        if (s->iep_group_num < 1 || s->iep_group_num > GP_MAX_IEP_GROUPS) return gp_fail(GP_EINVAL, "bad IEP group count");
        kp->ngroups = s->iep_group_num;
        for (int g = 0; g < s->iep_group_num; ++g) {
            int nb = 0;
            for (int i = 0; i < k; ++i) nb = std::max(nb, (int)s->iep_block[g][i] + 1);
            if (nb < 1 || nb > k) return gp_fail(GP_EINVAL, "bad IEP partition");
            kp->group_nb[g] = (uint8_t)nb;
            for (int b = 0; b < nb; ++b) {
                uint32_t m = 0;
                for (int i = 0; i < k; ++i)
                    if (s->iep_block[g][i] == b) m |= 1u << i;
                if (!m) return gp_fail(GP_EINVAL, "empty IEP block");
                int idx = -1;
                for (int j = 0; j < kp->nblocks; ++j)
                    if (kp->block_mask[j] == m) idx = j;
                if (idx < 0) {
                    if (kp->nblocks >= GP_MAX_BLOCKS) return gp_fail(GP_ERANGE, "too many IEP blocks");
                    idx = kp->nblocks++;
                    kp->block_mask[idx] = (uint8_t)m;
                }
                kp->group_blk[g][b] = (uint8_t)idx;
            }
            kp->group_val[g] = s->iep_val[g];
        }
    }

    // ---- batched leaf blocks ----
    if (kp->batched) {
        const int d = kp->leaf - 1;
        auto add_source = [&](int b, int p) {
            for (int i = 0; i < kp->bl_nsrc[b]; ++i)
                if (kp->bl_src[b][i] == p) return;
            kp->bl_src[b][kp->bl_nsrc[b]++] = (int8_t)p;
        };
        auto fill_block = [&](int b, uint32_t members) {
            bool uses_nv = false;
            for (int i = 0; i < (k ? k : 1); ++i) {
                if (!(members >> i & 1)) continue;
                const int q = kp->loop_set[kp->leaf + i];
                if (kp->depth_of[q] == d) {
                    uses_nv = true;
                    if (kp->father[q] >= 0) add_source(b, kp->father[q]);
                } else {
                    add_source(b, q);
                }
            }
            kp->bl_kind[b] = (uint8_t)(uses_nv ? (kp->bl_nsrc[b] ? 1 : 2) : 0);
            kp->bl_slot[b] = -1;
            if (kp->bl_nsrc[b] >= 2) kp->bl_slot[b] = (int8_t)kp->nbslots++;
        };
        if (k == 0) {
            kp->nblocks = 1;
            kp->block_mask[0] = 1;
            kp->ngroups = 1;
            kp->group_nb[0] = 1;
            kp->group_blk[0][0] = 0;
            kp->group_val[0] = 1;
            fill_block(0, 1);
        } else {
            for (int b = 0; b < kp->nblocks; ++b) fill_block(b, kp->block_mask[b]);
        }
    }

    // ---- vertex sets of the prefixes: bit j when the prefix lies in N(v_j) ----
    {
        uint32_t vset[GP_MAX_PREFIX] = {0};
        for (int d = 0; d < n; ++d)
            for (int i = 0; i < kp->nbuild[d]; ++i) {
                const int p = kp->build[d][i];
                vset[p] = (1u << d) | (kp->father[p] >= 0 ? vset[kp->father[p]] : 0u);
            }
        for (int b = 0; b < GP_MAX_BLOCKS; ++b) kp->bl_vmask[b] = 0;
        if (kp->batched)
            for (int b = 0; b < kp->nblocks; ++b) {
                uint32_t m = 0;
                for (int i = 0; i < kp->bl_nsrc[b]; ++i) m |= vset[kp->bl_src[b][i]];
                kp->bl_vmask[b] = (uint8_t)m;
            }
        // a loop set built with the single bound mask its depth uses
        kp->cut_done = 0;
        if (!faithful)
            for (int d = 1; d < n; ++d) {
                const int q = kp->loop_set[d];
                if (q >= 0 && kp->rmask[d] && kp->ub_n[q] == 1 && kp->ub_mask[q][0] == kp->rmask[d])
                    kp->cut_done |= 1u << d;
            }
    }

    // ---- shared search for the kind-1 blocks ----
    kp->bl_base = -1;
    kp->nk1 = 0;
    for (int b = 0; b < GP_MAX_BLOCKS; ++b) kp->bl_bit[b] = -1;
    if (kp->batched) {
        for (int b = 0; b < kp->nblocks; ++b)
            if (kp->bl_kind[b] == 1) kp->bl_bit[b] = (int8_t)kp->nk1++;
        // is p equal to q or an ancestor of q (so that q's set is inside p's,
        // truncation included: q is built from p's stored set)?
        auto covers = [&](int p, int q) {
            for (int x = q; x >= 0; x = kp->father[x])
                if (x == p) return true;
            return false;
        };
        if (kp->nk1 >= 2 && kp->nk1 <= 8) {
            for (int bb = 0; bb < kp->nblocks && kp->bl_base < 0; ++bb) {
                if (kp->bl_kind[bb] != 1) continue;
                bool ok = true;
                for (int b = 0; b < kp->nblocks && ok; ++b) {
                    if (b == bb || kp->bl_kind[b] != 1) continue;
                    for (int i = 0; i < kp->bl_nsrc[bb] && ok; ++i) {
                        bool found = false;
                        for (int j = 0; j < kp->bl_nsrc[b]; ++j) found = found || covers(kp->bl_src[bb][i], kp->bl_src[b][j]);
                        ok = found;
                    }
                }
                if (ok) kp->bl_base = bb;
            }
        }
    }

    // ---- pair level ----
    kp->pair = 0;
    kp->pair_g = -1;
    if (allow_pair && kp->batched && k == 0 && kp->leaf >= 4) {
        const int L = kp->leaf;
        const int q = kp->loop_set[L];
        const int f = kp->father[q];
        const int c = kp->loop_set[L - 1];
        if (kp->depth_of[q] == L - 1 && f >= 0 && kp->depth_of[f] == L - 2 && kp->nbuild[L - 2] == 1 &&
            kp->nbuild[L - 1] == 1 && kp->depth_of[c] < L - 2 && !(kp->rmask[L - 1] >> (L - 2) & 1) &&
            kp->rmask[L] < (1u << (L - 2))) {
            kp->pair = 1;
            kp->pair_g = kp->father[f];
        }
    }
    kp->pair_v1 = 0;
    if (kp->pair && kp->leaf == 4 && kp->pair_g < 0) {
        const int c = kp->loop_set[3];
        if (kp->rmask[1] == 1u && c >= 0 && kp->father[c] < 0 && kp->depth_of[c] == 1 && kp->rmask[3] == 1u &&
            (kp->rmask[4] & ~3u) == 0)
            kp->pair_v1 = 1;
    }
    return GP_OK;
}

extern "C" int gp_kernel_plan(const gp_sched_flat* sched, gp_plan_info* out) {
    if (!sched || !out) return gp_fail(GP_EINVAL, "null argument");
    KPlan kp;
    const int rc = gp_make_kplan(sched, false, true, true, &kp);
    if (rc != GP_OK) return rc;
    std::memset(out, 0, sizeof(*out));
    out->leaf = kp.leaf;
    out->iep_k = kp.k;
    out->root_tasks = kp.root_mode;
    out->batched = kp.batched;
    out->pair = kp.pair;
    out->pair_grouped_v1 = kp.pair_v1;
    out->shared_search = kp.batched && kp.bl_base >= 0;
    out->blocks = kp.batched ? kp.nblocks : 0;
    out->kind1_blocks = kp.batched ? kp.nk1 : 0;
    out->scratch_slots = kp.nslots;
    out->shape = gp_sched_shape(sched);
    return GP_OK;
}

// The loop nest a flattened schedule describes, compared with the shapes
// that have a specialised kernel.  Each prefix's set is the intersection of
// N(v_j) over the depths j on its father chain (a prefix whose last element
// is depth d adds N(v_d)); the loop set of depth d must be the intersection
// over d's earlier neighbours in the schedule-ordered adjacency, as the
// planner builds it (src/schedule.cpp:329-362).
namespace {

// the schedule's loop sets and IEP groups are the ones `want` (row-major
// n x n, schedule order) describes, with a two-vertex IEP leaf
bool loop_nest_is(const gp_sched_flat* s, int n, const int* want) {
    if (s->size != n || s->iep_num != 2 || s->iep_group_num != 2) return false;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if ((s->adj[i * n + j] != 0) != (want[i * n + j] != 0)) return false;
    // prefix p -> depth mask of the neighbourhoods it intersects
    const int P = s->total_prefix_num;
    if (P < 1 || P > GP_MAX_PREFIX) return false;
    int depth_of[GP_MAX_PREFIX];
    for (int p = 0; p < P; ++p) depth_of[p] = -1;
    for (int d = 0; d < n; ++d)
        for (int p = s->last[d], guard = 0; p != -1 && guard <= P; p = s->next[p], ++guard) {
            if (p < 0 || p >= P) return false;
            depth_of[p] = d;
        }
    uint32_t mask[GP_MAX_PREFIX];
    for (int p = 0; p < P; ++p) {
        uint32_t m = 0;
        for (int q = p, guard = 0; q != -1 && guard <= P; q = s->father[q], ++guard) {
            if (q < 0 || q >= P || depth_of[q] < 0) return false;
            m |= 1u << depth_of[q];
        }
        mask[p] = m;
    }
    for (int d = 1; d < n; ++d) {
        const int q = s->loop_set[d];
        if (q < 0 || q >= P) return false;
        uint32_t w = 0;
        for (int j = 0; j < d; ++j)
            if (want[d * n + j]) w |= 1u << j;
        if (mask[q] != w) return false;
    }
    // |A||B| - |A ∩ B| (two singleton blocks with weight +1, one block of
    // both with weight -1), in either group order
    int singles = -1, joint = -1;
    for (int g = 0; g < 2; ++g) {
        const int b0 = s->iep_block[g][0], b1 = s->iep_block[g][1];
        if (b0 != b1 && s->iep_val[g] == 1) singles = g;
        if (b0 == b1 && s->iep_val[g] == -1) joint = g;
    }
    return singles >= 0 && joint >= 0;
}

// the restrictions the loops check (depths below the IEP leaf): exactly
// `want_index` at depth `want_depth` (or none when want_depth < 0)
bool loop_restrictions_are(const gp_sched_flat* s, int leaf, int want_depth, int want_index) {
    for (int d = 0; d < leaf; ++d) {
        const int r = s->restrict_last[d];
        if (d == want_depth) {
            if (r < 0 || r >= GP_MAX_RESTRICT || s->restrict_index[r] != want_index || s->restrict_next[r] != -1)
                return false;
        } else if (r != -1) {
            return false;
        }
    }
    return true;
}

} // namespace

int gp_sched_shape(const gp_sched_flat* s) {
    // two triangles joined by an edge: triangle (v0, v1, v5), bridge v1-v2,
    // triangle (v2, v3, v4); loops v0, v1, v2, v3 with v3 < v0; the
    // restrictions on v4 and v5 sit at IEP depths, which the loop nest does
    // not check, and the redundancy 4 makes up for them
    static const int tet[36] = {0, 1, 0, 0, 0, 1, 1, 0, 1, 0, 0, 1, 0, 1, 0, 1, 1, 0,
                                0, 0, 1, 0, 1, 0, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0};
    if (loop_nest_is(s, 6, tet) && loop_restrictions_are(s, 4, 3, 0) && s->redundancy == 4)
        return GP_SHAPE_TRI_EDGE_TRI;
    const int n = s->size;
    if (n != 5 || s->iep_num != 2 || s->iep_group_num != 2) return GP_SHAPE_GENERIC;
    static const int house[5][5] = {{0, 1, 1, 1, 0}, {1, 0, 0, 1, 1}, {1, 0, 0, 0, 1}, {1, 1, 0, 0, 0}, {0, 1, 1, 0, 0}};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if ((s->adj[i * n + j] != 0) != (house[i][j] != 0)) return GP_SHAPE_GENERIC;
    // prefix p -> depth mask of the neighbourhoods it intersects
    const int P = s->total_prefix_num;
    if (P < 1 || P > GP_MAX_PREFIX) return GP_SHAPE_GENERIC;
    int depth_of[GP_MAX_PREFIX];
    for (int p = 0; p < P; ++p) depth_of[p] = -1;
    for (int d = 0; d < n; ++d)
        for (int p = s->last[d], guard = 0; p != -1 && guard <= P; p = s->next[p], ++guard) {
            if (p < 0 || p >= P) return GP_SHAPE_GENERIC;
            depth_of[p] = d;
        }
    uint32_t mask[GP_MAX_PREFIX];
    for (int p = 0; p < P; ++p) {
        uint32_t m = 0;
        for (int q = p, guard = 0; q != -1 && guard <= P; q = s->father[q], ++guard) {
            if (q < 0 || q >= P || depth_of[q] < 0) return GP_SHAPE_GENERIC;
            m |= 1u << depth_of[q];
        }
        mask[p] = m;
    }
    for (int d = 1; d < n; ++d) {
        const int q = s->loop_set[d];
        if (q < 0 || q >= P) return GP_SHAPE_GENERIC;
        uint32_t want = 0;
        for (int j = 0; j < d; ++j)
            if (house[d][j]) want |= 1u << j;
        if (mask[q] != want) return GP_SHAPE_GENERIC;
    }
    // restrictions: none, or only v1 < v0 (checked at depth 1)
    if (s->total_restrict_num != 0 && s->total_restrict_num != 1) return GP_SHAPE_GENERIC;
    for (int d = 0; d < n; ++d) {
        const int r = s->restrict_last[d];
        if (d == 1 && s->total_restrict_num == 1) {
            if (r < 0 || r >= GP_MAX_RESTRICT || s->restrict_index[r] != 0 || s->restrict_next[r] != -1)
                return GP_SHAPE_GENERIC;
        } else if (r != -1) {
            return GP_SHAPE_GENERIC;
        }
    }
    // IEP over v3, v4: |A||B| - |A ∩ B| (two singleton blocks with weight +1,
    // one block of both with weight -1), in either group order
    int singles = -1, joint = -1;
    for (int g = 0; g < 2; ++g) {
        const int b0 = s->iep_block[g][0], b1 = s->iep_block[g][1];
        if (b0 != b1 && s->iep_val[g] == 1) singles = g;
        if (b0 == b1 && s->iep_val[g] == -1) joint = g;
    }
    if (singles < 0 || joint < 0) return GP_SHAPE_GENERIC;
    return s->total_restrict_num == 1 ? GP_SHAPE_HOUSE : GP_SHAPE_HOUSE_ALL;
}
