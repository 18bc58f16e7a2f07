/*
 * gp_oracle.c — TEST INFRASTRUCTURE ONLY (see gp_oracle.h).
 *
 * A plain-C restatement of GraphPi's CPU counting loop nest, kept structurally
 * close to the reference so that it is a faithful checker and a fair CPU
 * baseline: recursive per-depth loops, per-thread prefix-set buffers, scalar
 * merge intersections, binary-search subtraction and OpenMP dynamic chunk-1
 * scheduling over roots.
 *
 * Reference lines restated:
 *   gpo_pattern_matching   src/graph.cpp:199-245 (root loop), :247-398 (levels,
 *                          IEP leaf :306-344, plain leaf :346-364, inner :366-396)
 *   set_intersect          src/vertex_set.cpp:56-120 (merge :105-119), :122-156
 *   build_prefix           src/vertex_set.cpp:158-170
 *   subtraction_size       src/vertex_set.cpp:196-209
 *   gpo_triangle_counting  src/graph.cpp:66-95, :112-139
 *
 * Parity pins (tests/test_oracle.py):
 *   - golden counts written by the unmodified reference compiled from source
 *     (tests/golden/*.json: wiki-vote and R-MAT graphs, many patterns and
 *     planner settings), compared bit for bit;
 *   - wiki-vote triangles = 608,389 (include/dataloader.h:23 of the reference);
 *   - gpo_brute_force (no schedule, no restrictions, no IEP) on random small
 *     graphs for every pattern of 3-5 vertices, and networkx subgraph
 *     monomorphism counts for the brute force itself;
 *   - closed forms: K_n, K_{a,b}, cycles, stars;
 *   - invariants: IEP on/off, restricted x |Aut| = unrestricted, relabelling.
 * B_alg (the algorithmic byte count) has no reference counterpart; it is
 * pinned by hand-computed values on tiny graphs in the same test file.
 */
#include "gp_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const int32_t* data;
    int32_t size;
    int32_t* buf; /* owned storage when the prefix is an intersection */
} vset;

typedef struct {
    const int32_t* edge;
    const uint32_t* vertex;
    int32_t v_cnt;
    const gp_sched_flat* s;
    vset sets[GP_MAX_PREFIX];
    int32_t emb[GP_MAX_SIZE];
    int32_t emb_n;
    int32_t* tmp; /* IEP scratch, max_degree entries */
    int32_t v1_begin, v1_end; /* depth-1 filter (test slicing; the reference has none) */
    double deadline;          /* timed sampling: abandon the root once omp_get_wtime() passes it (0: never) */
    int abandoned;
    uint32_t tick;
    uint64_t ans, balg, expanded;
} ctx_t;

/* sorted merge; out may alias a (in-place intersection_with) */
static int32_t set_intersect(const int32_t* a, int32_t na, const int32_t* b, int32_t nb,
                             int32_t* out) {
    int32_t i = 0, j = 0, n = 0;
    while (i < na && j < nb) {
        int32_t x = a[i], y = b[j];
        if (x < y)
            ++i;
        else if (x > y)
            ++j;
        else {
            out[n++] = x;
            ++i;
            ++j;
        }
    }
    return n;
}

static int binary_search(const int32_t* d, int32_t n, int32_t x) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (d[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && d[lo] == x;
}

static int32_t lower_bound(const int32_t* d, int32_t n, int32_t x) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (d[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* |set[0:n) \ embedding| by one binary search per embedded vertex */
static int32_t subtraction_size(const int32_t* d, int32_t n, const ctx_t* c) {
    int32_t ret = n;
    for (int j = 0; j < c->emb_n; ++j)
        if (binary_search(d, n, c->emb[j])) --ret;
    return ret;
}

static void build_prefix(ctx_t* c, int p, const int32_t* in, int32_t in_n) {
    int f = c->s->father[p];
    if (f == -1) {
        c->sets[p].data = in;
        c->sets[p].size = in_n;
    } else {
        c->balg += 4ull * ((uint64_t)c->sets[f].size + (uint64_t)in_n);
        c->sets[p].size = set_intersect(c->sets[f].data, c->sets[f].size, in, in_n, c->sets[p].buf);
        c->sets[p].data = c->sets[p].buf;
    }
}

static int32_t bound_at(const ctx_t* c, int depth) {
    int32_t m = c->v_cnt;
    for (int r = c->s->restrict_last[depth]; r != -1; r = c->s->restrict_next[r]) {
        int32_t x = c->emb[c->s->restrict_index[r]];
        if (x < m) m = x;
    }
    return m;
}

static void level(ctx_t* c, int depth) {
    const gp_sched_flat* s = c->s;
    int lp = s->loop_set[depth];
    int32_t loop_size = c->sets[lp].size;
    if (loop_size <= 0) return;
    const int32_t* loop = c->sets[lp].data;
    int k = s->iep_num;

    if (depth == s->size - k) {
        int ids[GP_MAX_SIZE];
        ids[0] = lp;
        for (int i = 1; i < k; ++i) ids[i] = s->loop_set[depth + i];
        for (int g = 0; g < s->iep_group_num; ++g) {
            uint64_t val = (uint64_t)s->iep_val[g];
            int nblocks = 0;
            for (int i = 0; i < k; ++i)
                if (s->iep_block[g][i] + 1 > nblocks) nblocks = s->iep_block[g][i] + 1;
            for (int b = 0; b < nblocks; ++b) {
                int mem[GP_MAX_SIZE], nm = 0;
                for (int i = 0; i < k; ++i)
                    if (s->iep_block[g][i] == b) mem[nm++] = i;
                int32_t sz;
                if (nm == 1) {
                    const vset* v = &c->sets[ids[mem[0]]];
                    sz = subtraction_size(v->data, v->size, c);
                } else {
                    const vset* a = &c->sets[ids[mem[0]]];
                    const vset* bb = &c->sets[ids[mem[1]]];
                    c->balg += 4ull * ((uint64_t)a->size + (uint64_t)bb->size);
                    int32_t tn = set_intersect(a->data, a->size, bb->data, bb->size, c->tmp);
                    for (int m = 2; m < nm; ++m) {
                        const vset* x = &c->sets[ids[mem[m]]];
                        c->balg += 4ull * ((uint64_t)tn + (uint64_t)x->size);
                        tn = set_intersect(c->tmp, tn, x->data, x->size, c->tmp);
                    }
                    sz = subtraction_size(c->tmp, tn, c);
                }
                val = val * (uint64_t)(int64_t)sz;
                if (val == 0) break;
            }
            c->ans += val;
        }
        return;
    }

    if (depth == s->size - 1) {
        if (s->total_restrict_num > 0) {
            int32_t mv = bound_at(c, depth);
            int32_t sar = lower_bound(loop, loop_size, mv);
            if (sar > 0) c->ans += (uint64_t)(int64_t)subtraction_size(loop, sar, c);
        } else {
            c->ans += (uint64_t)(int64_t)subtraction_size(loop, loop_size, c);
        }
        return;
    }

    int32_t mv = bound_at(c, depth);
    for (int32_t i = 0; i < loop_size; ++i) {
        if (c->deadline > 0 && (++c->tick & 1023) == 0 && omp_get_wtime() > c->deadline) c->abandoned = 1;
        if (c->abandoned) return;
        int32_t v = loop[i];
        if (mv <= v) break;
        if (depth == 1 && (v < c->v1_begin || v >= c->v1_end)) continue;
        int dup = 0;
        for (int j = 0; j < c->emb_n; ++j)
            if (c->emb[j] == v) {
                dup = 1;
                break;
            }
        if (dup) continue;
        uint32_t l = c->vertex[v], r = c->vertex[v + 1];
        c->expanded++;
        c->balg += 8;
        int zero = 0;
        for (int p = s->last[depth]; p != -1; p = s->next[p]) {
            build_prefix(c, p, c->edge + l, (int32_t)(r - l));
            if (c->sets[p].size == 0) {
                zero = 1;
                break;
            }
        }
        if (zero) continue;
        c->emb[c->emb_n++] = v;
        level(c, depth + 1);
        c->emb_n--;
    }
}

int64_t gpo_pattern_matching(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                             const gp_sched_flat* s, int threads, int64_t root_begin,
                             int64_t root_end, int64_t root_stride, uint64_t* raw_sum,
                             uint64_t* balg, uint64_t* expanded) {
    return gpo_pattern_matching_v1(edge, vertex, v_cnt, s, threads, root_begin, root_end, root_stride, 0,
                                   -1, raw_sum, balg, expanded);
}

int64_t gpo_pattern_matching_v1(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                                const gp_sched_flat* s, int threads, int64_t root_begin,
                                int64_t root_end, int64_t root_stride, int64_t v1_begin,
                                int64_t v1_end, uint64_t* raw_sum, uint64_t* balg,
                                uint64_t* expanded) {
    if (root_end < 0 || root_end > v_cnt) root_end = v_cnt;
    if (v1_end < 0 || v1_end > v_cnt) v1_end = v_cnt;
    if (v1_begin < 0) v1_begin = 0;
    if (root_stride < 1) root_stride = 1;
    /* buffer sizes: second-largest and largest degree (src/dataloader.cpp:110-115) */
    int32_t d1 = 0, d2 = 0;
    for (int32_t v = 0; v < v_cnt; ++v) {
        int32_t d = (int32_t)(vertex[v + 1] - vertex[v]);
        if (d > d1) {
            d2 = d1;
            d1 = d;
        } else if (d > d2)
            d2 = d;
    }
    if (d2 < 1) d2 = 1;
    if (d1 < 1) d1 = 1;
    uint64_t total = 0, tb = 0, te = 0;
    if (threads < 1) threads = 1;
    int64_t n_roots = root_end > root_begin ? (root_end - root_begin + root_stride - 1) / root_stride : 0;
#pragma omp parallel num_threads(threads) reduction(+ : total, tb, te)
    {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.edge = edge;
        c.vertex = vertex;
        c.v_cnt = v_cnt;
        c.s = s;
        c.v1_begin = (int32_t)v1_begin;
        c.v1_end = (int32_t)v1_end;
        for (int p = 0; p < s->total_prefix_num; ++p)
            if (s->father[p] != -1) c.sets[p].buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)d2);
        c.tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)d1);
#pragma omp for schedule(dynamic) nowait
        for (int64_t t = 0; t < n_roots; ++t) {
            int32_t v = (int32_t)(root_begin + t * root_stride);
            uint32_t l = vertex[v], r = vertex[v + 1];
            c.balg += 8;
            for (int p = s->last[0]; p != -1; p = s->next[p])
                build_prefix(&c, p, edge + l, (int32_t)(r - l));
            c.emb[0] = v;
            c.emb_n = 1;
            level(&c, 1);
            c.emb_n = 0;
        }
        for (int p = 0; p < s->total_prefix_num; ++p) free(c.sets[p].buf);
        free(c.tmp);
        total += c.ans;
        tb += c.balg;
        te += c.expanded;
    }
    if (raw_sum) *raw_sum = total;
    if (balg) *balg = tb;
    if (expanded) *expanded = te;
    int64_t red = s->redundancy ? s->redundancy : 1;
    return (int64_t)total / red;
}

int64_t gpo_pattern_matching_timed(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                                   const gp_sched_flat* s, int threads, const int32_t* roots,
                                   int64_t n_roots, double time_limit_s, uint64_t* raw_sum,
                                   int64_t* roots_done, double* busy_s) {
    int32_t d1 = 0, d2 = 0;
    for (int32_t v = 0; v < v_cnt; ++v) {
        int32_t d = (int32_t)(vertex[v + 1] - vertex[v]);
        if (d > d1) {
            d2 = d1;
            d1 = d;
        } else if (d > d2)
            d2 = d;
    }
    if (d2 < 1) d2 = 1;
    if (d1 < 1) d1 = 1;
    if (threads < 1) threads = 1;
    const double deadline = omp_get_wtime() + time_limit_s;
    uint64_t total = 0;
    int64_t done = 0;
    double busy = 0;
#pragma omp parallel num_threads(threads) reduction(+ : total, done, busy)
    {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.edge = edge;
        c.vertex = vertex;
        c.v_cnt = v_cnt;
        c.s = s;
        c.v1_begin = 0;
        c.v1_end = v_cnt;
        c.deadline = deadline;
        for (int p = 0; p < s->total_prefix_num; ++p)
            if (s->father[p] != -1) c.sets[p].buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)d2);
        c.tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)d1);
#pragma omp for schedule(dynamic) nowait
        for (int64_t t = 0; t < n_roots; ++t) {
            if (omp_get_wtime() > deadline) continue;
            const double t0 = omp_get_wtime();
            const uint64_t before = c.ans;
            int32_t v = roots[t];
            uint32_t l = vertex[v], r = vertex[v + 1];
            for (int p = s->last[0]; p != -1; p = s->next[p])
                build_prefix(&c, p, edge + l, (int32_t)(r - l));
            c.emb[0] = v;
            c.emb_n = 1;
            c.abandoned = 0;
            level(&c, 1);
            c.emb_n = 0;
            if (c.abandoned) {
                c.ans = before; /* an unfinished root counts for nothing */
            } else {
                total += c.ans - before;
                ++done;
                busy += omp_get_wtime() - t0;
            }
            c.ans = before;
        }
        for (int p = 0; p < s->total_prefix_num; ++p) free(c.sets[p].buf);
        free(c.tmp);
    }
    if (raw_sum) *raw_sum = total;
    if (roots_done) *roots_done = done;
    if (busy_s) *busy_s = busy;
    return done;
}

int64_t gpo_triangle_counting(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt,
                              int threads) {
    int64_t total = 0;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic) reduction(+ : total)
    for (int32_t v = 0; v < v_cnt; ++v) {
        for (uint32_t e = vertex[v]; e < vertex[v + 1]; ++e) {
            int32_t u = edge[e];
            if (v <= u) break;
            /* common neighbours w < u of v and u */
            uint32_t i = vertex[v], ie = vertex[v + 1], j = vertex[u], je = vertex[u + 1];
            while (i < ie && j < je) {
                int32_t x = edge[i], y = edge[j];
                if (x >= u || y >= u) break;
                if (x < y)
                    ++i;
                else if (x > y)
                    ++j;
                else {
                    ++total;
                    ++i;
                    ++j;
                }
            }
        }
    }
    return total;
}

/* ---------------- independent brute force ---------------- */

typedef struct {
    const int32_t* edge;
    const uint32_t* vertex;
    int32_t v_cnt, n;
    const int32_t* adj;
    int32_t map[GP_MAX_SIZE];
    uint64_t count;
} bf_t;

static int adjacent(const bf_t* b, int32_t x, int32_t y) {
    return binary_search(b->edge + b->vertex[x], (int32_t)(b->vertex[x + 1] - b->vertex[x]), y);
}

static void bf_rec(bf_t* b, int i) {
    if (i == b->n) {
        b->count++;
        return;
    }
    int anchor = -1;
    for (int j = 0; j < i; ++j)
        if (b->adj[i * b->n + j]) {
            anchor = j;
            break;
        }
    int32_t lo = 0, hi = b->v_cnt;
    const int32_t* cand = NULL;
    if (anchor >= 0) {
        int32_t a = b->map[anchor];
        cand = b->edge + b->vertex[a];
        lo = 0;
        hi = (int32_t)(b->vertex[a + 1] - b->vertex[a]);
    }
    for (int32_t t = lo; t < hi; ++t) {
        int32_t x = cand ? cand[t] : t;
        int ok = 1;
        for (int j = 0; j < i && ok; ++j) {
            if (b->map[j] == x) ok = 0;
            else if (b->adj[i * b->n + j] && !adjacent(b, b->map[j], x)) ok = 0;
        }
        if (!ok) continue;
        b->map[i] = x;
        bf_rec(b, i + 1);
    }
}

static int64_t automorphisms(int n, const int32_t* adj) {
    int p[GP_MAX_SIZE];
    for (int i = 0; i < n; ++i) p[i] = i;
    int64_t cnt = 0;
    for (;;) {
        int ok = 1;
        for (int i = 0; i < n && ok; ++i)
            for (int j = 0; j < n; ++j)
                if (adj[i * n + j] != adj[p[i] * n + p[j]]) {
                    ok = 0;
                    break;
                }
        cnt += ok;
        /* next permutation */
        int i = n - 2;
        while (i >= 0 && p[i] >= p[i + 1]) --i;
        if (i < 0) break;
        int j = n - 1;
        while (p[j] <= p[i]) --j;
        int t = p[i];
        p[i] = p[j];
        p[j] = t;
        for (int a = i + 1, z = n - 1; a < z; ++a, --z) {
            t = p[a];
            p[a] = p[z];
            p[z] = t;
        }
    }
    return cnt;
}

int64_t gpo_brute_force(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt, int32_t size,
                        const int32_t* adj) {
    if (size < 1 || size > GP_MAX_SIZE) return -1;
    bf_t b;
    memset(&b, 0, sizeof(b));
    b.edge = edge;
    b.vertex = vertex;
    b.v_cnt = v_cnt;
    b.n = size;
    b.adj = adj;
    bf_rec(&b, 0);
    return (int64_t)(b.count / (uint64_t)automorphisms(size, adj));
}
