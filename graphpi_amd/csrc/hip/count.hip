// Counting kernels and the device half of the C ABI.
//
// One wave64 per workgroup; a persistent grid of waves pulls chunks of tasks
// from a global atomic counter.  A task is a directed edge (v0, v1) = the
// first two loop levels of the reference's nest (src/graph.cpp:213-225 and the
// depth-1 loop of :373-396), or a root vertex when the pattern's leaf is at
// depth 1.  From there the wave walks the remaining loop levels depth-first
// with an explicit stack; every lane follows the same control path and the set
// algebra of src/vertex_set.cpp is done cooperatively by the 64 lanes:
//   - intersection: lanes take 64 elements of the shorter list at a time and
//     binary-search the longer one, then ballot + popcount place the hits in
//     order (a sorted result, like the reference's merge);
//   - unorderd_subtraction_size: one lane per embedded vertex, binary search;
//   - lower_bound for restriction cut-offs: 64-way search, about 3 probes
//     for a 100k-element list.
// Counts accumulate in unsigned 64-bit (two's-complement wrap-around, as the
// reference's long long sums do in practice); the host divides by the IEP
// redundancy once, after the global sum (src/graph.cpp:244).
//
// The FAITHFUL instantiation disables restriction truncation and leaf fusion
// and replays the reference's intersection sequence (prefix build order with
// its early exit, IEP group order with its zero early exit) so that it can
// tally the algorithmic byte count B_alg: 4 bytes x (|A| + |B|) for every
// intersection the reference performs plus 8 bytes per neighbour-list lookup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../gp_internal.h"
#include "../range_plan.h"
#include "special.h"
#include "kplan.h"
#include "schedule.h"

#define WAVE 64
#define GP_MAX_BATCH_BLOCKS 15  // leaf blocks the batched level handles
#define GP_LONG_LIST 64         // candidate lists at least this long are streamed by the whole wave
#define GP_NO_HALT (~0ull)      // halt tick of a plan without a hard stop
#define GP_HALT_STEPS 16        // long loops read the clock every this many 64-element steps
#define GP_UNROLL 4             // independent loads in flight per lane on a long list
// LDS hash pool per wave (dynamic shared memory, a power of two): the staged
// leaf sets of the batched level, or the pair level's count map (keys in the
// first half, counts in the second)
#define GP_HASH_SLOTS_BATCH 1024
#define GP_HASH_SLOTS_PAIR 2048
#define GP_HASH_SLOTS_MAX 8192
// Heavy tasks: a task whose depth-2 loop has more than split_min candidates
// is deferred to a second launch that runs it as up to GP_SPLIT_PARTS parts
// (part p takes candidates p, p + parts, p + 2 parts, ...), one part per
// work fetch, so one hub edge no longer decides the kernel's length.
#define GP_SPLIT_DEFAULT 1024     // candidates; per part the depth-1 sets are rebuilt
#define GP_SPLIT_DEFAULT_PAIR 256 // pair-level tasks cost far more per candidate
#define GP_SPLIT_GRAIN 64         // candidates per part, at least
#define GP_SPLIT_PARTS 64
#define GP_HEAVY_CAP (1u << 20) // deferred tasks per call; the rest run in place
#define GP_MAX_BATCH 256          // schedules per launch (8 bits of a deferred task)
#define GP_DENSE_MAX_BYTES (64ull << 30) // pair level: dense per-wave counts only below this total
#define GP_SCRATCH_CAP (32ull << 30)     // per-wave scratch of one call, at most (fewer waves beyond it)
#define GP_EXT_NONE 0
#define GP_EXT_SEARCH 1
#define GP_EXT_PAIR 2
#define GP_EXT_MERGE 3
#define GP_SHARED_MIN_DEG 4096 // the flagged shared search only on graphs with a larger max degree
#define GP_SAMPLE_MIN 64   // LDS entries a sampled index needs at least
#define GP_BITMAP_RATIO 16 // a block set goes into an LDS bitmap when its words <= this x its sorted table
#define GP_FLAG_VMAX ((1 << 23) - 1) // flagged LDS tables pack vertex ids below this with 8 flag bits
#define GP_MERGE_MIN 256   // merge-path intersection: the shorter list has at least this many elements
#define GP_MERGE_TILE 256  // merged elements per merge-path tile, at most (LDS: two lists of this many)
#define GP_FETCH_US 1000   // adaptive fetch size: microseconds of work a fetch aims at
#define GP_GRAPH_CACHE 8          // instantiated launch graphs kept per call context

namespace {

struct KArgs {
    const KPlan* plans;       // one per schedule counted by this launch
    int32_t nplans;
    const uint64_t* task_off;   // plan p owns work items [task_off[p], task_off[p+1]), a multiple of chunk
    const uint64_t* task_first; // first task (edge slot or root) of plan p
    const uint64_t* task_cnt;   // tasks of plan p (its items past that are padding)
    const uint32_t* vertex;
    const int32_t* edge;
    const int32_t* src;
    int32_t v_cnt;
    int32_t chunk;
    uint64_t n_tasks;         // task_off[nplans]
    unsigned long long* counter;
    int32_t* scratch;
    int64_t slot_stride;
    int64_t wave_stride;
    int64_t tmp_stride;
    int64_t bslot_stride;
    int32_t* dense;          // pair level: per-wave count array over all vertex ids (kept zero), or null
    int64_t dense_stride;
    int32_t probe_pct;    // probing is chosen when its cost x probe_pct / 100 is below streaming's
    int32_t bitmap_ratio; // bitmap staging when its words <= bitmap_ratio x the sorted table (0: never)
    int32_t hash_slots;   // LDS hash pool entries (power of two, >= 64)
    int32_t hash_lg;
    int32_t split_min;    // defer when the depth-2 loop has more candidates (INT_MAX: never)
    int32_t merge_ratio;  // prefix builds: merge-path intersection up to this length ratio (0: never)
    int64_t fetch_ticks;  // adaptive fetch size: wall-clock ticks a fetch aims at (0: always chunk items)
    int64_t deadline_ticks; // first launch: stop fetching after this many ticks (0: run every item)
    unsigned long long* aborted; // per plan: set when a wave abandoned one of its tasks (KPlan::abort_ticks)
    unsigned long long* origin;  // start time of the call's first launch (zero before it)
    unsigned long long* heavy;     // deferred: plan << 56 | task << 24 | min(candidates, 2^24 - 1)
    unsigned long long* heavy_cnt; // number deferred (may exceed heavy_cap)
    uint64_t heavy_cap;            // entries of the deferred list; the rest run in place
    unsigned long long* counter2;  // work counter of the parts launch
    unsigned long long* out; // per plan p: [3p] raw sum, [3p+1] B_alg bytes, [3p+2] tasks taken
    unsigned long long* clocks; // per wave: start, finish (wall clock ticks)
};

__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }

// Every set pointer in this kernel points into device memory (the CSR or the
// per-wave scratch), but many pass through LDS on their way, which makes the
// compiler fall back to flat addressing; reading through these types gives
// plain global loads that do not wait on LDS traffic.
typedef const int32_t __attribute__((address_space(1)))* gcptr;
typedef int32_t __attribute__((address_space(1)))* gptr;
#define GL(p) ((gcptr)(p))
#define GLW(p) ((gptr)(p))

__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// lane 0's value as a wave-uniform (scalar) value: lets the compiler keep
// what is derived from it, such as the current plan, in scalar registers
__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (uint64_t)hi << 32 | lo;
}

// sorted-list membership, one lane one probe sequence
__device__ __forceinline__ bool contains(const int32_t* __restrict__ ap, int32_t n, int32_t x) {
    const gcptr a = GL(ap);
    int32_t lo = 0, len = n;
    while (len > 0) {
        int32_t half = len >> 1;
        int32_t mid = lo + half;
        bool right = a[mid] < x;
        lo = right ? mid + 1 : lo;
        len = right ? len - half - 1 : half;
    }
    return lo < n && a[lo] == x;
}

// first index with a[i] >= x; wave-uniform result
__device__ int32_t wave_lower_bound(const int32_t* __restrict__ ap, int32_t n, int32_t x) {
    const gcptr a = GL(ap);
    const int lane = lane_id();
    int32_t lo = 0, hi = n;
    while (hi - lo > WAVE) {
        int32_t step = (hi - lo + WAVE - 1) / WAVE;
        int32_t pos = lo + lane * step;
        bool lt = pos < hi && a[pos] < x;
        int c = popc64(__ballot(lt));
        if (c == 0) return lo;
        int32_t nlo = lo + (c - 1) * step + 1;
        int32_t nhi = min(hi, lo + c * step);
        lo = nlo;
        hi = nhi;
    }
    int32_t pos = lo + lane;
    bool lt = pos < hi && a[pos] < x;
    return lo + popc64(__ballot(lt));
}

// |{emb[j] : j < m} ∩ a[0:n)|
__device__ int32_t count_emb_in(const int32_t* __restrict__ a, int32_t n, const int32_t* emb, int m) {
    const int lane = lane_id();
    bool hit = false;
    if (lane < m && n > 0) hit = contains(a, n, emb[lane]);
    return popc64(__ballot(hit));
}

// A ∩ B restricted to elements < ub.  COUNT: number of such elements not in
// emb[0:m).  Otherwise the sorted result is written to out (no exclusion).
__device__ __forceinline__ int ilog2_ceil(int32_t x);
__device__ void stage_sorted(int32_t* tab, int lg, const int32_t* __restrict__ g, int32_t n);
__device__ __forceinline__ bool staged_has(const int32_t* tab, int lg, int32_t x);

__device__ int32_t wave_excl_scan(int32_t x, int32_t* total);

// Merge-path intersection of two sorted lists (every element below the
// caller's bound), written sorted to out.  The wave walks the merged order
// of S and L in tiles of T elements (T a multiple of 64, tile: 2T LDS
// entries): it stages the next T elements of each list, finds where the
// tile's diagonal splits them (S first on ties), and each lane merges its
// T/64 consecutive elements of the tile from LDS.  An element x of S is a
// hit when it equals the first element of L not yet consumed; with ties
// taking S first that element is at most T - 1 past the tile's start in L,
// so it is staged even when it falls in the next tile.  A lane counts its
// hits, an exclusive scan places them, and the lane merges again to write
// them.  Every list element is read from HBM with coalesced loads, about
// twice (what is staged past the split is read again by the next tile),
// where a binary search reads log2 |L| scattered elements per element of S:
// cheaper when both lists are long and of similar length (hub lists).
__device__ int32_t intersect_merge(const int32_t* __restrict__ Sp, int32_t ns, const int32_t* __restrict__ Lp,
                                   int32_t nl, int32_t* __restrict__ outp, int32_t* tile, int32_t T) {
    const gcptr S = GL(Sp), Lg = GL(Lp);
    const gptr out = GLW(outp);
    const int lane = lane_id();
    const int32_t per = T / WAVE;
    int32_t* const ts = tile;
    int32_t* const tl = tile + T;
    int32_t i = 0, j = 0, total = 0;
    while (i < ns && j < nl) { // once either list is used up nothing else can match
        for (int32_t t = lane; t < T; t += WAVE) {
            ts[t] = i + t < ns ? S[i + t] : INT_MAX;
            tl[t] = j + t < nl ? Lg[j + t] : INT_MAX;
        }
        __syncthreads();
        const int32_t dt = min(T, (ns - i) + (nl - j));
        // elements of S among the first d of the tile's merged order
        auto split = [&](int32_t d) -> int32_t {
            int32_t lo = 0, hi = d;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (ts[mid] <= tl[d - mid - 1])
                    lo = mid + 1;
                else
                    hi = mid;
            }
            return lo;
        };
        const int32_t d0 = min(dt, lane * per), d1 = min(dt, d0 + per);
        const int32_t a0 = split(d0), ti = split(dt);
        // k < dt keeps a + b < T, so both reads stay inside the tile; a
        // padded INT_MAX never ties a real element (a + b < the elements left)
        int32_t a = a0, b = d0 - a0, c = 0;
        for (int32_t k = d0; k < d1; ++k) {
            const int32_t sv = ts[a], lv = tl[b];
            const bool take = sv <= lv;
            c += (take && sv == lv) ? 1 : 0;
            a += take ? 1 : 0;
            b += take ? 0 : 1;
        }
        int32_t sum;
        int32_t off = total + wave_excl_scan(c, &sum);
        if (sum > 0 && c > 0) {
            a = a0;
            b = d0 - a0;
            for (int32_t k = d0; k < d1; ++k) {
                const int32_t sv = ts[a], lv = tl[b];
                const bool take = sv <= lv;
                if (take && sv == lv) out[off++] = sv;
                a += take ? 1 : 0;
                b += take ? 0 : 1;
            }
        }
        total += sum;
        i += ti;
        j += dt - ti;
        __syncthreads(); // every lane is done with the tile before it is refilled
    }
    return total;
}

// pool (LDS, pool_n entries, may be null): when the longer list, cut at ub,
// fits, it is staged there sorted and searched in LDS instead of HBM.
// Otherwise, with merge_ratio > 0, lists of at least GP_MERGE_MIN elements
// whose longer one is at most merge_ratio times the shorter are merged
// (intersect_merge, tiles in the pool) instead of searched.
// halt_at: a wall-clock tick past which the caller's plan is abandoned
// (GP_NO_HALT: never).  Every GP_HALT_STEPS x 64 elements of the shorter list
// the search checks the clock; past halt_at it raises *halt_flag and returns
// the hits found so far (a prefix of the result, so every later access stays
// inside the true set's bounds; the plan's sums are discarded).
template <bool COUNT>
__device__ int32_t intersect(const int32_t* __restrict__ A, int32_t na, const int32_t* __restrict__ B,
                             int32_t nb, int32_t ub, int32_t* __restrict__ outp, const int32_t* emb,
                             int m, int32_t* pool = nullptr, int32_t pool_n = 0, int32_t merge_ratio = 0,
                             uint64_t halt_at = GP_NO_HALT, unsigned long long* halt_flag = nullptr) {
    const int lane = lane_id();
    const gptr out = GLW(outp);
    const int32_t* S = A;
    const int32_t* Lg = B;
    int32_t ns = na, nl = nb;
    if (na > nb) {
        S = B;
        Lg = A;
        ns = nb;
        nl = na;
    }
    // hits are below ub: search only that part of the longer list
    if (ub != INT_MAX && nl > WAVE) nl = wave_lower_bound(Lg, nl, ub);
    int lgp = -1; // staged in the pool: 2^lgp entries
    if (pool && ns >= 16 && nl >= WAVE) {
        const int lg = ilog2_ceil(nl + 1);
        if ((1 << lg) <= pool_n) {
            stage_sorted(pool, lg, Lg, nl);
            lgp = lg;
        } else if (!COUNT && merge_ratio > 0 && ns >= GP_MERGE_MIN && nl <= (int64_t)ns * merge_ratio &&
                   pool_n >= 2 * WAVE) {
            if (ub != INT_MAX) ns = wave_lower_bound(S, ns, ub);
            return intersect_merge(S, ns, Lg, nl, outp, pool, min(GP_MERGE_TILE, (pool_n / 2) & ~(WAVE - 1)));
        }
    }
    const uint64_t below_me = (1ull << lane) - 1;
    int32_t total = 0;
    for (int32_t base = 0; base < ns; base += WAVE) {
        int32_t i = base + lane;
        bool in = i < ns;
        int32_t x = in ? GL(S)[i] : 0;
        bool keep = in && x < ub;
        bool hit = keep && (lgp >= 0 ? staged_has(pool, lgp, x) : contains(Lg, nl, x));
        if (COUNT)
            for (int j = 0; j < m; ++j) hit = hit && x != emb[j];
        uint64_t mk = __ballot(hit);
        if (!COUNT && hit) out[total + popc64(mk & below_me)] = x;
        total += popc64(mk);
        if (__ballot(in && !keep)) break; // the rest of S is >= ub
        if (halt_at != GP_NO_HALT && ((base / WAVE) & (GP_HALT_STEPS - 1)) == GP_HALT_STEPS - 1 &&
            first_lane_u64(wall_clock64()) > halt_at) {
            if (halt_flag && lane == 0) atomicOr(halt_flag, 1ull);
            break;
        }
    }
    return total;
}

// ---- wave utilities ----

// exclusive prefix sum over the wave; *total receives the sum
__device__ __forceinline__ int32_t wave_excl_scan(int32_t x, int32_t* total) {
    const int lane = lane_id();
    int32_t v = x;
    for (int off = 1; off < WAVE; off <<= 1) {
        int32_t y = __shfl_up(v, off);
        if (lane >= off) v += y;
    }
    *total = __shfl(v, WAVE - 1);
    return v - x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        uint32_t l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
        uint64_t s = ((uint64_t)hi << 32 | lo) + ((uint64_t)h2 << 32 | l2);
        lo = (uint32_t)s;
        hi = (uint32_t)(s >> 32);
    }
    return (uint64_t)hi << 32 | lo;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
    for (int off = WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// lane-private lower bound (first index with a[i] >= x)
__device__ __forceinline__ int32_t lower_bound_lane(const int32_t* __restrict__ ap, int32_t n, int32_t x) {
    const gcptr a = GL(ap);
    int32_t lo = 0, len = n;
    while (len > 0) {
        int32_t half = len >> 1;
        int32_t mid = lo + half;
        bool right = a[mid] < x;
        lo = right ? mid + 1 : lo;
        len = right ? len - half - 1 : half;
    }
    return lo;
}

// open-addressing hash set of vertex ids in LDS (empty slot = -1)
__device__ __forceinline__ uint32_t hash_slot(int32_t x, int lg) {
    return ((uint32_t)x * 2654435761u) >> (32 - lg);
}

__device__ void hash_fill(int32_t* tab, int lg, const int32_t* __restrict__ g, int32_t n) {
    const int lane = lane_id();
    const int cap = 1 << lg;
    for (int i = lane; i < cap; i += WAVE) tab[i] = -1;
    __syncthreads();
    for (int32_t base = 0; base < n; base += WAVE) {
        int32_t i = base + lane;
        if (i < n) {
            int32_t x = g[i];
            uint32_t h = hash_slot(x, lg);
            for (;;) {
                int32_t old = atomicCAS(&tab[h], -1, x);
                if (old == -1 || old == x) break;
                h = (h + 1) & (cap - 1);
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool hash_has(const int32_t* tab, int lg, int32_t x) {
    const uint32_t mask = (1u << lg) - 1;
    uint32_t h = hash_slot(x, lg);
    for (;;) {
        int32_t y = tab[h];
        if (y == x) return true;
        if (y == -1) return false;
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ int ilog2_ceil(int32_t x) { return x <= 1 ? 0 : 32 - __clz(x - 1); }

// A sorted set copied into LDS and padded with INT_MAX to 2^lg entries
// (2^lg > n).  Membership is a branch-free binary search of exactly lg steps,
// the same for every lane, so a wave never diverges on it.
__device__ void stage_sorted(int32_t* tab, int lg, const int32_t* __restrict__ g, int32_t n) {
    const gcptr src = GL(g);
    for (int32_t i = lane_id(); i < (1 << lg); i += WAVE) tab[i] = i < n ? src[i] : INT_MAX;
    __syncthreads();
}

__device__ __forceinline__ bool staged_has(const int32_t* tab, int lg, int32_t x) {
    int32_t pos = 0;
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) pos = tab[pos + step - 1] < x ? pos + step : pos;
    return tab[pos] == x;
}

// GP_UNROLL searches interleaved step by step, so their LDS reads overlap
__device__ __forceinline__ uint32_t staged_has_n(const int32_t* tab, int lg, const int32_t* x) {
    int32_t pos[GP_UNROLL];
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) pos[u] = 0;
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) {
#pragma unroll
        for (int u = 0; u < GP_UNROLL; ++u) pos[u] = tab[pos[u] + step - 1] < x[u] ? pos[u] + step : pos[u];
    }
    uint32_t hits = 0;
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) hits |= (tab[pos[u]] == x[u] ? 1u : 0u) << u;
    return hits;
}

// Bitmap of a sorted set in LDS: bit r of word r / 32 is set when lo + r is
// in the set; bits values are covered.  x outside [lo, lo + bits) is absent.
__device__ __forceinline__ bool bitmap_has(const int32_t* tab, int32_t lo, uint32_t bits, int32_t x) {
    const uint32_t r = (uint32_t)(x - lo);
    const uint32_t w = r < bits ? (uint32_t)tab[r >> 5] : 0u;
    return (w >> (r & 31)) & 1u;
}

// Sampled index of a sorted set g[0:gn) too large for LDS: tab holds
// g[i << sk] for i < m, padded with INT_MAX to 2^lg (> m) entries.  The
// number of samples <= x names the only block of 2^sk elements that can
// hold x, which is then searched in HBM (a cache line or a few, reused by
// every lookup of the task).  x = -1 is never found.
__device__ __forceinline__ bool sampled_has(const int32_t* tab, int lg, int sk, const int32_t* gp, int32_t gn, int32_t x) {
    int32_t pos = 0;
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) pos = tab[pos + step - 1] <= x ? pos + step : pos;
    if (pos == 0) return false;
    const gcptr g = GL(gp);
    int32_t q = (pos - 1) << sk;
    for (int32_t step = (1 << sk) >> 1; step > 0; step >>= 1) {
        const int32_t i = q + step - 1;
        q = (i < gn ? g[i] : INT_MAX) < x ? q + step : q;
    }
    return q < gn && g[q] == x;
}

// GP_UNROLL sampled lookups interleaved step by step
__device__ __forceinline__ uint32_t sampled_has_n(const int32_t* tab, int lg, int sk, const int32_t* gp, int32_t gn,
                                                  const int32_t* x) {
    int32_t pos[GP_UNROLL];
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) pos[u] = 0;
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) {
#pragma unroll
        for (int u = 0; u < GP_UNROLL; ++u) pos[u] = tab[pos[u] + step - 1] <= x[u] ? pos[u] + step : pos[u];
    }
    const gcptr g = GL(gp);
    int32_t q[GP_UNROLL];
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) q[u] = pos[u] > 0 ? (pos[u] - 1) << sk : 0;
    for (int32_t step = (1 << sk) >> 1; step > 0; step >>= 1) {
#pragma unroll
        for (int u = 0; u < GP_UNROLL; ++u) {
            const int32_t i = q[u] + step - 1;
            q[u] = (i < gn ? g[i] : INT_MAX) < x[u] ? q[u] + step : q[u];
        }
    }
    uint32_t hits = 0;
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) hits |= (pos[u] > 0 && q[u] < gn && g[q[u]] == x[u] ? 1u : 0u) << u;
    return hits;
}

// Flagged tables: entry = value << 8 | flags, sorted by value, padded with
// INT_MAX (value 2^23 - 1, never a vertex id below GP_FLAG_VMAX).  A lookup
// searches x << 8; the result is the flags of x, or 0 when x is absent.
__device__ __forceinline__ uint32_t staged_flags(const int32_t* tab, int lg, int32_t x) {
    const int32_t key = x < 0 ? INT_MIN : (int32_t)((uint32_t)x << 8);
    int32_t pos = 0;
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) pos = tab[pos + step - 1] < key ? pos + step : pos;
    const int32_t e = tab[pos];
    return (e >> 8) == x ? (uint32_t)e & 0xffu : 0u;
}

// GP_UNROLL flagged lookups interleaved step by step
__device__ __forceinline__ void staged_flags_n(const int32_t* tab, int lg, const int32_t* x, uint32_t* fl) {
    int32_t key[GP_UNROLL], pos[GP_UNROLL];
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) {
        key[u] = x[u] < 0 ? INT_MIN : (int32_t)((uint32_t)x[u] << 8);
        pos[u] = 0;
    }
    for (int32_t step = (1 << lg) >> 1; step > 0; step >>= 1) {
#pragma unroll
        for (int u = 0; u < GP_UNROLL; ++u) pos[u] = tab[pos[u] + step - 1] < key[u] ? pos[u] + step : pos[u];
    }
#pragma unroll
    for (int u = 0; u < GP_UNROLL; ++u) {
        const int32_t e = tab[pos[u]];
        fl[u] = (e >> 8) == x[u] ? (uint32_t)e & 0xffu : 0u;
    }
}

// a plan's abort time (from the call's first launch, origin) or its
// per-task limit (from the wave's current task, ttask) has passed; then the
// plan's abort flag is raised (its sums are discarded), so every caller that
// abandons work on this answer leaves the plan flagged
__device__ __forceinline__ bool plan_expired(const KPlan* pl, uint64_t origin, uint64_t ttask,
                                             unsigned long long* flag) {
    const uint64_t now = first_lane_u64(wall_clock64());
    const bool gone = (pl->abort_ticks > 0 && now - origin > (uint64_t)pl->abort_ticks) ||
                      (pl->task_ticks > 0 && now - ttask > (uint64_t)pl->task_ticks);
    if (gone && threadIdx.x == 0) atomicOr(flag, 1ull);
    return gone;
}

// FAITHFUL: replay the reference's intersections and tally B_alg.
// PARTS: run the parts of deferred heavy tasks instead of the task list.
// MULTI: several plans share the launch (one plan when false).
// EXT: the extended paths, kept out of the other instantiations, whose
// register allocation they would otherwise burden:
//   GP_EXT_SEARCH: the batched level's flagged shared search
//                  (KPlan::bl_base) and sampled indexes of large sets;
//   GP_EXT_PAIR:   the pair level grouped by v1 (KPlan::pair_v1);
//   GP_EXT_MERGE:  merge-path prefix builds (gp_tuning.merge_ratio > 0).
// A single-schedule launch needs at most one of them (a pair plan has a
// plain leaf, a flagged plan an IEP leaf).
// HALT: some plan of the launch has a hard stop (KPlan::abort_ticks or
// task_ticks); only then are the clock checks compiled in, down to the inner
// loops of intersections and long lists, so that an abandoned plan's waves
// leave within about a millisecond.  Without a hard stop none is read.
template <bool FAITHFUL, bool PARTS, bool MULTI, int EXT, bool HALT>
__global__ void __launch_bounds__(WAVE, MULTI ? 5 : 6) gp_count_kernel(KArgs a) {
    __shared__ const int32_t* s_ptr[GP_MAX_PREFIX];
    __shared__ int32_t s_len[GP_MAX_PREFIX];
    __shared__ int32_t s_emb[GP_MAX_SIZE];
    __shared__ const int32_t* s_lp[GP_MAX_SIZE];
    __shared__ int32_t s_ll[GP_MAX_SIZE];
    __shared__ int32_t s_li[GP_MAX_SIZE];
    __shared__ int32_t s_blk[GP_MAX_BLOCKS];
    // batched leaf level
    // dynamic: the hash pool, then one row of 64 per-candidate counts per leaf block
    extern __shared__ int32_t s_dyn[];
    int32_t* const s_hash = s_dyn;
    int32_t(*const s_bc)[WAVE] = reinterpret_cast<int32_t(*)[WAVE]>(s_dyn + a.hash_slots);
    __shared__ const int32_t* s_bg[GP_MAX_BATCH_BLOCKS];
    __shared__ int32_t s_bn[GP_MAX_BATCH_BLOCKS];
    __shared__ int32_t s_bh[GP_MAX_BATCH_BLOCKS];  // LDS hash offset, -1 = not staged
    __shared__ int32_t s_blg[GP_MAX_BATCH_BLOCKS];
    __shared__ int32_t s_bmo[GP_MAX_BATCH_BLOCKS];  // LDS offset of the block's bitmap, -1 = none
    __shared__ int32_t s_blo[GP_MAX_BATCH_BLOCKS];  // bitmap: smallest element (bit 0)
    __shared__ uint32_t s_bbits[GP_MAX_BATCH_BLOCKS]; // bitmap: bits covered (largest - smallest + 1)
    __shared__ int32_t s_bso[GP_MAX_BATCH_BLOCKS];  // LDS offset of the block's sampled index, -1 = none
    __shared__ int32_t s_bsk[GP_MAX_BATCH_BLOCKS];  // sampled index: log2 of the sampling stride
    __shared__ int32_t s_bemb[GP_MAX_BATCH_BLOCKS]; // bit j: emb[j] in G (kinds 0 and 1)
    __shared__ int32_t s_off[WAVE + 1];
    __shared__ int32_t s_off2[WAVE + 1];
    __shared__ uint64_t s_torigin; // start of the call's first launch (plans' abort times count from it)
    __shared__ uint64_t s_ttask;   // start of the wave's current fetch (a part, in the parts launch)
    __shared__ int32_t s_plane[WAVE]; // lane of the r-th probed candidate

    // A plan with an abort time (KPlan::abort_ticks, for counts that must
    // end by a deadline even when a single task would run past it) or a
    // per-task limit (KPlan::task_ticks, timed per fetch, so that a wave's
    // tail after the launch's deadline stays within it): once either has
    // passed, a wave
    // abandons the plan's task it is in, wherever it is, skips the plan's
    // other tasks and raises a.aborted[plan].  That plan's
    // sums are then meaningless and the host discards them; the launch's
    // other plans are unaffected.
#define GP_PLAN_EXPIRED()                                                                                \
    (HALT && (pl->abort_ticks > 0 || pl->task_ticks > 0) &&                                             \
     plan_expired(pl, first_lane_u64(s_torigin), first_lane_u64(s_ttask), &a.aborted[cur_plan]))
    // (checked per task, per loop iteration and per 64 candidates of a
    // batched or pair level)

    // the plan of the current work item; switching plans flushes the sums
    int cur_plan = 0;
    const KPlan* pl = a.plans;
    const int lane = lane_id();
    if (lane == 0) a.clocks[2 * blockIdx.x] = wall_clock64();
    int n = pl->n, L = pl->leaf, K = pl->k;
    const int32_t v_cnt = a.v_cnt;
    int32_t* wbase = a.scratch + (int64_t)blockIdx.x * a.wave_stride;
    int32_t* bbase = wbase + (int64_t)pl->nslots * a.slot_stride + 2 * a.tmp_stride;
    auto set_plan = [&](int p) {
        cur_plan = p;
        pl = a.plans + p;
        n = pl->n;
        L = pl->leaf;
        K = pl->k;
        bbase = wbase + (int64_t)pl->nslots * a.slot_stride + 2 * a.tmp_stride;
    };
    uint64_t acc = 0, balg = 0, taken = 0;
    uint64_t lacc = 0; // per-lane partial sums of the batched leaf level
    // the depth-2 loop visits candidates dstart, dstart + dstride, ... (a part)
    int32_t dstart = 0, dstride = 1; // stay 0 and 1 unless PARTS

    auto bound = [&](uint32_t mask) -> int32_t {
        int32_t b = v_cnt;
        for (int j = 0; j < n; ++j)
            if (mask >> j & 1) b = min(b, s_emb[j]);
        return b;
    };
    auto ub_of = [&](int p) -> int32_t {
        if (FAITHFUL || pl->ub_n[p] == 0) return INT_MAX;
        int32_t ub = 0;
        for (int i = 0; i < pl->ub_n[p]; ++i) ub = max(ub, bound(pl->ub_mask[p][i]));
        return ub;
    };
    // the wall-clock tick past which the current plan is abandoned
    // (GP_NO_HALT when it has no hard stop), for intersections to check
    auto halt_at = [&]() -> uint64_t {
        uint64_t h = GP_NO_HALT;
        if (pl->abort_ticks > 0) h = first_lane_u64(s_torigin) + (uint64_t)pl->abort_ticks;
        if (pl->task_ticks > 0) h = min(h, first_lane_u64(s_ttask) + (uint64_t)pl->task_ticks);
        return h;
    };
    auto build_prefix = [&](int p, const int32_t* nv, int32_t deg) -> int32_t {
        const int f = pl->father[p];
        const int32_t ub = ub_of(p);
        const int32_t* ptr;
        int32_t len;
        if (f < 0) {
            ptr = nv;
            len = ub == INT_MAX ? deg : wave_lower_bound(nv, deg, ub);
        } else {
            int32_t* out = wbase + (int64_t)pl->slot[p] * a.slot_stride;
            const int32_t* fp = s_ptr[f];
            const int32_t fl = s_len[f];
            if (FAITHFUL) balg += 4ull * (uint64_t)((int64_t)fl + deg);
            len = intersect<false>(fp, fl, nv, deg, ub, out, nullptr, 0, s_hash, a.hash_slots,
                                   EXT == GP_EXT_MERGE ? a.merge_ratio : 0, HALT ? halt_at() : GP_NO_HALT,
                                   &a.aborted[cur_plan]);
            ptr = out;
            __syncthreads(); // make the new set visible to every lane
        }
        s_ptr[p] = ptr;
        s_len[p] = len;
        return len;
    };
    // build the prefixes whose last vertex is depth d; false when one is empty
    auto build_level = [&](int d, const int32_t* nv, int32_t deg) -> bool {
        for (int i = 0; i < pl->nbuild[d]; ++i)
            if (build_prefix(pl->build[d][i], nv, deg) == 0) return false;
        return true;
    };
    // count of the common elements of the block's sets, minus embedded ones
    auto block_count = [&](uint32_t bm) -> int32_t {
        int best = -1;
        int32_t best_len = INT_MAX;
        for (int i = 0; i < K; ++i)
            if (bm >> i & 1) {
                int q = pl->loop_set[L + i];
                if (s_len[q] < best_len) {
                    best_len = s_len[q];
                    best = q;
                }
            }
        const gcptr S = GL(s_ptr[best]);
        int32_t total = 0;
        for (int32_t base = 0; base < best_len; base += WAVE) {
            if (((base / WAVE) & (GP_HALT_STEPS - 1)) == GP_HALT_STEPS - 1 && GP_PLAN_EXPIRED()) break;
            int32_t i = base + lane;
            bool hit = i < best_len;
            int32_t x = hit ? S[i] : 0;
            for (int t = 0; t < K; ++t) {
                if (!(bm >> t & 1)) continue;
                int q = pl->loop_set[L + t];
                if (q == best) continue;
                hit = hit && contains(s_ptr[q], s_len[q], x);
            }
            for (int j = 0; j < L; ++j) hit = hit && x != s_emb[j];
            total += popc64(__ballot(hit));
        }
        return total;
    };
    // count at depth L with emb[0:L) placed and every set built (one candidate)
    auto leaf = [&]() -> uint64_t {
        if (K == 0) {
            const int q = pl->loop_set[L];
            const int32_t b = bound(pl->rmask[L]);
            const int32_t* S = s_ptr[q];
            int32_t c = s_len[q];
            if (pl->rmask[L]) c = wave_lower_bound(S, c, b);
            if (c > 0) c -= count_emb_in(S, c, s_emb, L);
            return (uint64_t)(int64_t)c;
        }
        if (!FAITHFUL) {
            for (int bi = 0; bi < pl->nblocks; ++bi) {
                const uint32_t bm = pl->block_mask[bi];
                int32_t c;
                if (__popc(bm) == 1) {
                    const int q = pl->loop_set[L + __ffs(bm) - 1];
                    c = s_len[q] - count_emb_in(s_ptr[q], s_len[q], s_emb, L);
                } else {
                    c = block_count(bm);
                }
                s_blk[bi] = c;
            }
            uint64_t sum = 0;
            for (int g = 0; g < pl->ngroups; ++g) {
                uint64_t val = (uint64_t)pl->group_val[g];
                for (int j = 0; j < pl->group_nb[g]; ++j) val *= (uint64_t)(int64_t)s_blk[pl->group_blk[g][j]];
                sum += val;
            }
            return sum;
        }
        // reference order: group by group, block by block, stop at zero
        uint64_t sum = 0;
        int32_t* tmp0 = wbase + (int64_t)pl->nslots * a.slot_stride;
        int32_t* tmp1 = tmp0 + a.tmp_stride;
        if (s_len[pl->loop_set[L]] <= 0) return 0; // src/graph.cpp:250-252
        for (int g = 0; g < pl->ngroups; ++g) {
            uint64_t val = (uint64_t)pl->group_val[g];
            for (int j = 0; j < pl->group_nb[g]; ++j) {
                const uint32_t bm = pl->block_mask[pl->group_blk[g][j]];
                int32_t c;
                if (__popc(bm) == 1) {
                    const int q = pl->loop_set[L + __ffs(bm) - 1];
                    c = s_len[q] - count_emb_in(s_ptr[q], s_len[q], s_emb, L);
                } else {
                    int first = -1;
                    const int32_t* cur = nullptr;
                    int32_t cn = 0;
                    int32_t* dst = tmp0;
                    for (int t = 0; t < K; ++t) {
                        if (!(bm >> t & 1)) continue;
                        const int q = pl->loop_set[L + t];
                        if (first < 0) {
                            first = q;
                            continue;
                        }
                        const int32_t* A = cur ? cur : s_ptr[first];
                        const int32_t an = cur ? cn : s_len[first];
                        balg += 4ull * (uint64_t)((int64_t)an + s_len[q]);
                        cn = intersect<false>(A, an, s_ptr[q], s_len[q], INT_MAX, dst, nullptr, 0);
                        __syncthreads();
                        cur = dst;
                        dst = dst == tmp0 ? tmp1 : tmp0;
                    }
                    c = cn - (cn > 0 ? count_emb_in(cur, cn, s_emb, L) : 0);
                }
                val *= (uint64_t)(int64_t)c;
                if (val == 0) break;
            }
            sum += val;
        }
        return sum;
    };
    auto enter = [&](int d) {
        const int q = pl->loop_set[d];
        s_lp[d] = s_ptr[q];
        int32_t len = s_len[q];
        if (pl->rmask[d] && !(pl->cut_done >> d & 1)) len = wave_lower_bound(s_ptr[q], len, bound(pl->rmask[d]));
        s_ll[d] = len;
        s_li[d] = d == 2 ? dstart : 0;
    };

    // The last loop level (depth d = L-1) as one batch: every candidate v of
    // the loop set gets its leaf value from block counts (see KPlan); the
    // sets shared by all candidates are intersected once and staged in an
    // LDS hash, and the candidates' neighbour lists are streamed flattened
    // across the 64 lanes.  Candidates whose list is far longer than the
    // staged sets are handled by probing the sets into the list instead.
    auto leaf_level = [&](int d) __attribute__((always_inline)) {
        const int32_t cstart = d == 2 ? dstart : 0, cstep = d == 2 ? dstride : 1;
        const int32_t ncand = s_ll[d] > cstart ? (s_ll[d] - cstart + cstep - 1) / cstep : 0;
        if (ncand <= 0) return;
        const int32_t* cand = s_lp[d] + cstart;
        const int nb = pl->nblocks;
        // 1. the fixed set G of every block
        int32_t smax = -1;       // largest element of any streamed set
        int nstream = 0;
        // stream or probe: per streamed element, the lookup steps into every
        // block's set (LDS steps count a quarter); per probed set element, one
        // binary search of the list at twice the weight
        int stream_cost_w = 0;
        int32_t stream_total = 0;
        int hoff = 0;
        for (int b = 0; b < nb; ++b) {
            const int kind = pl->bl_kind[b];
            const int32_t* g = nullptr;
            int32_t gn = 0;
            if (kind != 2) {
                const int ns = pl->bl_nsrc[b];
                const int p0 = pl->bl_src[b][0];
                g = s_ptr[p0];
                gn = s_len[p0];
                if (ns >= 2) {
                    int32_t* out = bbase + (int64_t)pl->bl_slot[b] * a.bslot_stride;
                    int32_t* spare = bbase + (int64_t)pl->nbslots * a.bslot_stride;
                    for (int i = 1; i < ns; ++i) {
                        const int pi = pl->bl_src[b][i];
                        // alternate so the last intersection lands in 'out'
                        int32_t* dst = ((ns - 1 - i) & 1) ? spare : out;
                        gn = intersect<false>(g, gn, s_ptr[pi], s_len[pi], INT_MAX, dst, nullptr, 0, nullptr, 0, 0,
                                              HALT ? halt_at() : GP_NO_HALT, &a.aborted[cur_plan]);
                        __syncthreads();
                        g = dst;
                    }
                }
            }
            s_bg[b] = g;
            s_bn[b] = gn;
            s_bh[b] = -1;
            s_blg[b] = 0;
            s_bmo[b] = -1;
            s_bso[b] = -1;
            if (kind != 2) {
                bool hit = lane < d && !(pl->bl_vmask[b] >> lane & 1) && gn > 0 && contains(g, gn, s_emb[lane]);
                s_bemb[b] = (int32_t)__ballot(hit);
            }
            if (kind == 1) {
                ++nstream;
                stream_total += gn;
                if (gn > 0) smax = max(smax, GL(g)[gn - 1]);
            }
        }
        // stage the kind-1 sets: the base with membership flags when it
        // fits (one search per streamed element answers every kind-1
        // block), otherwise each set on its own while the pool lasts
        bool flagged = false;
        if (EXT == GP_EXT_SEARCH && pl->bl_base >= 0 && v_cnt < GP_FLAG_VMAX) {
            const int bb = pl->bl_base;
            const int32_t gn = s_bn[bb];
            const int lg = ilog2_ceil(gn + 1);
            if (gn > 0 && (1 << lg) <= a.hash_slots) {
                const gcptr src = GL(s_bg[bb]);
                for (int32_t i = lane; i < (1 << lg); i += WAVE) {
                    int32_t e = INT_MAX;
                    if (i < gn) {
                        const int32_t y = src[i];
                        uint32_t f = 0;
                        for (int b = 0; b < nb; ++b) {
                            if (pl->bl_kind[b] != 1) continue;
                            if (b == bb || contains(s_bg[b], s_bn[b], y)) f |= 1u << pl->bl_bit[b];
                        }
                        e = (int32_t)((uint32_t)y << 8 | f);
                    }
                    s_hash[i] = e;
                }
                __syncthreads();
                // s_bh[bb] stays -1: the probe path reads plain ids, so it
                // takes the base set from global memory
                s_blg[bb] = lg;
                stream_cost_w = 1 + lg / 4;
                flagged = true;
            }
        }
        if (!flagged) {
            // smallest sets first, so that they keep full tables; a set that
            // no longer fits gets a sampled index (EXT) or is searched in HBM
            uint32_t todo = 0;
            for (int b = 0; b < nb; ++b)
                if (pl->bl_kind[b] == 1) todo |= 1u << b;
            while (todo) {
                int b = __ffs(todo) - 1;
                for (uint32_t r = todo & (todo - 1); r; r &= r - 1) {
                    const int c = __ffs(r) - 1;
                    if (s_bn[c] < s_bn[b]) b = c;
                }
                todo &= ~(1u << b);
                const int32_t gn = s_bn[b];
                const int lg = ilog2_ceil(gn + 1);
                if (gn > 0) {
                    // a set spanning few ids goes into a bitmap: one LDS read
                    // per lookup instead of a search
                    const int32_t lo = GL(s_bg[b])[0], hi = GL(s_bg[b])[gn - 1];
                    const uint32_t bits = (uint32_t)(hi - lo) + 1u;
                    const int32_t words = (int32_t)((bits + 31u) >> 5);
                    if (words <= a.bitmap_ratio << lg && hoff + words <= a.hash_slots) {
                        int32_t* tab = s_hash + hoff;
                        for (int32_t i = lane; i < words; i += WAVE) tab[i] = 0;
                        __syncthreads();
                        const gcptr src = GL(s_bg[b]);
                        for (int32_t i = lane; i < gn; i += WAVE) {
                            const uint32_t r = (uint32_t)(src[i] - lo);
                            atomicOr(reinterpret_cast<uint32_t*>(&tab[r >> 5]), 1u << (r & 31));
                        }
                        __syncthreads();
                        s_bmo[b] = hoff;
                        s_blo[b] = lo;
                        s_bbits[b] = bits;
                        hoff += words;
                        stream_cost_w += 1;
                        continue;
                    }
                }
                if (gn > 0 && hoff + (1 << lg) <= a.hash_slots) {
                    stage_sorted(s_hash + hoff, lg, s_bg[b], gn);
                    s_bh[b] = hoff;
                    s_blg[b] = lg;
                    hoff += 1 << lg;
                    stream_cost_w += 1 + lg / 4;
                    continue;
                }
                if (EXT == GP_EXT_SEARCH && gn > 0 && a.hash_slots - hoff >= GP_SAMPLE_MIN) {
                    // sampled index: every 2^sk-th element, the last sample
                    // <= x names the block of 2^sk elements that may hold x
                    int32_t room = GP_SAMPLE_MIN;
                    while (room * 2 <= a.hash_slots - hoff) room *= 2;
                    const int sk = ilog2_ceil((gn + room - 2) / (room - 1));
                    const int32_t m = (gn + (1 << sk) - 1) >> sk; // samples, m + 1 <= room
                    const int slg = ilog2_ceil(m + 1);
                    int32_t* tab = s_hash + hoff;
                    const gcptr src = GL(s_bg[b]);
                    for (int32_t i = lane; i < (1 << slg); i += WAVE) tab[i] = i < m ? src[(int64_t)i << sk] : INT_MAX;
                    __syncthreads();
                    s_bso[b] = hoff;
                    s_blg[b] = slg;
                    s_bsk[b] = sk;
                    hoff += 1 << slg;
                    stream_cost_w += 1 + slg / 4 + sk + 1;
                    continue;
                }
                stream_cost_w += lg + 1;
            }
        }
        const int32_t* const ftab = s_hash;                  // flagged base table
        const int flg = flagged ? s_blg[pl->bl_base] : 0;
        const bool plain = K == 0;
        const uint32_t leaf_mask = plain ? pl->rmask[L] : 0;
        // 2. candidates in batches of 64
        for (int32_t base = 0; base < ncand; base += WAVE) {
            if (GP_PLAN_EXPIRED()) break;
            const int32_t ci = base + lane;
            bool valid = ci < ncand;
            const int32_t v = valid ? cand[(int64_t)ci * cstep] : -1;
            for (int j = 0; j < d; ++j) valid = valid && v != s_emb[j];
            int32_t nvs = 0, deg = 0;
            if (valid) {
                const uint32_t l = a.vertex[v];
                nvs = (int32_t)l;
                deg = (int32_t)(a.vertex[v + 1] - l);
            }
            const int32_t* nv = a.edge + nvs;
            // leaf bound with this candidate placed at depth d
            int32_t b = INT_MAX;
            if (leaf_mask) {
                b = v_cnt;
                for (int j = 0; j < n; ++j)
                    if (leaf_mask >> j & 1) b = min(b, j == d ? v : s_emb[j]);
            }
            // streamed part of N(v): elements below b and below max(G)+1
            const int32_t scut = smax < 0 ? -1 : min(b, smax + 1);
            int32_t work = 0;
            int32_t pcut = 0; // probed lists: elements below scut
            bool probe = false;
            if (valid && nstream > 0 && scut >= 0) {
                // a list short enough for the flattened path is streamed
                // whole (it filters x < scut), saving the dependent search
                int32_t cut = deg < GP_LONG_LIST ? deg : lower_bound_lane(nv, deg, scut);
                pcut = cut;
                // stream the list, or probe the sets into it when cheaper
                const int64_t stream_cost = (int64_t)cut * stream_cost_w;
                const int64_t probe_cost = (int64_t)stream_total * (ilog2_ceil(cut + 1) + 1) * 2;
                probe = cut > 64 && probe_cost * a.probe_pct < stream_cost * 100;
                work = probe ? 0 : cut;
            }
            // long lists go through the whole-wave path below
            const bool longc = work >= GP_LONG_LIST && deg > 16;
            const int32_t lwork = work;
            if (longc) work = 0;
            for (int b2 = 0; b2 < nb; ++b2)
                if (pl->bl_kind[b2] == 1) s_bc[b2][lane] = 0;
            int32_t total = 0;
            const int32_t off = wave_excl_scan(work, &total);
            s_off[lane] = off;
            if (lane == WAVE - 1) s_off[WAVE] = total;
            __syncthreads();
            for (int32_t f0 = 0; f0 < total; f0 += WAVE) {
                const int32_t f = f0 + lane;
                const bool act = f < total;
                // owner: last lane whose offset is <= f
                int lo = 0, hi = WAVE;
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (s_off[mid] <= f) lo = mid; else hi = mid;
                }
                const int owner = act ? lo : WAVE - 1;
                const int32_t o_nvs = __shfl(nvs, owner);
                const int32_t o_off = __shfl(off, owner);
                const int32_t o_cut = __shfl(scut, owner);
                const int32_t x = act ? a.edge[o_nvs + (f - o_off)] : -1;
                bool keep = act && x < o_cut;
                for (int j = 0; j < d; ++j) keep = keep && x != s_emb[j];
                const int prev = __shfl_up(owner, 1);
                const bool start = act && (lane == 0 || prev != owner);
                const uint64_t starts = __ballot(start || !act);
                const uint64_t above = lane == WAVE - 1 ? 0ull : starts & (~0ull << (lane + 1));
                const int next = above ? __ffsll((long long)above) - 1 : WAVE;
                const uint64_t seg = (next == WAVE ? ~0ull : ((1ull << next) - 1)) & (~0ull << lane);
                const uint32_t fl = flagged && keep ? staged_flags(ftab, flg, x) : 0u;
                for (int b2 = 0; b2 < nb; ++b2) {
                    if (pl->bl_kind[b2] != 1) continue;
                    bool hit = false;
                    if (flagged) {
                        hit = fl >> pl->bl_bit[b2] & 1;
                    } else if (keep) {
                        const int h = s_bh[b2];
                        const int bmo = s_bmo[b2];
                        const int so = EXT == GP_EXT_SEARCH ? s_bso[b2] : -1;
                        hit = bmo >= 0  ? bitmap_has(s_hash + bmo, s_blo[b2], s_bbits[b2], x)
                              : h >= 0  ? staged_has(s_hash + h, s_blg[b2], x)
                              : so >= 0 ? sampled_has(s_hash + so, s_blg[b2], s_bsk[b2], s_bg[b2], s_bn[b2], x)
                                        : contains(s_bg[b2], s_bn[b2], x);
                    }
                    const uint64_t hm = __ballot(hit);
                    if (start) {
                        const int c = popc64(hm & seg);
                        if (c) atomicAdd(&s_bc[b2][owner], c);
                    }
                }
            }
            __syncthreads();
            // long lists: the whole wave streams one list, GP_UNROLL loads in
            // flight per lane, and counts hits block by block (the list stays
            // in cache for the next block).  Every element is below the cut,
            // and embedded vertices in G are subtracted once per candidate.
            uint64_t lm = __ballot(longc);
            while (lm) {
                const int j = __ffsll((long long)lm) - 1;
                lm &= lm - 1;
                const gcptr pv = GL(a.edge) + __shfl(nvs, j);
                const int32_t pn = __shfl(lwork, j);
                // before a list of many steps (the shorter ones of a batch
                // add up to little)
                if (pn >= GP_HALT_STEPS * GP_UNROLL * WAVE && GP_PLAN_EXPIRED()) break; // (flagged)
                if (flagged) {
                    // one pass, one search per element, every kind-1 block at
                    // once: per-lane counts, two 16-bit fields per register
                    // (a lane sees at most pn / 64 + 4 elements of a list)
                    uint32_t pk0 = 0, pk1 = 0, pk2 = 0, pk3 = 0;
                    const int nk1 = pl->nk1;
                    for (int32_t f0 = 0; f0 < pn; f0 += GP_UNROLL * WAVE) {
                        const int32_t f = f0 + lane;
                        int32_t x[GP_UNROLL];
#pragma unroll
                        for (int u = 0; u < GP_UNROLL; ++u) x[u] = f + u * WAVE < pn ? pv[f + u * WAVE] : -1;
                        uint32_t fl[GP_UNROLL];
                        staged_flags_n(ftab, flg, x, fl);
#pragma unroll
                        for (int u = 0; u < GP_UNROLL; ++u) {
                            const uint32_t m = fl[u];
                            pk0 += (m & 1u) | ((m & 2u) << 15);
                            if (nk1 > 2) pk1 += ((m >> 2) & 1u) | ((m & 8u) << 13);
                            if (nk1 > 4) pk2 += ((m >> 4) & 1u) | ((m & 32u) << 11);
                            if (nk1 > 6) pk3 += ((m >> 6) & 1u) | ((m & 128u) << 9);
                        }
                    }
                    for (int b2 = 0; b2 < nb; ++b2) {
                        if (pl->bl_kind[b2] != 1) continue;
                        const int q = pl->bl_bit[b2];
                        const uint32_t w = q < 2 ? pk0 : q < 4 ? pk1 : q < 6 ? pk2 : pk3;
                        int32_t c = (int32_t)wave_sum_u32((w >> (16 * (q & 1))) & 0xffffu);
                        const uint32_t em = (uint32_t)s_bemb[b2];
                        if (em) {
                            bool sub = lane < d && (em >> lane & 1) && contains((const int32_t*)pv, pn, s_emb[lane]);
                            c -= popc64(__ballot(sub));
                        }
                        if (lane == j) s_bc[b2][j] = c;
                    }
                    continue;
                }
                for (int b2 = 0; b2 < nb; ++b2) {
                    if (pl->bl_kind[b2] != 1) continue;
                    // wave-uniform block parameters, kept in scalar registers
                    const int h = __builtin_amdgcn_readfirstlane(s_bh[b2]);
                    const int lg = __builtin_amdgcn_readfirstlane(s_blg[b2]);
                    const int bmo = __builtin_amdgcn_readfirstlane(s_bmo[b2]);
                    const int32_t blo = __builtin_amdgcn_readfirstlane(s_blo[b2]);
                    const uint32_t bbits = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_bbits[b2]);
                    const int so = EXT == GP_EXT_SEARCH ? __builtin_amdgcn_readfirstlane(s_bso[b2]) : -1;
                    const int32_t* g = s_bg[b2];
                    const int32_t gn = __builtin_amdgcn_readfirstlane(s_bn[b2]);
                    int32_t c = 0;
                    // wave-uniform trip count: every lane takes part in every ballot
                    for (int32_t f0 = 0; f0 < pn; f0 += GP_UNROLL * WAVE) {
                        const int32_t f = f0 + lane;
                        int32_t x[GP_UNROLL];
#pragma unroll
                        for (int u = 0; u < GP_UNROLL; ++u) x[u] = f + u * WAVE < pn ? pv[f + u * WAVE] : -1;
                        if (bmo >= 0) {
                            // x = -1 (past the end) falls outside the bitmap
#pragma unroll
                            for (int u = 0; u < GP_UNROLL; ++u)
                                c += popc64(__ballot(bitmap_has(s_hash + bmo, blo, bbits, x[u])));
                        } else if (h >= 0) {
                            // x = -1 (past the end) never matches a staged entry
                            const uint32_t hits = staged_has_n(s_hash + h, lg, x);
#pragma unroll
                            for (int u = 0; u < GP_UNROLL; ++u) c += popc64(__ballot(hits >> u & 1));
                        } else if (EXT == GP_EXT_SEARCH && so >= 0) {
                            const uint32_t hits = sampled_has_n(s_hash + so, lg, __builtin_amdgcn_readfirstlane(s_bsk[b2]), g, gn, x);
#pragma unroll
                            for (int u = 0; u < GP_UNROLL; ++u) c += popc64(__ballot(hits >> u & 1));
                        } else {
#pragma unroll
                            for (int u = 0; u < GP_UNROLL; ++u) c += popc64(__ballot(x[u] >= 0 && contains(g, gn, x[u])));
                        }
                    }
                    const uint32_t em = (uint32_t)s_bemb[b2];
                    if (em) {
                        bool sub = lane < d && (em >> lane & 1) && contains((const int32_t*)pv, pn, s_emb[lane]);
                        c -= popc64(__ballot(sub));
                    }
                    if (lane == j) s_bc[b2][j] = c;
                }
            }
            // candidates whose lists are probed: every (candidate, element of
            // G) pair is one binary search, and the pairs of all probed
            // candidates are spread over the lanes together
            const uint64_t pm = __ballot(probe);
            if (pm) {
                const int np = popc64(pm);
                if (probe) s_plane[popc64(pm & ((1ull << lane) - 1))] = lane;
                __syncthreads();
                bool halted = false;
                for (int b2 = 0; b2 < nb && !halted; ++b2) {
                    if (pl->bl_kind[b2] != 1) continue;
                    const int32_t gn = s_bn[b2];
                    const gcptr g = GL(s_bg[b2]);
                    const int h = s_bh[b2];
                    if (probe) s_bc[b2][lane] = 0;
                    __syncthreads();
                    const int32_t items = np * gn;
                    for (int32_t t0 = 0; t0 < items; t0 += WAVE) {
                        if (((t0 / WAVE) & (GP_HALT_STEPS - 1)) == GP_HALT_STEPS - 1 && GP_PLAN_EXPIRED()) {
                            halted = true; // (flagged)
                            break;
                        }
                        const int32_t t = t0 + lane;
                        const bool act = t < items;
                        const int32_t r = act ? t / gn : 0;
                        const int cl = act ? s_plane[r] : lane;
                        const int32_t o_nvs = __shfl(nvs, cl);
                        const int32_t o_cut = __shfl(pcut, cl);
                        bool hit = false;
                        if (act) {
                            const int32_t i = t - r * gn;
                            const int32_t x = h >= 0 ? s_hash[h + i] : g[i];
                            hit = contains(a.edge + o_nvs, o_cut, x);
                            for (int j = 0; j < d; ++j) hit = hit && x != s_emb[j];
                        }
                        if (hit) atomicAdd(&s_bc[b2][cl], 1);
                    }
                }
            }
            __syncthreads();
            // 3. leaf value of each candidate
            if (valid) {
                // kind-1 counts are already in s_bc; fill in the others
                for (int b2 = 0; b2 < nb; ++b2) {
                    const int kind = pl->bl_kind[b2];
                    if (kind == 2) {
                        // |N(v) ∩ [0, b) \ emb|
                        const int32_t cut = b == INT_MAX ? deg : lower_bound_lane(nv, deg, b);
                        int32_t c = cut;
                        for (int j = 0; j < d; ++j)
                            if (s_emb[j] < b && contains(nv, cut, s_emb[j])) --c;
                        s_bc[b2][lane] = c;
                    } else if (kind == 0) {
                        // |G ∩ [0, b) \ emb|, emb including v itself
                        const int32_t* g = s_bg[b2];
                        const int32_t gn = s_bn[b2];
                        int32_t c = b == INT_MAX ? gn : lower_bound_lane(g, gn, b);
                        const uint32_t em = (uint32_t)s_bemb[b2];
                        for (int j = 0; j < d; ++j)
                            if ((em >> j & 1) && s_emb[j] < b) --c;
                        if (v < b && contains(g, gn, v)) --c;
                        s_bc[b2][lane] = c;
                    }
                }
                uint64_t sum = 0;
                for (int g = 0; g < pl->ngroups; ++g) {
                    uint64_t val = (uint64_t)pl->group_val[g];
                    for (int j = 0; j < pl->group_nb[g]; ++j) val *= (uint64_t)(int64_t)s_bc[pl->group_blk[g][j]][lane];
                    sum += val;
                }
                lacc += sum;
            }
            __syncthreads();
        }
    };

    // Depths L-2 and L-1 together (KPlan::pair).  The depth-(L-2) candidates
    // a are taken 64 at a time and split into sub-batches whose sets
    // G(a) ∩ X fit an LDS count map (key -> number of a whose set holds it);
    // then the lists N(v) of the depth-(L-1) candidates are streamed once per
    // sub-batch and each element adds its count.  Pairs with v == a, which the
    // loop nest skips, are subtracted.  An a whose set is too large for the
    // map goes through the batched leaf level on its own.
    // Pair level with a dense per-wave count array cnt[] over all vertex ids
    // instead of the LDS map: scatter every kept element of every N(a) once
    // (cnt[x] += 1), stream every candidate list N(v) once summing cnt[x],
    // subtract the pairs v == a, then clear what was scattered.  Each list is
    // read once or twice, where the LDS map re-streams the N(v) lists once per
    // sub-batch of a's.  Same filters and result as pair_level below.
    auto pair_level_dense = [&]() {
        const int da = L - 2, dv = L - 1;
        const int32_t astart = da == 2 ? dstart : 0, astep = da == 2 ? dstride : 1;
        const int32_t* D = s_lp[da] + astart;
        const int32_t Dn = s_ll[da] > astart ? (s_ll[da] - astart + astep - 1) / astep : 0;
        if (Dn <= 0) return;
        const int qc = pl->loop_set[dv];
        const int32_t* C = s_ptr[qc];
        int32_t Cn = s_len[qc];
        if (pl->rmask[dv]) Cn = wave_lower_bound(C, Cn, bound(pl->rmask[dv]));
        if (Cn <= 0) return;
        const int32_t b = bound(pl->rmask[L]);
        const int gpf = pl->pair_g;
        const int32_t* gptr = gpf >= 0 ? s_ptr[gpf] : nullptr;
        const int32_t glen = gpf >= 0 ? s_len[gpf] : 0;
        int32_t* const cnt = a.dense + (int64_t)blockIdx.x * a.dense_stride;
        const gcptr E = GL(a.edge);
        // mode 0: scatter kept elements; 1: gather counts; 2: clear
        auto visit = [&](int mode, int32_t x, bool act) {
            if (mode == 0) {
                bool keep = act && x < b;
                for (int j = 0; j < da; ++j) keep = keep && x != s_emb[j];
                if (keep && gpf >= 0) keep = contains(gptr, glen, x);
                if (keep) atomicAdd(&cnt[x], 1);
            } else if (mode == 1) {
                if (act && x < b) lacc += (uint64_t)__hip_atomic_load(&cnt[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (act && x < b) cnt[x] = 0;
            }
        };
        // every list N(u) ∩ [0, b) of the 64 lanes' vertices u (u < 0: none)
        auto stream = [&](int mode, int32_t u) {
            int32_t us = 0, ucut = 0;
            if (u >= 0) {
                const uint32_t l = a.vertex[u];
                us = (int32_t)l;
                const int32_t dg = (int32_t)(a.vertex[u + 1] - l);
                // every mode filters x < b: short lists need no exact cut
                ucut = (b >= v_cnt || dg < GP_LONG_LIST) ? dg : lower_bound_lane(a.edge + l, dg, b);
            }
            const bool lng = ucut >= GP_LONG_LIST;
            int32_t tot = 0;
            const int32_t ex = wave_excl_scan(lng ? 0 : ucut, &tot);
            s_off2[lane] = ex;
            __syncthreads();
            for (int32_t f0 = 0; f0 < tot; f0 += WAVE) {
                const int32_t f = f0 + lane;
                const bool act = f < tot;
                int lo = 0, hi = WAVE;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_off2[mid] <= f) lo = mid; else hi = mid;
                }
                const int32_t os = __shfl(us, lo);
                const int32_t oe = __shfl(ex, lo);
                visit(mode, act ? E[os + (f - oe)] : -1, act);
            }
            uint64_t lm = __ballot(lng);
            while (lm) {
                // a gather may stop (flagged); a scatter or clear runs to the
                // end so that the array is all zero again afterwards
                if (mode == 1 && GP_PLAN_EXPIRED()) break;
                const int j = __ffsll((long long)lm) - 1;
                lm &= lm - 1;
                const gcptr pv = E + __shfl(us, j);
                const int32_t pn = __shfl(ucut, j);
                for (int32_t f0 = 0; f0 < pn; f0 += GP_UNROLL * WAVE) {
                    int32_t x[GP_UNROLL];
#pragma unroll
                    for (int q = 0; q < GP_UNROLL; ++q) x[q] = f0 + q * WAVE + lane < pn ? pv[f0 + q * WAVE + lane] : -1;
#pragma unroll
                    for (int q = 0; q < GP_UNROLL; ++q) visit(mode, x[q], x[q] >= 0);
                }
            }
            __syncthreads();
        };
        auto a_of = [&](int32_t ai) -> int32_t {
            bool va = ai < Dn;
            const int32_t av = va ? D[(int64_t)ai * astep] : -1;
            for (int j = 0; j < da; ++j) va = va && av != s_emb[j];
            return va ? av : -1;
        };
        for (int32_t abase = 0; abase < Dn; abase += WAVE) stream(0, a_of(abase + lane));
        __threadfence();
        for (int32_t cbase = 0; cbase < Cn; cbase += WAVE) {
            if (GP_PLAN_EXPIRED()) break;
            const int32_t ci = cbase + lane;
            bool vc = ci < Cn;
            const int32_t vv = vc ? C[ci] : -1;
            for (int j = 0; j < da; ++j) vc = vc && vv != s_emb[j];
            stream(1, vc ? vv : -1);
        }
        // pairs v == a (the loop nest skips them): each kept element of N(a) once
        for (int32_t abase = 0; abase < Dn; abase += WAVE) {
            const int32_t av = a_of(abase + lane);
            if (av >= 0 && contains(C, Cn, av)) {
                const uint32_t l = a.vertex[av];
                const int32_t dg = (int32_t)(a.vertex[av + 1] - l);
                const gcptr na = E + l;
                uint64_t own = 0;
                if (gpf < 0) {
                    // the elements below b, less the placed vertices among them
                    const int32_t cut = b >= v_cnt ? dg : lower_bound_lane(a.edge + l, dg, b);
                    int32_t c = cut;
                    for (int j = 0; j < da; ++j)
                        if (s_emb[j] < b && contains(a.edge + l, cut, s_emb[j])) --c;
                    own = (uint64_t)(int64_t)c;
                } else {
                    for (int32_t i = 0; i < dg; ++i) {
                        const int32_t x = na[i];
                        if (x >= b) break;
                        bool keep = true;
                        for (int j = 0; j < da; ++j) keep = keep && x != s_emb[j];
                        if (keep) keep = contains(gptr, glen, x);
                        own += keep ? 1 : 0;
                    }
                }
                lacc -= own;
            }
        }
        __threadfence();
        for (int32_t abase = 0; abase < Dn; abase += WAVE) stream(2, a_of(abase + lane));
        __threadfence();
    };

    auto pair_level = [&]() {
        if (a.dense) {
            pair_level_dense();
            return;
        }
        const int da = L - 2, dv = L - 1;
        const int32_t astart = da == 2 ? dstart : 0, astep = da == 2 ? dstride : 1;
        const int32_t* D = s_lp[da] + astart;
        const int32_t Dn = s_ll[da] > astart ? (s_ll[da] - astart + astep - 1) / astep : 0;
        if (Dn <= 0) return;
        const int qc = pl->loop_set[dv];
        const int32_t* C = s_ptr[qc];
        int32_t Cn = s_len[qc];
        if (pl->rmask[dv]) Cn = wave_lower_bound(C, Cn, bound(pl->rmask[dv]));
        if (Cn <= 0) return;
        const int32_t b = bound(pl->rmask[L]);
        const int gpf = pl->pair_g;
        const int32_t* gptr = gpf >= 0 ? s_ptr[gpf] : nullptr;
        const int32_t glen = gpf >= 0 ? s_len[gpf] : 0;
        const int32_t pair_slots = a.hash_slots >> 1;
        const int pair_lg = a.hash_lg - 1;
        int32_t* keys = s_hash;
        int32_t* cnts = s_hash + pair_slots;
        const int32_t cap_elems = pair_slots / 2;
        bool halted = false; // a gather stopped at the plan's hard stop
        for (int32_t abase = 0; abase < Dn && !halted; abase += WAVE) {
            if (GP_PLAN_EXPIRED()) break;
            const int32_t ai = abase + lane;
            bool va = ai < Dn;
            const int32_t av = va ? D[(int64_t)ai * astep] : -1;
            for (int j = 0; j < da; ++j) va = va && av != s_emb[j];
            int32_t as = 0, acut = 0;
            if (va) {
                const uint32_t l = a.vertex[av];
                as = (int32_t)l;
                const int32_t dg = (int32_t)(a.vertex[av + 1] - l);
                acut = (b >= v_cnt || dg <= 16) ? dg : lower_bound_lane(a.edge + l, dg, b);
            }
            const bool big = va && acut > cap_elems;
            const int32_t sz = (va && !big) ? acut : 0;
            int32_t tot = 0;
            const int32_t ex = wave_excl_scan(sz, &tot);
            s_off[lane] = ex;
            __syncthreads();
            int start = 0;
            while (start < WAVE) {
                const int32_t base_e = __shfl(ex, start);
                const bool fits = lane >= start && (ex + sz - base_e) <= cap_elems;
                const uint64_t fm = __ballot(fits);
                int end = fm ? 64 - __clzll((long long)fm) : start + 1;
                if (end <= start) end = start + 1;
                const int32_t batch_elems = __shfl(ex + sz, end - 1) - base_e;
                if (batch_elems > 0) {
                    for (int i = lane; i < pair_slots; i += WAVE) {
                        keys[i] = -1;
                        cnts[i] = 0;
                    }
                    __syncthreads();
                    for (int32_t f0 = 0; f0 < batch_elems; f0 += WAVE) {
                        const int32_t f = f0 + lane;
                        const bool act = f < batch_elems;
                        int lo = start, hi = end;
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (s_off[mid] - base_e <= f) lo = mid; else hi = mid;
                        }
                        const int32_t os = __shfl(as, lo);
                        const int32_t oe = __shfl(ex, lo);
                        const int32_t x = act ? a.edge[os + (f - (oe - base_e))] : -1;
                        bool keep = act && x < b;
                        for (int j = 0; j < da; ++j) keep = keep && x != s_emb[j];
                        if (keep && gpf >= 0) keep = contains(gptr, glen, x);
                        if (keep) {
                            uint32_t h = hash_slot(x, pair_lg);
                            for (;;) {
                                const int32_t k = atomicCAS(&keys[h], -1, x);
                                if (k == -1 || k == x) {
                                    atomicAdd(&cnts[h], 1);
                                    break;
                                }
                                h = (h + 1) & (pair_slots - 1);
                            }
                        }
                    }
                    __syncthreads();
                    for (int32_t cbase = 0; cbase < Cn && !halted; cbase += WAVE) {
                        const int32_t ci = cbase + lane;
                        bool vc = ci < Cn;
                        const int32_t vv = vc ? C[ci] : -1;
                        for (int j = 0; j < da; ++j) vc = vc && vv != s_emb[j];
                        int32_t vs = 0, vcut = 0;
                        if (vc) {
                            const uint32_t l = a.vertex[vv];
                            vs = (int32_t)l;
                            const int32_t dg = (int32_t)(a.vertex[vv + 1] - l);
                            vcut = (b >= v_cnt || dg <= 16) ? dg : lower_bound_lane(a.edge + l, dg, b);
                        }
                        int32_t vt = 0;
                        const int32_t vex = wave_excl_scan(vcut, &vt);
                        s_off2[lane] = vex;
                        __syncthreads();
                        for (int32_t f0 = 0; f0 < vt; f0 += WAVE) {
                            if (((f0 / WAVE) & (GP_HALT_STEPS - 1)) == GP_HALT_STEPS - 1 && GP_PLAN_EXPIRED()) {
                                halted = true; // (flagged)
                                break;
                            }
                            const int32_t f = f0 + lane;
                            const bool act = f < vt;
                            int lo = 0, hi = WAVE;
                            while (hi - lo > 1) {
                                const int mid = (lo + hi) >> 1;
                                if (s_off2[mid] <= f) lo = mid; else hi = mid;
                            }
                            const int32_t os = __shfl(vs, lo);
                            const int32_t oe = __shfl(vex, lo);
                            if (act) {
                                const int32_t x = a.edge[os + (f - oe)];
                                if (x < b) {
                                    uint32_t h = hash_slot(x, pair_lg);
                                    for (;;) {
                                        const int32_t k = keys[h];
                                        if (k == x) {
                                            lacc += (uint64_t)cnts[h];
                                            break;
                                        }
                                        if (k == -1) break;
                                        h = (h + 1) & (pair_slots - 1);
                                    }
                                }
                            }
                        }
                        __syncthreads();
                    }
                    // pairs v == a
                    if (lane >= start && lane < end && va && !big && contains(C, Cn, av)) {
                        const int32_t* na = a.edge + as;
                        uint64_t own = 0;
                        for (int32_t i = 0; i < acut; ++i) {
                            const int32_t x = na[i];
                            bool keep = x < b;
                            for (int j = 0; j < da; ++j) keep = keep && x != s_emb[j];
                            if (keep && gpf >= 0) keep = contains(gptr, glen, x);
                            own += keep ? 1 : 0;
                        }
                        lacc -= own;
                    }
                    __syncthreads();
                }
                start = end;
            }
            // sets too large for the map
            uint64_t bm = __ballot(big);
            while (bm) {
                const int j = __ffsll((long long)bm) - 1;
                bm &= bm - 1;
                const int32_t aj = __shfl(av, j);
                s_emb[da] = aj;
                const uint32_t l = a.vertex[aj];
                if (build_level(da, a.edge + l, (int32_t)(a.vertex[aj + 1] - l))) {
                    enter(dv);
                    leaf_level(dv);
                }
            }
        }
    };

    // add this wave's sums for the current plan to its outputs and reset them
    // Pair level grouped by v1 (KPlan::pair_v1).  The tasks of one fetch
    // share v1 (edge slots read reversed) with v0 ascending.  The wave's dense
    // array holds cnt[x] = |{v3 in N(v1) ∩ [0, pv_hi-th element) : x in N(v3)}|
    // (x != v1), grown as v0 grows; each task streams only its depth-2 lists:
    //   sum_{a in D} sum_{x in N(a), x < b, x != v0} cnt[x]
    // minus the pairs v3 == a (a in D ∩ C), which the loop nest skips.  The
    // array is cleared when the fetch ends, so it is all zero between fetches.
    int32_t pv_v1 = -1; // v1 whose counts the array holds, -1 = none
    int32_t pv_hi = 0;  // N(v1)[0:pv_hi) scattered
    // every list N(u) of the 64 lanes' vertices u (u < 0: none), x != skip:
    // mode 0 adds 1 at cnt[x], mode 2 writes 0, mode 1 (lists cut at b)
    // adds cnt[x] to the lane sums
    auto pv_stream = [&](int mode, int32_t u, int32_t skip, int32_t b) {
        int32_t* const cnt = a.dense + (int64_t)blockIdx.x * a.dense_stride;
        const gcptr E = GL(a.edge);
        int32_t us = 0, un = 0;
        if (u >= 0) {
            const uint32_t l = a.vertex[u];
            us = (int32_t)l;
            const int32_t dg = (int32_t)(a.vertex[u + 1] - l);
            un = (mode != 1 || b >= v_cnt || dg < GP_LONG_LIST) ? dg : lower_bound_lane(a.edge + l, dg, b);
        }
        auto put = [&](int32_t x) {
            if (x < 0 || x == skip) return;
            if (mode == 0)
                atomicAdd(&cnt[x], 1);
            else if (mode == 2)
                cnt[x] = 0;
            else if (x < b)
                lacc += (uint64_t)__hip_atomic_load(&cnt[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        };
        const bool lng = un >= GP_LONG_LIST;
        int32_t tot = 0;
        const int32_t ex = wave_excl_scan(lng ? 0 : un, &tot);
        s_off2[lane] = ex;
        __syncthreads();
        for (int32_t f0 = 0; f0 < tot; f0 += WAVE) {
            const int32_t f = f0 + lane;
            const bool act = f < tot;
            int lo = 0, hi = WAVE;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off2[mid] <= f) lo = mid; else hi = mid;
            }
            const int32_t os = __shfl(us, lo);
            const int32_t oe = __shfl(ex, lo);
            put(act ? E[os + (f - oe)] : -1);
        }
        uint64_t lm = __ballot(lng);
        while (lm) {
            if (mode == 1 && GP_PLAN_EXPIRED()) break; // gathers only (flagged)
            const int j = __ffsll((long long)lm) - 1;
            lm &= lm - 1;
            const gcptr pv = E + __shfl(us, j);
            const int32_t pn = __shfl(un, j);
            for (int32_t f0 = 0; f0 < pn; f0 += GP_UNROLL * WAVE) {
                int32_t x[GP_UNROLL];
#pragma unroll
                for (int q = 0; q < GP_UNROLL; ++q) x[q] = f0 + q * WAVE + lane < pn ? pv[f0 + q * WAVE + lane] : -1;
#pragma unroll
                for (int q = 0; q < GP_UNROLL; ++q) put(x[q]);
            }
        }
        __syncthreads();
    };
    // N(pv_v1)[lo:hi): scatter (mode 0) or clear (mode 2) their lists
    auto pv_range = [&](int mode, int32_t lo, int32_t hi) {
        const gcptr C = GL(a.edge) + a.vertex[pv_v1];
        for (int32_t i0 = lo; i0 < hi; i0 += WAVE) {
            const int32_t i = i0 + lane;
            pv_stream(mode, i < hi ? C[i] : -1, pv_v1, 0);
        }
        __threadfence();
    };
    auto pv_clear = [&]() {
        if (pv_v1 < 0) return;
        pv_range(2, 0, pv_hi);
        pv_v1 = -1;
        pv_hi = 0;
    };
    auto pair_level_v1 = [&]() {
        const int32_t v0 = s_emb[0], v1 = s_emb[1];
        const uint32_t l1 = a.vertex[v1];
        const int32_t hi = wave_lower_bound(a.edge + l1, (int32_t)(a.vertex[v1 + 1] - l1), v0);
        if (pv_v1 != v1 || hi < pv_hi) {
            pv_clear();
            pv_v1 = v1;
            pv_hi = 0;
        }
        if (hi > pv_hi) {
            pv_range(0, pv_hi, hi);
            pv_hi = hi;
        }
        const int32_t* D = s_lp[2];
        const int32_t Dn = s_ll[2];
        if (Dn <= 0 || pv_hi <= 0) return;
        const int32_t b = bound(pl->rmask[L]);
        const int32_t* C = a.edge + l1;
        for (int32_t abase = 0; abase < Dn; abase += WAVE) {
            if (GP_PLAN_EXPIRED()) break; // (flagged)
            const int32_t ai = abase + lane;
            int32_t av = ai < Dn ? GL(D)[ai] : -1;
            if (av == v0 || av == v1) av = -1;
            pv_stream(1, av, v0, b);
            // the pair v3 == av: every kept element of N(av), i.e. those
            // below b other than v0 and v1
            if (av >= 0 && contains(C, pv_hi, av)) {
                const uint32_t l = a.vertex[av];
                const int32_t dg = (int32_t)(a.vertex[av + 1] - l);
                const int32_t* na = a.edge + l;
                const int32_t cut = b >= v_cnt ? dg : lower_bound_lane(na, dg, b);
                int32_t own = cut;
                if (v0 < b && contains(na, cut, v0)) --own;
                if (v1 < b && contains(na, cut, v1)) --own;
                lacc -= (uint64_t)(int64_t)own;
            }
        }
    };

    auto flush = [&]() {
        const uint64_t lsum = wave_sum_u64(lacc);
        if (lane == 0) {
            unsigned long long* o = a.out + 3 * cur_plan;
            if (acc + lsum) atomicAdd(&o[0], (unsigned long long)(acc + lsum));
            if (FAITHFUL && balg) atomicAdd(&o[1], (unsigned long long)balg);
            if (taken) atomicAdd(&o[2], (unsigned long long)taken);
        }
        acc = lacc = balg = taken = 0;
    };

    // Work loop.  First launch (PARTS = false): chunks of tasks, deferring
    // heavy ones.  Second launch (PARTS = true): parts of the deferred tasks,
    // one per fetch.
    const uint64_t n_items = !PARTS ? a.n_tasks : min((uint64_t)*a.heavy_cnt, a.heavy_cap) * GP_SPLIT_PARTS;
    // [cur, cend): tasks of the current plan still to run from this fetch;
    // [cur, fend): the whole fetch, which may run on into the next plan's items.
    // The fetch size adapts per wave (a.fetch_ticks > 0): the first fetch
    // takes one item, each later one as many as the last fetch's time per
    // item says fit in fetch_ticks, at most a.chunk.  Cheap tasks then come
    // a.chunk at a time (few atomics on the shared counter), expensive ones
    // one at a time (no wave holds a run of expensive tasks at the end).
    // (the fetch's end, size and start time live in LDS: they are touched
    // once per fetch, and registers held across the whole loop spill)
    // With a.deadline_ticks > 0 a wave stops fetching once that many ticks
    // have passed since the launch's first wave started (s_torigin).  The
    // grid can hold more waves than fit on the device at once (LDS bounds
    // residency), and a wave that starts only when another exits then finds
    // the deadline passed: it fetches nothing once any wave has fetched (so
    // every launch makes progress), instead of running a deadline of its own
    // after the first ones, which made launches last two deadlines.  Items
    // are handed out in order from one counter and a fetched item is always
    // finished, so the launch then has counted exactly the items below the
    // counter's final value: a prefix, which the host resumes from.
    __shared__ uint64_t s_fend, s_tfetch, s_tstart;
    __shared__ int32_t s_chunk;
    if (lane == 0) {
        s_fend = 0;
        s_tfetch = 0;
        s_tstart = wall_clock64();
        s_chunk = a.fetch_ticks > 0 ? 1 : a.chunk;
        // plans' hard stops count from the start of the call's first launch:
        // its first wave to start records the time, the parts launch reads it
        s_torigin = s_tstart;
        s_ttask = s_tstart;
        if (!PARTS) {
            const unsigned long long old = atomicCAS(a.origin, 0ull, (unsigned long long)s_tstart);
            if (old != 0) s_torigin = old;
        } else {
            const unsigned long long o = __hip_atomic_load(a.origin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o != 0) s_torigin = o;
        }
    }
    __syncthreads();
    uint64_t cur = 0, cend = 0;
    for (;;) {
        uint64_t t;
        if (!PARTS) {
            bool done = false;
            while (cur >= cend) {
                if (cur >= first_lane_u64(s_fend)) {
                    if (EXT == GP_EXT_PAIR) pv_clear(); // the dense array is zero between fetches
                    if (a.deadline_ticks > 0 &&
                        first_lane_u64(wall_clock64()) - first_lane_u64(s_torigin) > (uint64_t)a.deadline_ticks &&
                        (first_lane_u64(s_fend) > 0 ||
                         first_lane_u64(__hip_atomic_load(a.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) > 0)) {
                        done = true;
                        break;
                    }
                    int32_t chunk_now = __builtin_amdgcn_readfirstlane(s_chunk);
                    if (a.fetch_ticks > 0 && first_lane_u64(s_fend) > 0) {
                        const uint64_t el = first_lane_u64(wall_clock64()) - first_lane_u64(s_tfetch);
                        const uint64_t per = max(el / (uint64_t)chunk_now, (uint64_t)1);
                        chunk_now = (int32_t)min((uint64_t)a.chunk, max((uint64_t)a.fetch_ticks / per, (uint64_t)1));
                        chunk_now = __builtin_amdgcn_readfirstlane(chunk_now);
                    }
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd(a.counter, (unsigned long long)chunk_now);
                    base = first_lane_u64(base);
                    if (base >= n_items) {
                        done = true;
                        break;
                    }
                    __syncthreads(); // every lane has read the old fetch state
                    if (lane == 0) {
                        s_chunk = chunk_now;
                        s_fend = min((uint64_t)base + (uint64_t)chunk_now, n_items);
                        s_tfetch = wall_clock64();
                        s_ttask = s_tfetch; // the per-task limit counts per fetch
                    }
                    __syncthreads();
                    cur = base;
                }
                // the plan owning item cur
                int p = 0;
                if (MULTI) {
                    while (p + 1 < a.nplans && a.task_off[p + 1] <= cur) ++p;
                    if (p != cur_plan) {
                        flush();
                        set_plan(p);
                    }
                }
                const uint64_t pend = a.task_off[p] + a.task_cnt[p];
                const uint64_t fend = first_lane_u64(s_fend);
                if (cur < pend) {
                    cend = min(fend, pend);
                } else { // padding after the plan's tasks: on to the next plan's items
                    cur = min(fend, (uint64_t)a.task_off[p + 1]);
                    cend = cur;
                }
            }
            if (done) break;
            t = a.task_first[cur_plan] + (cur - a.task_off[cur_plan]);
            ++cur;
            ++taken;
        } else {
            unsigned long long w = 0;
            __syncthreads(); // every lane is done with the last part's start time
            if (lane == 0) {
                w = atomicAdd(a.counter2, 1ull);
                s_ttask = wall_clock64();
            }
            __syncthreads();
            w = first_lane_u64(w);
            if (w >= n_items) break;
            const unsigned long long e = a.heavy[w / GP_SPLIT_PARTS];
            const int32_t part = (int32_t)(w % GP_SPLIT_PARTS);
            const int32_t parts = max(2, min(GP_SPLIT_PARTS, (int32_t)(e & 0xFFFFFF) / GP_SPLIT_GRAIN));
            if (part >= parts) continue;
            if (MULTI) {
                const int p = (int)(e >> 56);
                if (p != cur_plan) {
                    flush();
                    set_plan(p);
                }
            }
            dstart = part;
            dstride = parts;
            t = (e >> 24) & 0xFFFFFFFFull;
        }
        if (GP_PLAN_EXPIRED()) { // (flagged)
            if (!PARTS) {
                // its results are discarded: the rest of its items are
                // skipped at once, here and for every wave still to fetch
                // (the items of the plans after it keep their order, so
                // their counted prefixes stay exact)
                if (lane == 0) atomicMax(a.counter, (unsigned long long)a.task_off[cur_plan + 1]);
                cur = cend;
            } else if (!MULTI) {
                // every part is the one plan's: skip them all at once (the
                // deferred list may hold 2^20 tasks of 64 parts each)
                if (lane == 0) atomicMax(a.counter2, (unsigned long long)n_items);
                break;
            }
            continue;
        }
        // the task: depth 0 and 1 from the edge (v0, v1), then depth 2 on
        int32_t v0, v1 = -1;
        if (pl->root_mode) {
            v0 = (int32_t)t;
        } else {
            v0 = a.src[t];
            v1 = a.edge[t];
            if (pl->pair_v1) { // the slot read reversed: its row is v1
                v1 = v0;
                v0 = a.edge[t];
            }
            if (v1 == v0) continue;
            if ((pl->rmask[1] & 1) && v1 >= v0) continue;
        }
        s_emb[0] = v0;
        const uint32_t l0 = a.vertex[v0];
        const int32_t* n0 = a.edge + l0;
        const int32_t d0 = (int32_t)(a.vertex[v0 + 1] - l0);
        for (int i = 0; i < pl->nbuild[0]; ++i) build_prefix(pl->build[0][i], n0, d0);
        if (pl->root_mode) {
            if (L == 1) {
                acc += leaf();
            } else {
                enter(1);
                leaf_level(1);
            }
            continue;
        }
        s_emb[1] = v1;
        if (FAITHFUL) balg += 8;
        const uint32_t l1 = a.vertex[v1];
        const int32_t* n1 = a.edge + l1;
        const int32_t d1 = (int32_t)(a.vertex[v1 + 1] - l1);
        if (!build_level(1, n1, d1)) continue;
        if (L == 2) {
            acc += leaf();
            continue;
        }
        enter(2);
        if (!PARTS && !FAITHFUL && s_ll[2] > pl->split_min) {
            unsigned long long slot = 0;
            if (lane == 0) slot = atomicAdd(a.heavy_cnt, 1ull);
            slot = __shfl(slot, 0);
            if (slot < a.heavy_cap) {
                if (lane == 0)
                    a.heavy[slot] = (unsigned long long)cur_plan << 56 | (unsigned long long)t << 24 |
                                    (unsigned long long)min(s_ll[2], 0xFFFFFF);
                continue;
            }
        }
        if (pl->pair && L == 4) {
            if (EXT == GP_EXT_PAIR && !PARTS && pl->pair_v1 && a.dense)
                pair_level_v1();
            else
                pair_level();
            continue;
        }
        if (pl->batched && L == 3) {
            leaf_level(2);
            continue;
        }
        // depth-first walk over depths 2 .. L-1 (L-2 when batched)
        int d = 2;
        for (;;) {
            if (GP_PLAN_EXPIRED()) break; // (flagged)
            const int32_t idx = s_li[d];
            if (idx >= s_ll[d]) {
                if (d == 2) break;
                --d;
                continue;
            }
            s_li[d] = idx + (d == 2 ? dstride : 1);
            const int32_t v = s_lp[d][idx];
            bool dup = false;
            for (int j = 0; j < d; ++j) dup = dup || s_emb[j] == v;
            if (dup) continue;
            s_emb[d] = v;
            if (FAITHFUL) balg += 8;
            const uint32_t lv = a.vertex[v];
            const int32_t* nv = a.edge + lv;
            const int32_t dv = (int32_t)(a.vertex[v + 1] - lv);
            if (!build_level(d, nv, dv)) continue;
            if (pl->pair && d == L - 3) {
                enter(L - 2);
                pair_level();
                continue;
            }
            if (pl->batched && d == L - 2) {
                enter(L - 1);
                leaf_level(L - 1);
                continue;
            }
            if (!pl->batched && d == L - 1) {
                acc += leaf();
                continue;
            }
            ++d;
            enter(d);
        }
    }
#undef GP_PLAN_EXPIRED
    flush();
    if (lane == 0) a.clocks[2 * blockIdx.x + 1] = wall_clock64();
}

// src[e] = the vertex whose list holds edge slot e
__global__ void build_src_kernel(const uint32_t* __restrict__ vertex, int32_t v_cnt, uint32_t e_cnt,
                                 int32_t* __restrict__ src) {
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e_cnt) return;
    int32_t lo = 0, hi = v_cnt; // last v with vertex[v] <= e
    while (hi - lo > 1) {
        int32_t mid = lo + (hi - lo) / 2;
        if (vertex[mid] <= (uint32_t)e)
            lo = mid;
        else
            hi = mid;
    }
    src[e] = lo;
}

__global__ void zero_i32_kernel(int32_t* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void fill_u64_kernel(unsigned long long* p, int n) {
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) p[i] = 0;
}

} // namespace

// ---------------------------------------------------------------------------
// runtime
// ---------------------------------------------------------------------------

struct CallCtx {
    unsigned long long* d_clocks = nullptr;
    size_t clocks_cap = 0; // waves
    // plans, then task_off[nplans + 1] and task_first[nplans] (uint64)
    char* d_meta = nullptr;
    size_t meta_cap = 0;
    // [0] task counter, [1] part counter, [2] deferred tasks, then 3 outputs per plan
    unsigned long long* d_cnt = nullptr;
    size_t cnt_cap = 0;
    unsigned long long* d_heavy = nullptr; // GP_HEAVY_CAP deferred tasks
    int32_t* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    int32_t* d_dense = nullptr; // pair level counts, all zero between calls
    size_t dense_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // graph mode: pinned staging for the plan upload and the read-back, and
    // the instantiated graph of the last launch shape (see GraphKey)
    char* h_meta = nullptr;
    size_t h_meta_cap = 0;
    unsigned long long* h_out = nullptr; // counters (cnt_cap entries), then 4 clock stamps per wave
    size_t h_out_cap = 0;                // entries
    struct GraphEntry {
        std::vector<char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipGraphNode_t count = nullptr, parts = nullptr;
    };
    std::vector<GraphEntry> graphs; // most recently used last, at most GP_GRAPH_CACHE
    std::vector<std::vector<char> > seen; // launch shapes seen once (not yet instantiated)
};

struct gp_graph {
    int device = 0;
    int num_cu = 0;
    int wall_khz = 100000; // wall-clock ticks per millisecond (wall_clock64 in kernels)
    int32_t v_cnt = 0;
    uint32_t e_cnt = 0;
    int32_t max_deg = 0, max_deg2 = 0;
    uint32_t* d_vertex = nullptr;
    int32_t* d_edge = nullptr;
    int32_t* d_src = nullptr;
    int32_t* d_sup = nullptr;   // per edge slot |N(src) ∩ N(dst)|, built on first use
    double sup_build_ms = 0;    // time of that build
    uint64_t* d_vtri = nullptr; // per vertex: triangles containing it, built on first use
    std::mutex mu;
    std::mutex sup_mu;
    std::vector<CallCtx*> idle;
};

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t err_ = (expr);                                                                \
        if (err_ != hipSuccess)                                                                  \
            return gp_fail((int)err_, std::string(#expr) + ": " + hipGetErrorString(err_));      \
    } while (0)

namespace {

void destroy_ctx(CallCtx* c) {
    if (!c) return;
    if (c->d_meta) (void)hipFree(c->d_meta);
    if (c->d_clocks) (void)hipFree(c->d_clocks);
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    if (c->d_heavy) (void)hipFree(c->d_heavy);
    if (c->d_scratch) (void)hipFree(c->d_scratch);
    if (c->d_dense) (void)hipFree(c->d_dense);
    for (CallCtx::GraphEntry& ge : c->graphs) {
        if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
        if (ge.graph) (void)hipGraphDestroy(ge.graph);
    }
    if (c->h_meta) (void)hipHostFree(c->h_meta);
    if (c->h_out) (void)hipHostFree(c->h_out);
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

hipError_t ensure(void** p, size_t* cap, size_t bytes, hipStream_t stream);

int acquire_ctx(gp_graph* g, size_t scratch_bytes, CallCtx** out) {
    CallCtx* c = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (!g->idle.empty()) {
            c = g->idle.back();
            g->idle.pop_back();
        }
    }
    if (!c) {
        c = new CallCtx();
        hipError_t e = hipSuccess;
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipMallocAsync((void**)&c->d_heavy, GP_HEAVY_CAP * sizeof(unsigned long long), c->stream);
        if (e == hipSuccess) e = hipEventCreate(&c->e0);
        if (e == hipSuccess) e = hipEventCreate(&c->e1);
        if (e != hipSuccess) {
            destroy_ctx(c);
            return gp_fail((int)e, std::string("call context: ") + hipGetErrorString(e));
        }
    }
    if (c->scratch_bytes < scratch_bytes) {
        size_t cap = c->scratch_bytes;
        hipError_t e = ensure((void**)&c->d_scratch, &cap, scratch_bytes, c->stream);
        if (e != hipSuccess) {
            destroy_ctx(c);
            return gp_fail(GP_ENOMEM, std::string("scratch allocation of ") + std::to_string(scratch_bytes) +
                                          " bytes failed: " + hipGetErrorString(e));
        }
        c->scratch_bytes = scratch_bytes;
    }
    *out = c;
    return GP_OK;
}

void release_ctx(gp_graph* g, CallCtx* c) {
    std::lock_guard<std::mutex> lk(g->mu);
    g->idle.push_back(c);
}

// grow a device buffer to at least `bytes`, stream-ordered on the call's
// stream: hipFree and hipMalloc would wait for the whole device, so one
// call growing its buffers would stall every other call in flight (the
// serving loop runs several at once)
hipError_t ensure(void** p, size_t* cap, size_t bytes, hipStream_t stream) {
    if (*cap >= bytes) return hipSuccess;
    if (*p) (void)hipFreeAsync(*p, stream);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMallocAsync(p, bytes, stream);
    if (e == hipErrorOutOfMemory) {
        // the pool keeps what calls freed (smaller blocks that a larger
        // request cannot use): hand it back to the driver and try once more
        (void)hipGetLastError();
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipStreamSynchronize(stream) == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
            (void)hipMemPoolTrimTo(pool, 0);
            e = hipMallocAsync(p, bytes, stream);
        }
    }
    if (e == hipSuccess) *cap = bytes;
    return e;
}

// The launch of one counting call: kernel plans, work layout and resources.
struct Launch {
    bool faithful = false, any_split = false;
    std::vector<KPlan> kps;
    std::vector<uint64_t> task_off, task_first, task_cnt;
    uint64_t n_tasks = 0;
    int64_t waves = 0;
    int chunk = 1;            // tasks per fetch, or the most a fetch takes when fetch_ticks > 0
    int64_t fetch_ticks = 0;  // adaptive fetch size: ticks a fetch aims at (0: fixed chunk)
    int64_t slot_stride = 0, tmp_stride = 0, bslot_stride = 0, wave_stride = 0;
    size_t scratch_bytes = 0;
    int hash_slots = 0, hash_lg = 0;
    size_t lds_bytes = 0;
    int64_t dense_stride = 0; // ints per wave of the pair level's count array (0: LDS map)
    size_t dense_bytes = 0;
    // device memory the call needs besides the graph: scratch, dense pair
    // counts, plans and work layout, counters, clock stamps and the
    // deferred-task list
    size_t device_bytes() const {
        const size_t n = kps.size();
        return scratch_bytes + dense_bytes + sizeof(KPlan) * n + sizeof(uint64_t) * (3 * n + 1) +
               sizeof(unsigned long long) * (3 + 3 * n + 4 * (size_t)waves) + sizeof(unsigned long long) * GP_HEAVY_CAP;
    }
};

// rb, re (optional): plan p counts tasks [rb[p], re[p]) instead of the
// tuning's range (which applies to a single plan only)
int prepare(const gp_graph* g, const gp_sched_flat* scheds, int nplans, const gp_tuning* tuning, Launch* out,
            const uint64_t* rb = nullptr, const uint64_t* re = nullptr) {
    if (nplans < 1 || nplans > GP_MAX_BATCH) return gp_fail(GP_ERANGE, "schedule count out of range");
    const bool faithful = tuning && tuning->collect_balg;
    std::vector<KPlan> kps((size_t)nplans);
    const int split_req = tuning ? tuning->split_min : 0;
    bool any_split = false, any_batched = false, any_k = false;
    int hash_slots = 0, bc_rows = 0, nbslots = 0;
    int max_slots = 0;
    for (int p = 0; p < nplans; ++p) {
        KPlan& kp = kps[p];
        int rc = gp_make_kplan(&scheds[p], faithful, !(tuning && tuning->no_batch), !(tuning && tuning->no_pair), &kp);
        if (rc != GP_OK) return rc;
        if (tuning && tuning->no_truncate) {
            std::memset(kp.ub_n, 0, sizeof(kp.ub_n));
            kp.cut_done = 0; // no loop set is cut when it is built
        }
        // the flagged shared search pays off only on graphs with large
        // neighbour sets (measured: slower on R-MAT s10/s12 with max degree
        // 467/1,319 by 27-35 %, faster on s18 with 15,992); below the
        // threshold its plans stay on the plain kernel
        if ((tuning && tuning->no_shared_search) || g->max_deg < GP_SHARED_MIN_DEG) kp.bl_base = -1;
        // grouping by v1 needs the single-plan launch's 256-task fetches
        if ((tuning && tuning->no_pair_v1) || nplans > 1) kp.pair_v1 = 0;
        // heavy-task splitting only where a depth-2 loop exists
        const bool split = !faithful && !kp.root_mode && kp.leaf >= 3 && split_req >= 0;
        kp.split_min = split ? (split_req > 0 ? split_req : kp.pair ? GP_SPLIT_DEFAULT_PAIR : GP_SPLIT_DEFAULT) : INT_MAX;
        // the pair level grouped by v1 keeps its tasks in place by default: a
        // part would re-stream the whole N(v1) side, which grouping avoids
        if (kp.pair_v1 && split_req == 0) kp.split_min = INT_MAX;
        any_split = any_split || split;
        any_batched = any_batched || kp.batched;
        any_k = any_k || kp.k;
        if (kp.pair) hash_slots = std::max(hash_slots, GP_HASH_SLOTS_PAIR);
        if (kp.batched) {
            hash_slots = std::max(hash_slots, GP_HASH_SLOTS_BATCH);
            bc_rows = std::max(bc_rows, kp.nblocks);
            nbslots = std::max(nbslots, kp.nbslots);
        }
        max_slots = std::max(max_slots, kp.nslots);
    }
    const int wpc = tuning && tuning->waves_per_cu > 0 ? tuning->waves_per_cu : 24;
    int64_t waves = (int64_t)g->num_cu * wpc;
    // work items: each plan's task range, concatenated
    std::vector<uint64_t> task_off((size_t)nplans + 1, 0), task_first((size_t)nplans, 0), task_cnt((size_t)nplans, 0);
    for (int p = 0; p < nplans; ++p) {
        const uint64_t all_tasks = kps[p].root_mode ? (uint64_t)g->v_cnt : (uint64_t)g->e_cnt;
        uint64_t t_begin = 0, t_end = all_tasks;
        if (rb && re) {
            t_begin = std::min(rb[p], all_tasks);
            t_end = std::min(re[p], all_tasks);
        } else if (tuning && (tuning->task_begin != 0 || tuning->task_end > 0)) {
            if (nplans != 1) return gp_fail(GP_EINVAL, "task ranges apply to a single schedule");
            if (tuning->task_begin < 0 || (tuning->task_end > 0 && tuning->task_end < tuning->task_begin))
                return gp_fail(GP_EINVAL, "bad task range");
            t_begin = std::min<uint64_t>((uint64_t)tuning->task_begin, all_tasks);
            if (tuning->task_end > 0) t_end = std::min<uint64_t>((uint64_t)tuning->task_end, all_tasks);
        }
        task_first[p] = t_begin;
        task_cnt[p] = t_end > t_begin ? t_end - t_begin : 0;
    }
    uint64_t real_tasks = 0;
    for (int p = 0; p < nplans; ++p) real_tasks += task_cnt[p];
    // tasks per work fetch: at most 16, fewer when a plan has under 32
    // fetches of 16 per wave (small graphs: the last fetches then decide the
    // kernel's length); within that cap each wave sizes its fetches for about
    // GP_FETCH_US of work from the time its last fetch took.  Measured on
    // R-MAT s22: the 4-clique (about 30 us of a wave per task) took 317 ms in
    // fetches of 16 and 847 ms in fetches of 1 (6,144 waves queue on one
    // atomic counter); the 5-clique on the Chung-Lu 3M graph (about 17 ms
    // per task) took 10.3 s on a slice in fetches of 1 and 12.9 s in fetches
    // of 16; all 133 5- and 6-vertex motifs as one batch on R-MAT s12 took
    // 11.5 s in fetches of 1 and 23.4 s in fetches of 16.
    // (512 for a single plan whose pair level is grouped by v1: its counts
    // carry over between the tasks of one fetch; measured on R-MAT s16-s20,
    // 256 beat 64 and the default by 1.2-6x despite the coarser balance, and
    // 512 beat 256 by 6 % at s20)
    const bool grouped = nplans == 1 && kps[0].pair_v1;
    const uint64_t per_plan = real_tasks / (uint64_t)nplans;
    const int chunk = tuning && tuning->chunk > 0
                          ? tuning->chunk
                          : grouped ? 512
                                    : (int)std::max<uint64_t>(1, std::min<uint64_t>(16, per_plan / (uint64_t)(waves * 32)));
    const int64_t fetch_ticks =
        (tuning && tuning->chunk > 0) || grouped ? 0 : (int64_t)g->wall_khz * GP_FETCH_US / 1000;
    // each plan's items start on a chunk boundary (a fetch of a fixed chunk
    // then never spans two plans; an adaptive fetch may, and the kernel walks
    // on into the next plan's items)
    for (int p = 0; p < nplans; ++p)
        task_off[p + 1] = task_off[p] + (task_cnt[p] + (uint64_t)chunk - 1) / (uint64_t)chunk * (uint64_t)chunk;
    const uint64_t n_tasks = task_off[nplans];
    if (waves < 1) waves = 1;
    // (graph mode keeps the full grid, so that launch shapes, and with them
    // the cached graphs, stay the same whatever the task count)
    if (!(tuning && tuning->use_graph) && (uint64_t)waves > n_tasks / (uint64_t)chunk + 1)
        waves = (int64_t)(n_tasks / (uint64_t)chunk + 1);
    // scratch per wave: one slot per materialised prefix (a subset of one
    // neighbour list of a vertex other than the root, so at most the
    // second-largest degree), two IEP temporaries in faithful mode, and the
    // batched level's block sets (each bounded by a neighbour list)
    const int64_t slot_stride = ((int64_t)std::max(g->max_deg2, 1) + 63) / 64 * 64;
    const int64_t deg_stride = ((int64_t)std::max(g->max_deg, 1) + 63) / 64 * 64;
    const int64_t tmp_stride = faithful && any_k ? deg_stride : 0;
    const int64_t bslot_stride = any_batched && nbslots ? deg_stride : 0;
    const int64_t wave_stride = max_slots * slot_stride + 2 * tmp_stride + (nbslots + 1) * bslot_stride;
    // per-wave scratch is bounded by GP_SCRATCH_CAP over all waves: a plan
    // with many materialised sets on a graph with huge second-largest
    // degree (7 vertices on R-MAT s24 ef16: 12 MB a wave) runs on fewer
    // waves rather than claiming most of the device
    if (wave_stride > 0 && (size_t)waves * (size_t)wave_stride * sizeof(int32_t) > GP_SCRATCH_CAP)
        waves = std::max<int64_t>(std::min<int64_t>(waves, g->num_cu),
                                  (int64_t)(GP_SCRATCH_CAP / ((size_t)wave_stride * sizeof(int32_t))));
    const size_t scratch_bytes = (size_t)std::max<int64_t>(waves * wave_stride, 64) * sizeof(int32_t);
    // Several plans share one LDS size; the larger pair-level map would cut
    // residency for all of them, so a batch gives the pair level the batched
    // level's pool (its sub-batches get smaller instead).
    if (nplans > 1 && hash_slots > GP_HASH_SLOTS_BATCH) hash_slots = GP_HASH_SLOTS_BATCH;
    // LDS hash pool: tuning->lds_stage elements staged at load factor <= 1/2
    if (tuning && tuning->lds_stage > 0 && hash_slots > 0) {
        hash_slots = 64;
        while (hash_slots < 2 * tuning->lds_stage && hash_slots < GP_HASH_SLOTS_MAX) hash_slots <<= 1;
    }
    int hash_lg = 0;
    while ((1 << hash_lg) < hash_slots) ++hash_lg;
    const size_t lds_bytes = sizeof(int32_t) * ((size_t)hash_slots + (size_t)bc_rows * WAVE);
    // pair level: a dense count array per wave when every wave's fits the cap
    bool any_pair = false;
    for (const KPlan& kp : kps) any_pair = any_pair || kp.pair;
    int64_t dense_stride = 0;
    size_t dense_bytes = 0;
    if (any_pair && !(tuning && tuning->pair_map == 1)) {
        const int64_t stride = ((int64_t)std::max(g->v_cnt, 1) + 63) / 64 * 64;
        const size_t bytes = (size_t)waves * (size_t)stride * sizeof(int32_t);
        if (bytes <= GP_DENSE_MAX_BYTES) {
            dense_stride = stride;
            dense_bytes = bytes;
        }
    }

    out->faithful = faithful;
    out->any_split = any_split;
    out->kps = std::move(kps);
    out->task_off = std::move(task_off);
    out->task_first = std::move(task_first);
    out->task_cnt = std::move(task_cnt);
    out->n_tasks = n_tasks;
    out->waves = waves;
    out->chunk = chunk;
    out->fetch_ticks = fetch_ticks;
    out->slot_stride = slot_stride;
    out->tmp_stride = tmp_stride;
    out->bslot_stride = bslot_stride;
    out->wave_stride = wave_stride;
    out->scratch_bytes = scratch_bytes;
    out->hash_slots = hash_slots;
    out->hash_lg = hash_lg;
    out->lds_bytes = lds_bytes;
    out->dense_stride = dense_stride;
    out->dense_bytes = dense_bytes;
    return GP_OK;
}

// The count kernel instantiation for a launch (faithful runs have no parts
// launch and never take the extended paths).
#define GP_KFN(...) reinterpret_cast<const void*>(&gp_count_kernel<__VA_ARGS__>)
template <bool HALT>
const void* count_kernel_pick(bool parts, bool multi, int ext) {
    // multi-plan launches have no extended instantiation
    if (parts) {
        if (multi) return GP_KFN(false, true, true, GP_EXT_NONE, HALT);
        if (ext == GP_EXT_SEARCH) return GP_KFN(false, true, false, GP_EXT_SEARCH, HALT);
        if (ext == GP_EXT_PAIR) return GP_KFN(false, true, false, GP_EXT_PAIR, HALT);
        if (ext == GP_EXT_MERGE) return GP_KFN(false, true, false, GP_EXT_MERGE, HALT);
        return GP_KFN(false, true, false, GP_EXT_NONE, HALT);
    }
    if (multi) return GP_KFN(false, false, true, GP_EXT_NONE, HALT);
    if (ext == GP_EXT_SEARCH) return GP_KFN(false, false, false, GP_EXT_SEARCH, HALT);
    if (ext == GP_EXT_PAIR) return GP_KFN(false, false, false, GP_EXT_PAIR, HALT);
    if (ext == GP_EXT_MERGE) return GP_KFN(false, false, false, GP_EXT_MERGE, HALT);
    return GP_KFN(false, false, false, GP_EXT_NONE, HALT);
}

// halt: some plan has a hard stop (the faithful loop nest never has one)
const void* count_kernel_fn(bool faithful, bool parts, bool multi, int ext, bool halt) {
    if (faithful)
        return multi ? GP_KFN(true, false, true, GP_EXT_NONE, false) : GP_KFN(true, false, false, GP_EXT_NONE, false);
    return halt ? count_kernel_pick<true>(parts, multi, ext) : count_kernel_pick<false>(parts, multi, ext);
}
#undef GP_KFN

// Graph mode of one counting call.  The graph holds, in order: the plan
// upload from pinned staging (the whole meta buffer), the counter reset (a
// fill kernel: no memset nodes), the count launch, the deferred-parts launch
// when there is one, and the read-back of the counters and clock stamps into
// pinned staging.  It is built once per launch shape (kernel variants, grid,
// LDS size and the call context's buffers) and replayed; between replays only
// the two count kernels' arguments change (hipGraphExecKernelNodeSetParams).
struct GraphKey {
    const void* kfun;
    const void* pfun;
    unsigned grid;
    size_t lds;
    void* d_meta;
    size_t meta_cap;
    void* d_cnt;
    size_t cnt_cap;
    void* d_clocks;
    size_t n_clk;
    void* h_meta;
    void* h_out;
};

// A shape is instantiated the second time it is seen: one-off shapes run as
// stream launches (*launched = false) and pay no instantiation.
hipError_t graph_run(CallCtx* c, const void* kfun, const void* pfun, unsigned grid, size_t lds, KArgs& a, KArgs& a2,
                     const std::vector<char>& meta, size_t n_cnt, size_t n_clk, unsigned long long* host_out,
                     unsigned long long* clocks, float* ms, bool* launched) {
    *launched = false;
    hipError_t e = hipSuccess;
    if (c->h_meta_cap < c->meta_cap) {
        if (c->h_meta) (void)hipHostFree(c->h_meta);
        c->h_meta = nullptr;
        c->h_meta_cap = 0;
        e = hipHostMalloc((void**)&c->h_meta, c->meta_cap, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        c->h_meta_cap = c->meta_cap;
    }
    const size_t cnt_entries = c->cnt_cap / sizeof(unsigned long long);
    const size_t out_need = cnt_entries + n_clk;
    if (c->h_out_cap < out_need) {
        if (c->h_out) (void)hipHostFree(c->h_out);
        c->h_out = nullptr;
        c->h_out_cap = 0;
        e = hipHostMalloc((void**)&c->h_out, out_need * sizeof(unsigned long long), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        c->h_out_cap = out_need;
    }
    GraphKey key;
    std::memset(&key, 0, sizeof(key)); // compared bytewise, padding included
    key.kfun = kfun;
    key.pfun = pfun;
    key.grid = grid;
    key.lds = lds;
    key.d_meta = c->d_meta;
    key.meta_cap = c->meta_cap;
    key.d_cnt = c->d_cnt;
    key.cnt_cap = c->cnt_cap;
    key.d_clocks = c->d_clocks;
    key.n_clk = n_clk;
    key.h_meta = c->h_meta;
    key.h_out = c->h_out;
    const char* kb = reinterpret_cast<const char*>(&key);
    void* kargs[] = {&a};
    void* pargs[] = {&a2};
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(kfun);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(WAVE);
    kp.sharedMemBytes = (unsigned)lds;
    kp.kernelParams = kargs;
    kp.extra = nullptr;
    hipKernelNodeParams pp = kp;
    pp.func = const_cast<void*>(pfun);
    pp.kernelParams = pargs;
    int hit = -1;
    for (size_t i = 0; i < c->graphs.size(); ++i)
        if (c->graphs[i].key.size() == sizeof(key) && std::memcmp(c->graphs[i].key.data(), kb, sizeof(key)) == 0)
            hit = (int)i;
    if (hit >= 0) {
        // most recently used goes last
        std::rotate(c->graphs.begin() + hit, c->graphs.begin() + hit + 1, c->graphs.end());
        CallCtx::GraphEntry& ge = c->graphs.back();
        e = hipGraphExecKernelNodeSetParams(ge.exec, ge.count, &kp);
        if (e == hipSuccess && pfun) e = hipGraphExecKernelNodeSetParams(ge.exec, ge.parts, &pp);
        if (e != hipSuccess) return e;
    } else {
        bool again = false;
        for (const std::vector<char>& k : c->seen)
            if (k.size() == sizeof(key) && std::memcmp(k.data(), kb, sizeof(key)) == 0) again = true;
        if (!again) {
            if (c->seen.size() >= 4 * GP_GRAPH_CACHE) c->seen.erase(c->seen.begin());
            c->seen.push_back(std::vector<char>(kb, kb + sizeof(key)));
            return hipSuccess; // first sighting: the caller launches on the stream
        }
        if (c->graphs.size() >= GP_GRAPH_CACHE) {
            CallCtx::GraphEntry& old = c->graphs.front();
            if (old.exec) (void)hipGraphExecDestroy(old.exec);
            if (old.graph) (void)hipGraphDestroy(old.graph);
            c->graphs.erase(c->graphs.begin());
        }
        hipGraph_t gr = nullptr;
        e = hipGraphCreate(&gr, 0);
        if (e != hipSuccess) return e;
        hipGraphNode_t up = nullptr, fill = nullptr, cn = nullptr, pn = nullptr, r1 = nullptr, r2 = nullptr;
        e = hipGraphAddMemcpyNode1D(&up, gr, nullptr, 0, c->d_meta, c->h_meta, c->meta_cap, hipMemcpyHostToDevice);
        unsigned long long* fp = c->d_cnt;
        int fn = (int)cnt_entries;
        void* fargs[] = {&fp, &fn};
        hipKernelNodeParams fpar{};
        fpar.func = reinterpret_cast<void*>(&fill_u64_kernel);
        fpar.gridDim = dim3(1);
        fpar.blockDim = dim3(256);
        fpar.sharedMemBytes = 0;
        fpar.kernelParams = fargs;
        fpar.extra = nullptr;
        if (e == hipSuccess) e = hipGraphAddKernelNode(&fill, gr, nullptr, 0, &fpar);
        hipGraphNode_t deps[2] = {up, fill};
        if (e == hipSuccess) e = hipGraphAddKernelNode(&cn, gr, deps, 2, &kp);
        hipGraphNode_t last = cn;
        if (e == hipSuccess && pfun) {
            e = hipGraphAddKernelNode(&pn, gr, &cn, 1, &pp);
            last = pn;
        }
        if (e == hipSuccess)
            e = hipGraphAddMemcpyNode1D(&r1, gr, &last, 1, c->h_out, c->d_cnt, c->cnt_cap, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipGraphAddMemcpyNode1D(&r2, gr, &last, 1, c->h_out + cnt_entries, c->d_clocks,
                                        n_clk * sizeof(unsigned long long), hipMemcpyDeviceToHost);
        hipGraphExec_t ex = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(gr);
            return e;
        }
        CallCtx::GraphEntry ge;
        ge.key.assign(kb, kb + sizeof(key));
        ge.graph = gr;
        ge.exec = ex;
        ge.count = cn;
        ge.parts = pn;
        c->graphs.push_back(ge);
    }
    std::memcpy(c->h_meta, meta.data(), meta.size());
    *launched = true;
    e = hipEventRecord(c->e0, c->stream);
    if (e == hipSuccess) e = hipGraphLaunch(c->graphs.back().exec, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, c->e0, c->e1);
    if (e != hipSuccess) return e;
    std::memcpy(host_out, c->h_out, n_cnt * sizeof(unsigned long long));
    std::memcpy(clocks, c->h_out + cnt_entries, n_clk * sizeof(unsigned long long));
    return hipSuccess;
}

// The edge-support array of a graph (the House kernel reads it), built on
// the first call that needs it and kept with the graph.
int ensure_support(gp_graph* g, CallCtx* c) {
    std::lock_guard<std::mutex> lk(g->sup_mu);
    if (g->d_sup || g->e_cnt == 0) return GP_OK;
    int32_t* sup = nullptr;
    HIP_TRY(hipMalloc((void**)&sup, sizeof(int32_t) * (size_t)g->e_cnt));
    hipError_t e = ensure((void**)&c->d_cnt, &c->cnt_cap, 8 * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, c->stream, c->d_cnt, 8);
        e = hipGetLastError();
    }
    float ms = 0;
    if (e == hipSuccess) e = hipEventRecord(c->e0, c->stream);
    if (e == hipSuccess)
        e = gp_support_launch(g->d_vertex, g->d_edge, g->d_src, g->e_cnt, sup, c->d_cnt, (unsigned)(g->num_cu * 16),
                              c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->e0, c->e1);
    if (e != hipSuccess) {
        (void)hipFree(sup);
        return gp_fail((int)e, std::string("edge supports: ") + hipGetErrorString(e));
    }
    g->d_sup = sup;
    g->sup_build_ms = ms;
    return GP_OK;
}

int run_tri_edge_tri(gp_graph* g, const gp_sched_flat* sched, uint64_t* raw, int64_t* count, gp_stats* stats);
int count_ranged(gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning, uint64_t* raw_sum,
                 int64_t* count, gp_stats* stats);

// House under the GraphPi schedule's order, with the specialised two-hop
// kernel (house.hip): one launch over a task range of edge slots.
int run_house(gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning, bool all_v1, uint64_t* raw,
              int64_t* count, gp_stats* stats) {
    HIP_TRY(hipSetDevice(g->device));
    const uint64_t all_tasks = (uint64_t)g->e_cnt;
    uint64_t t_begin = 0, t_end = all_tasks;
    if (tuning && (tuning->task_begin != 0 || tuning->task_end > 0)) {
        if (tuning->task_begin < 0 || (tuning->task_end > 0 && tuning->task_end < tuning->task_begin))
            return gp_fail(GP_EINVAL, "bad task range");
        t_begin = std::min<uint64_t>((uint64_t)tuning->task_begin, all_tasks);
        if (tuning->task_end > 0) t_end = std::min<uint64_t>((uint64_t)tuning->task_end, all_tasks);
    }
    const int wpc = tuning && tuning->waves_per_cu > 0 ? tuning->waves_per_cu : 24;
    int64_t waves = (int64_t)g->num_cu * wpc;
    // one dense count array per wave: fewer waves when they would not fit
    const int64_t stride = ((int64_t)std::max(g->v_cnt, 1) + 63) / 64 * 64;
    const int64_t fit = (int64_t)(GP_DENSE_MAX_BYTES / ((uint64_t)stride * sizeof(int32_t)));
    waves = std::max<int64_t>(1, std::min<int64_t>(waves, fit));
    if (t_end > t_begin) waves = std::min<int64_t>(waves, (int64_t)(t_end - t_begin));
    const size_t dense_bytes = (size_t)waves * (size_t)stride * sizeof(int32_t);
    CallCtx* c = nullptr;
    int rc = acquire_ctx(g, 64, &c);
    if (rc != GP_OK) return rc;
    hipError_t e = ensure((void**)&c->d_cnt, &c->cnt_cap, 8 * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess && dense_bytes > c->dense_bytes) {
        e = ensure((void**)&c->d_dense, &c->dense_bytes, dense_bytes, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(zero_i32_kernel, dim3(4096), dim3(256), 0, c->stream, c->d_dense,
                               (uint64_t)(c->dense_bytes / sizeof(int32_t)));
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess && c->clocks_cap < (size_t)waves) {
        size_t cap = c->clocks_cap * 4 * sizeof(unsigned long long);
        e = ensure((void**)&c->d_clocks, &cap, sizeof(unsigned long long) * 4 * (size_t)waves, c->stream);
        c->clocks_cap = e == hipSuccess ? (size_t)waves : 0;
    }
    if (e != hipSuccess) {
        destroy_ctx(c);
        return gp_fail(GP_ENOMEM, std::string("call buffers: ") + hipGetErrorString(e));
    }
    const bool built_now = g->d_sup == nullptr;
    rc = ensure_support(g, c);
    if (rc != GP_OK) {
        destroy_ctx(c);
        return rc;
    }
    GpHouseArgs a;
    a.vertex = g->d_vertex;
    a.edge = g->d_edge;
    a.sup = g->d_sup;
    a.v_cnt = g->v_cnt;
    a.slot_begin = t_begin;
    a.slot_end = t_end;
    a.all_v1 = all_v1 ? 1 : 0;
    a.counter = c->d_cnt + 1;
    a.dense = c->d_dense;
    a.dense_stride = stride;
    a.out = c->d_cnt + 3;
    a.clocks = c->d_clocks;
    unsigned long long host_out[8] = {0};
    std::vector<unsigned long long> clocks(2 * (size_t)waves, 0);
    float ms = 0;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, c->stream, c->d_cnt, 8);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->e0, c->stream);
    if (e == hipSuccess && t_end > t_begin) e = gp_house_launch(a, (unsigned)waves, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, c->d_cnt, sizeof(host_out), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && t_end > t_begin)
        e = hipMemcpyAsync(clocks.data(), c->d_clocks, sizeof(unsigned long long) * 2 * (size_t)waves,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->e0, c->e1);
    if (e != hipSuccess) {
        destroy_ctx(c);
        return gp_fail((int)e, std::string("House kernel: ") + hipGetErrorString(e));
    }
    release_ctx(g, c);
    const uint64_t sum = host_out[3];
    if (raw) *raw = sum;
    if (count) *count = (int64_t)sum / (sched->redundancy ? sched->redundancy : 1);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->kernel_ms = ms + (built_now ? g->sup_build_ms : 0.0);
        stats->tasks = host_out[5];
        stats->waves = (int32_t)waves;
        stats->kernel_variant = 2;
        stats->launches = 1;
        stats->finished = 1;
        stats->tasks_total = t_end > t_begin ? t_end - t_begin : 0;
        if (t_end > t_begin) {
            int rate_khz = 0;
            if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess ||
                rate_khz <= 0)
                rate_khz = 100000;
            unsigned long long t0 = ~0ull;
            for (int64_t w = 0; w < waves; ++w) t0 = std::min(t0, clocks[2 * w]);
            std::vector<double> fin((size_t)waves);
            for (int64_t w = 0; w < waves; ++w) fin[w] = (double)(clocks[2 * w + 1] - t0) / rate_khz;
            std::sort(fin.begin(), fin.end());
            stats->finish_p50_ms = fin[fin.size() / 2];
            stats->finish_p90_ms = fin[fin.size() * 9 / 10];
            stats->finish_max_ms = fin.back();
        }
    }
    return GP_OK;
}

// Count nplans schedules over one graph in one persistent launch (plus the
// launch of deferred parts).  Results are per schedule; a task range in the
// tuning is only accepted for a single schedule.
// rb, re: per-plan task ranges (see prepare).  deadline_ms > 0: the task
// launch stops fetching after that long (the kernel's deadline), and
// plan_done[p] receives how many of plan p's tasks, from its first, were
// counted (all of them without a deadline).
// abort_ms (optional, per plan, > 0 to apply): plan p's tasks are abandoned
// that long after the first launch started (KPlan::abort_ticks); aborted[p]
// then receives 1 and plan p's raw sum and progress are meaningless.
int run_plans(gp_graph* g, const gp_sched_flat* scheds, int nplans, const gp_tuning* tuning, uint64_t* raw,
              int64_t* count, gp_stats* stats, const uint64_t* rb = nullptr, const uint64_t* re = nullptr,
              double deadline_ms = 0, uint64_t* plan_done = nullptr, const double* abort_ms = nullptr,
              int32_t* aborted = nullptr, double task_ms = 0) {
    // schedule-specialised kernels for single counts of shapes they cover
    if (nplans == 1 && !rb && deadline_ms <= 0 && !abort_ms && task_ms <= 0 &&
        !(tuning && (tuning->collect_balg || tuning->no_special))) {
        const int shape = gp_sched_shape(&scheds[0]);
        if (shape == GP_SHAPE_HOUSE || shape == GP_SHAPE_HOUSE_ALL)
            return run_house(g, scheds, tuning, shape == GP_SHAPE_HOUSE_ALL, raw, count, stats);
        if (shape == GP_SHAPE_TRI_EDGE_TRI && !(tuning && (tuning->task_begin != 0 || tuning->task_end > 0)))
            return run_tri_edge_tri(g, scheds, raw, count, stats);
    }
    Launch ln;
    {
        int rc = prepare(g, scheds, nplans, tuning, &ln, rb, re);
        if (rc != GP_OK) return rc;
    }
    HIP_TRY(hipSetDevice(g->device));
    if (abort_ms)
        for (int p = 0; p < nplans; ++p)
            ln.kps[p].abort_ticks =
                abort_ms[p] > 0 ? std::max<int64_t>(1, (int64_t)(abort_ms[p] * g->wall_khz)) : 0;
    if (task_ms > 0)
        for (int p = 0; p < nplans; ++p) ln.kps[p].task_ticks = std::max<int64_t>(1, (int64_t)(task_ms * g->wall_khz));
    bool halt = false; // the kernel instantiation with the hard-stop checks
    for (int p = 0; p < nplans; ++p) halt = halt || ln.kps[p].abort_ticks > 0 || ln.kps[p].task_ticks > 0;
    const bool faithful = ln.faithful, any_split = ln.any_split;
    const std::vector<KPlan>& kps = ln.kps;
    const std::vector<uint64_t>&task_off = ln.task_off, &task_first = ln.task_first, &task_cnt = ln.task_cnt;
    const uint64_t n_tasks = ln.n_tasks;
    const int64_t waves = ln.waves;
    const int chunk = ln.chunk;
    const int64_t slot_stride = ln.slot_stride, tmp_stride = ln.tmp_stride, bslot_stride = ln.bslot_stride,
                  wave_stride = ln.wave_stride;
    const size_t scratch_bytes = ln.scratch_bytes;
    const int hash_slots = ln.hash_slots, hash_lg = ln.hash_lg;
    const size_t lds_bytes = ln.lds_bytes;
    const int64_t dense_stride = ln.dense_stride;
    const size_t dense_bytes = ln.dense_bytes;
    CallCtx* c = nullptr;
    int rc = acquire_ctx(g, scratch_bytes, &c);
    if (rc != GP_OK) return rc;
    const size_t plans_bytes = sizeof(KPlan) * (size_t)nplans;
    const size_t meta_bytes = plans_bytes + sizeof(uint64_t) * (3 * (size_t)nplans + 1);
    // counters: fetch, parts fetch, deferred count, 3 per plan, an abort
    // flag per plan, the first launch's start time
    const size_t n_cnt = 4 + 4 * (size_t)nplans;
    hipError_t e = ensure((void**)&c->d_meta, &c->meta_cap, meta_bytes, c->stream);
    if (e == hipSuccess) e = ensure((void**)&c->d_cnt, &c->cnt_cap, n_cnt * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess && dense_bytes > c->dense_bytes) {
        // a new array starts at zero; the kernels leave it at zero
        e = ensure((void**)&c->d_dense, &c->dense_bytes, dense_bytes, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(zero_i32_kernel, dim3(4096), dim3(256), 0, c->stream, c->d_dense,
                               (uint64_t)(c->dense_bytes / sizeof(int32_t)));
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess && c->clocks_cap < (size_t)waves) {
        // start and finish stamps per wave, for each of the two launches
        size_t cap = c->clocks_cap * 4 * sizeof(unsigned long long);
        e = ensure((void**)&c->d_clocks, &cap, sizeof(unsigned long long) * 4 * (size_t)waves, c->stream);
        c->clocks_cap = e == hipSuccess ? (size_t)waves : 0;
    }
    if (e != hipSuccess) {
        destroy_ctx(c);
        return gp_fail(GP_ENOMEM, std::string("call buffers: ") + hipGetErrorString(e));
    }
    std::vector<char> meta(meta_bytes);
    std::memcpy(meta.data(), kps.data(), plans_bytes);
    std::memcpy(meta.data() + plans_bytes, task_off.data(), sizeof(uint64_t) * ((size_t)nplans + 1));
    std::memcpy(meta.data() + plans_bytes + sizeof(uint64_t) * ((size_t)nplans + 1), task_first.data(),
                sizeof(uint64_t) * (size_t)nplans);
    std::memcpy(meta.data() + plans_bytes + sizeof(uint64_t) * (2 * (size_t)nplans + 1), task_cnt.data(),
                sizeof(uint64_t) * (size_t)nplans);
    KArgs a{};
    a.plans = reinterpret_cast<const KPlan*>(c->d_meta);
    a.nplans = nplans;
    a.task_off = reinterpret_cast<const uint64_t*>(c->d_meta + plans_bytes);
    a.task_first = reinterpret_cast<const uint64_t*>(c->d_meta + plans_bytes + sizeof(uint64_t) * ((size_t)nplans + 1));
    a.task_cnt = reinterpret_cast<const uint64_t*>(c->d_meta + plans_bytes + sizeof(uint64_t) * (2 * (size_t)nplans + 1));
    a.vertex = g->d_vertex;
    a.edge = g->d_edge;
    a.src = g->d_src;
    a.v_cnt = g->v_cnt;
    a.chunk = chunk;
    a.fetch_ticks = ln.fetch_ticks;
    a.n_tasks = n_tasks;
    a.counter = c->d_cnt;
    a.counter2 = c->d_cnt + 1;
    a.heavy_cnt = c->d_cnt + 2;
    a.out = c->d_cnt + 3;
    a.heavy = c->d_heavy;
    a.heavy_cap = GP_HEAVY_CAP;
    a.scratch = c->d_scratch;
    a.slot_stride = slot_stride;
    a.wave_stride = wave_stride;
    a.tmp_stride = tmp_stride;
    a.bslot_stride = bslot_stride;
    a.hash_slots = hash_slots;
    a.hash_lg = hash_lg;
    a.probe_pct = tuning && tuning->probe_pct > 0 ? tuning->probe_pct : 100;
    a.bitmap_ratio = !tuning || tuning->bitmap_ratio == 0 ? GP_BITMAP_RATIO : std::max(0, tuning->bitmap_ratio);
    a.merge_ratio = tuning ? std::max(0, tuning->merge_ratio) : 0;
    a.deadline_ticks = deadline_ms > 0 ? std::max<int64_t>(1, (int64_t)(deadline_ms * g->wall_khz)) : 0;
    a.aborted = c->d_cnt + 3 + 3 * (size_t)nplans;
    a.origin = c->d_cnt + 3 + 4 * (size_t)nplans;
    a.dense = dense_bytes ? c->d_dense : nullptr;
    a.dense_stride = dense_stride;
    a.clocks = c->d_clocks;
    std::vector<unsigned long long> host_out(n_cnt, 0);
    std::vector<unsigned long long> clocks(4 * (size_t)waves, 0);
    const bool parts_launch = any_split && n_tasks > 0;
    const bool multi = nplans > 1;
    // the extended kernel for a plan that takes the flagged shared search or
    // the pair level grouped by v1, or for merge-path prefix builds when the
    // caller asks for them.  Multi-plan launches stay on the plain
    // instantiations: an extended multi-plan kernel spilled far more
    // (128 B/lane) than the shared search saved in a batch of mostly cheap
    // patterns.
    int ext = GP_EXT_NONE;
    if (!faithful && !multi) {
        const KPlan& kp = kps[0];
        if (kp.batched && kp.bl_base >= 0 && g->v_cnt < GP_FLAG_VMAX)
            ext = GP_EXT_SEARCH;
        else if (kp.pair_v1 && dense_bytes)
            ext = GP_EXT_PAIR;
        else if (tuning && tuning->merge_ratio > 0)
            ext = GP_EXT_MERGE;
    }
    float ms = 0;
    bool launched = false;
    if (tuning && tuning->use_graph && n_tasks > 0) {
        // the same device work as below, as one replayable graph
        const void* kfun = count_kernel_fn(faithful, false, multi, ext, halt);
        const void* pfun = count_kernel_fn(false, true, multi, ext, halt);
        KArgs a2 = a;
        a2.clocks = c->d_clocks + 2 * waves;
        const size_t n_clk = (parts_launch ? 4 : 2) * (size_t)waves;
        e = graph_run(c, kfun, parts_launch ? pfun : nullptr, (unsigned)waves, lds_bytes, a, a2, meta, n_cnt, n_clk,
                      host_out.data(), clocks.data(), &ms, &launched);
    }
    if (e == hipSuccess && !launched) {
        e = hipMemcpyAsync(c->d_meta, meta.data(), meta_bytes, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, c->stream, c->d_cnt, (int)n_cnt);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(c->e0, c->stream);
        if (e == hipSuccess && n_tasks > 0) {
            const dim3 grid((unsigned)waves), block(WAVE);
            void* kargs[] = {&a};
            e = hipLaunchKernel(count_kernel_fn(faithful, false, multi, ext, halt), grid, block, kargs, lds_bytes,
                                c->stream);
            if (e == hipSuccess && parts_launch) {
                // the deferred tasks, in parts; the kernel reads how many there are
                KArgs a2 = a;
                a2.clocks = c->d_clocks + 2 * waves;
                void* pargs[] = {&a2};
                e = hipLaunchKernel(count_kernel_fn(false, true, multi, ext, halt), grid, block, pargs, lds_bytes,
                                    c->stream);
            }
        }
        if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(host_out.data(), c->d_cnt, n_cnt * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               c->stream);
        if (e == hipSuccess && n_tasks > 0)
            e = hipMemcpyAsync(clocks.data(), c->d_clocks,
                               sizeof(unsigned long long) * (parts_launch ? 4 : 2) * (size_t)waves,
                               hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->e0, c->e1);
    }
    if (e != hipSuccess) {
        // the context may hold a poisoned stream; drop it
        destroy_ctx(c);
        return gp_fail((int)e, std::string("count kernel: ") + hipGetErrorString(e));
    }
    release_ctx(g, c);
    if (aborted)
        for (int p = 0; p < nplans; ++p) aborted[p] = host_out[3 + 3 * (size_t)nplans + p] ? 1 : 0;
    if (plan_done) {
        // the items below the counter's final value were counted
        const uint64_t upto = a.deadline_ticks > 0 ? std::min<uint64_t>(host_out[0], n_tasks) : n_tasks;
        for (int p = 0; p < nplans; ++p)
            plan_done[p] = upto > task_off[p] ? std::min<uint64_t>(upto - task_off[p], task_cnt[p]) : 0;
    }
    uint64_t balg_total = 0, tasks_total = 0;
    for (int p = 0; p < nplans; ++p) {
        const uint64_t sum = host_out[3 + 3 * p];
        if (raw) raw[p] = sum;
        if (count) count[p] = (int64_t)sum / (scheds[p].redundancy ? scheds[p].redundancy : 1);
        // every root's neighbour list is looked up once by the reference;
        // those lookups are attributed to the range that starts at task 0
        if (faithful)
            balg_total += host_out[4 + 3 * p] +
                          (kps[p].root_mode ? 8ull * task_cnt[p]
                                            : task_first[p] == 0 ? 8ull * (uint64_t)g->v_cnt : 0);
        tasks_total += host_out[5 + 3 * p];
    }
    if (stats) {
        stats->kernel_ms = ms;
        stats->balg_bytes = balg_total;
        stats->tasks = tasks_total;
        stats->waves = (int32_t)waves;
        stats->kernel_variant = faithful ? 1 : 0;
        stats->split_tasks = parts_launch ? std::min<uint64_t>(host_out[2], GP_HEAVY_CAP) : 0;
        stats->inplace_tasks = parts_launch && host_out[2] > GP_HEAVY_CAP ? host_out[2] - GP_HEAVY_CAP : 0;
        stats->launches = 1;
        stats->finished = 1;
        stats->tasks_total = 0;
        for (int p = 0; p < nplans; ++p) stats->tasks_total += task_cnt[p];
        stats->finish_p50_ms = stats->finish_p90_ms = stats->finish_max_ms = 0;
        if (n_tasks > 0) {
            int rate_khz = 0;
            if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess ||
                rate_khz <= 0)
                rate_khz = 100000;
            unsigned long long t0 = ~0ull;
            for (int64_t w = 0; w < waves; ++w) t0 = std::min(t0, clocks[2 * w]);
            std::vector<double> fin((size_t)waves);
            for (int64_t w = 0; w < waves; ++w) {
                unsigned long long f = clocks[2 * w + 1];
                if (parts_launch) f = std::max(f, clocks[2 * waves + 2 * w + 1]);
                fin[w] = (double)(f - t0) / rate_khz;
            }
            std::sort(fin.begin(), fin.end());
            stats->finish_p50_ms = fin[fin.size() / 2];
            stats->finish_p90_ms = fin[fin.size() * 9 / 10];
            stats->finish_max_ms = fin.back();
        }
    }
    return GP_OK;
}

// Two triangles joined by an edge under the GraphPi schedule (shape
// GP_SHAPE_TRI_EDGE_TRI), counted in closed form instead of by walking every
// 3-path.  An occurrence is a bridge edge {a, b} with a triangle T1 at a that
// avoids b and a triangle T2 at b that avoids a, all six vertices distinct:
//   C = sum over edges {a, b} of (t_a - s_ab)(t_b - s_ab) - X_ab,
// with t the triangles at a vertex, s the edge supports, and X_ab the pairs
// (T1, T2) that share a vertex.  By inclusion-exclusion over the shared
// vertices v (each in N(a) ∩ N(b), where T1 ∋ v has s_av - 1 choices and
// T2 ∋ v has s_bv - 1), pairs sharing both vertices being the edges inside
// N(a) ∩ N(b), i.e. the 4-cliques on {a, b}:
//   X_ab = sum_{v in N(a) ∩ N(b)} (s_av - 1)(s_bv - 1) - K4(a, b),
// and the K4 terms add up to 6 x the number of 4-cliques.  The reference's
// loop nest enforces only v3 < v0 (its other two restrictions sit at IEP
// depths), which keeps exactly 4 of each occurrence's 8 labelled embeddings,
// so its raw sum is 4 C (mod 2^64) and its result raw / 4.  Everything is u64
// arithmetic, so even a count past 2^63 matches the reference's wrapped sum.
int run_tri_edge_tri(gp_graph* g, const gp_sched_flat* sched, uint64_t* raw, int64_t* count, gp_stats* stats) {
    HIP_TRY(hipSetDevice(g->device));
    // 4-cliques, each once (v3 < v2 < v1 < v0), on the generic kernel in bounded launches
    const int adj4[16] = {0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0};
    Schedule s4(adj4, 4);
    s4.add_restrict({{0, 1}, {1, 2}, {2, 3}});
    gp_sched_flat flat4;
    int rc = s4.flatten(&flat4);
    if (rc != GP_OK) return gp_fail(rc, "4-clique schedule");
    uint64_t k4 = 0;
    gp_stats st4{};
    rc = count_ranged(g, &flat4, nullptr, &k4, nullptr, &st4);
    if (rc != GP_OK) return rc;
    CallCtx* c = nullptr;
    rc = acquire_ctx(g, 64, &c);
    if (rc != GP_OK) return rc;
    const unsigned waves = (unsigned)(g->num_cu * 16);
    hipError_t e = ensure((void**)&c->d_cnt, &c->cnt_cap, 8 * sizeof(unsigned long long), c->stream);
    if (e == hipSuccess && c->clocks_cap < (size_t)waves) { // the per-wave 128-bit sums go there
        size_t cap = c->clocks_cap * 4 * sizeof(unsigned long long);
        e = ensure((void**)&c->d_clocks, &cap, sizeof(unsigned long long) * 4 * (size_t)waves, c->stream);
        c->clocks_cap = e == hipSuccess ? (size_t)waves : 0;
    }
    if (e != hipSuccess) {
        destroy_ctx(c);
        return gp_fail(GP_ENOMEM, std::string("call buffers: ") + hipGetErrorString(e));
    }
    const bool sup_now = g->d_sup == nullptr;
    rc = ensure_support(g, c);
    if (rc != GP_OK) {
        destroy_ctx(c);
        return rc;
    }
    float ms_t = 0, ms_e = 0;
    {
        std::lock_guard<std::mutex> lk(g->sup_mu);
        if (!g->d_vtri && g->v_cnt > 0) {
            uint64_t* t = nullptr;
            e = hipMalloc((void**)&t, sizeof(uint64_t) * (size_t)g->v_cnt);
            if (e == hipSuccess) e = hipEventRecord(c->e0, c->stream);
            if (e == hipSuccess) e = gp_vertex_triangles_launch(g->d_vertex, g->d_sup, g->v_cnt, t, c->stream);
            if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms_t, c->e0, c->e1);
            if (e != hipSuccess) {
                if (t) (void)hipFree(t);
                destroy_ctx(c);
                return gp_fail((int)e, std::string("vertex triangles: ") + hipGetErrorString(e));
            }
            g->d_vtri = t;
        }
    }
    unsigned long long host_out[8] = {0};
    std::vector<unsigned long long> wide(2 * (size_t)waves, 0);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, c->stream, c->d_cnt, 8);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->e0, c->stream);
    if (e == hipSuccess && g->e_cnt)
        e = gp_tri_edge_launch(g->d_vertex, g->d_edge, g->d_src, g->d_sup, g->d_vtri, g->e_cnt, c->d_cnt + 1,
                               c->d_cnt + 3, c->d_clocks, waves, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->e1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, c->d_cnt, sizeof(host_out), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && g->e_cnt)
        e = hipMemcpyAsync(wide.data(), c->d_clocks, sizeof(unsigned long long) * wide.size(), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_e, c->e0, c->e1);
    if (e != hipSuccess) {
        destroy_ctx(c);
        return gp_fail((int)e, std::string("tri-edge-tri kernel: ") + hipGetErrorString(e));
    }
    release_ctx(g, c);
    const uint64_t cnt = host_out[3] + 6ull * k4;       // C mod 2^64
    const uint64_t r = cnt * (uint64_t)sched->redundancy; // the loop nest's raw sum, 4 C
    if (raw) *raw = r;
    if (count) *count = (int64_t)r / sched->redundancy;
    // the count itself in 128 bits (the reference's long long wraps past 2^63)
    __int128 exact = (__int128)k4 * 6;
    if (g->e_cnt)
        for (unsigned w = 0; w < waves; ++w)
            exact += (__int128)(((unsigned __int128)wide[2 * w + 1] << 64) | wide[2 * w]);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->exact_count_valid = 1;
        stats->exact_count_lo = (uint64_t)(unsigned __int128)exact;
        stats->exact_count_hi = (uint64_t)((unsigned __int128)exact >> 64);
        stats->kernel_ms = st4.kernel_ms + ms_t + ms_e + (sup_now ? g->sup_build_ms : 0.0);
        stats->tasks = g->e_cnt;
        stats->tasks_total = g->e_cnt;
        stats->waves = g->num_cu * 16;
        stats->kernel_variant = 3;
        stats->launches = st4.launches + 1;
        stats->finished = 1;
    }
    return GP_OK;
}

} // namespace

extern "C" int gp_device_count(int32_t* n) {
    if (!n) return gp_fail(GP_EINVAL, "null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = e == hipSuccess ? c : 0;
    return GP_OK;
}

extern "C" int gp_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return gp_fail(GP_EINVAL, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return gp_fail(GP_ENODEV, "no HIP device");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return GP_OK;
}

extern "C" int gp_graph_upload(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt, uint32_t e_cnt,
                               gp_graph** out) {
    if (!out || !vertex || (!edge && e_cnt) || v_cnt < 0) return gp_fail(GP_EINVAL, "bad CSR arguments");
    {
        const int rc = gp_csr_validate(edge, vertex, v_cnt, e_cnt);
        if (rc != GP_OK) return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return gp_fail(GP_ENODEV, "no HIP device");
    gp_graph* g = new gp_graph();
    HIP_TRY(hipGetDevice(&g->device));
    HIP_TRY(hipDeviceGetAttribute(&g->num_cu, hipDeviceAttributeMultiprocessorCount, g->device));
    {
        // call buffers come from the device's default pool (stream-ordered);
        // keep what calls free in the pool instead of handing it back to the
        // driver at every synchronisation, so that calls reuse it
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, g->device) == hipSuccess && pool) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) == hipSuccess && khz > 0)
            g->wall_khz = khz;
    }
    g->v_cnt = v_cnt;
    g->e_cnt = e_cnt;
    for (int32_t v = 0; v < v_cnt; ++v) {
        int32_t d = (int32_t)(vertex[v + 1] - vertex[v]);
        if (d > g->max_deg) {
            g->max_deg2 = g->max_deg;
            g->max_deg = d;
        } else if (d > g->max_deg2) {
            g->max_deg2 = d;
        }
    }
    auto fail = [&](hipError_t e, const char* what) {
        gp_graph_free(g);
        return gp_fail((int)e, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&g->d_vertex, sizeof(uint32_t) * ((size_t)v_cnt + 1))) != hipSuccess) return fail(e, "hipMalloc vertex");
    if ((e = hipMalloc((void**)&g->d_edge, sizeof(int32_t) * std::max<size_t>(e_cnt, 1))) != hipSuccess) return fail(e, "hipMalloc edge");
    if ((e = hipMalloc((void**)&g->d_src, sizeof(int32_t) * std::max<size_t>(e_cnt, 1))) != hipSuccess) return fail(e, "hipMalloc src");
    if ((e = hipMemcpy(g->d_vertex, vertex, sizeof(uint32_t) * ((size_t)v_cnt + 1), hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "copy vertex");
    if (e_cnt && (e = hipMemcpy(g->d_edge, edge, sizeof(int32_t) * (size_t)e_cnt, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "copy edge");
    if (e_cnt) {
        unsigned blocks = (unsigned)(((uint64_t)e_cnt + 255) / 256);
        hipLaunchKernelGGL(build_src_kernel, dim3(blocks), dim3(256), 0, 0, g->d_vertex, v_cnt, e_cnt, g->d_src);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "build_src_kernel");
        if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "build_src_kernel sync");
    }
    *out = g;
    return GP_OK;
}

extern "C" int gp_graph_free(gp_graph* g) {
    if (!g) return GP_OK;
    (void)hipSetDevice(g->device);
    for (CallCtx* c : g->idle) destroy_ctx(c);
    if (g->d_vertex) (void)hipFree(g->d_vertex);
    if (g->d_edge) (void)hipFree(g->d_edge);
    if (g->d_src) (void)hipFree(g->d_src);
    if (g->d_sup) (void)hipFree(g->d_sup);
    if (g->d_vtri) (void)hipFree(g->d_vtri);
    delete g;
    return GP_OK;
}

extern "C" int gp_graph_bytes(const gp_graph* g, uint64_t* bytes) {
    if (!g || !bytes) return gp_fail(GP_EINVAL, "null argument");
    *bytes = sizeof(uint32_t) * ((uint64_t)g->v_cnt + 1) + (g->d_sup ? 3 : 2) * sizeof(int32_t) * (uint64_t)g->e_cnt +
             (g->d_vtri ? sizeof(uint64_t) * (uint64_t)g->v_cnt : 0);
    return GP_OK;
}

namespace {

double wall_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One schedule's count as consecutive bounded launches, so that no single
// launch runs for long and a host-side time limit is checked between them.
// The generic kernel gets a deadline of slice_ms per launch: each launch
// counts a prefix of the remaining tasks and the next starts where it
// stopped, so a count that fits one slice is one launch and a long one
// becomes launches of about slice_ms whatever the cost of its tasks.  The
// House kernel (a root per fetch) is sliced by GpRangePlan instead.  Raw
// sums of consecutive ranges add up exactly (u64 wrap-around), so the
// result is the one-launch result.
int count_ranged(gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning, uint64_t* raw_sum,
                 int64_t* count, gp_stats* stats) {
    KPlan kp;
    int rc = gp_make_kplan(sched, false, true, true, &kp);
    if (rc != GP_OK) return rc;
    const uint64_t all_tasks = kp.root_mode ? (uint64_t)g->v_cnt : (uint64_t)g->e_cnt;
    gp_tuning t{};
    if (tuning) t = *tuning;
    if (t.task_begin < 0 || (t.task_end > 0 && t.task_end < t.task_begin)) return gp_fail(GP_EINVAL, "bad task range");
    const uint64_t begin = std::min<uint64_t>((uint64_t)t.task_begin, all_tasks);
    const uint64_t end = t.task_end > 0 ? std::min<uint64_t>((uint64_t)t.task_end, all_tasks) : all_tasks;
    // the House kernel's unit of work is a whole root, the generic kernel's an
    // edge slot; a hard time limit takes the generic kernel, which can stop
    // inside a task
    const bool hard = t.abort_s > 0;
    const int shape = (t.collect_balg || t.no_special || hard) ? GP_SHAPE_GENERIC : gp_sched_shape(sched);
    if (shape == GP_SHAPE_TRI_EDGE_TRI && t.task_begin == 0 && t.task_end <= 0)
        return run_tri_edge_tri(g, sched, raw_sum, count, stats);
    const bool per_root = shape == GP_SHAPE_HOUSE || shape == GP_SHAPE_HOUSE_ALL;
    const double t0 = wall_seconds();
    if ((t.slice_ms < 0 && !hard) || end <= begin) return run_plans(g, sched, 1, &t, raw_sum, count, stats);
    const double slice = t.slice_ms > 0 ? (double)t.slice_ms : 2000.0;
    gp_stats acc{};
    uint64_t raw = 0, cur = begin;
    bool timed_out = false;
    auto add = [&](const gp_stats& st) {
        const double before_ms = acc.kernel_ms;
        acc.kernel_ms += st.kernel_ms;
        acc.balg_bytes += st.balg_bytes;
        acc.tasks += st.tasks;
        acc.split_tasks += st.split_tasks;
        acc.inplace_tasks += st.inplace_tasks;
        acc.waves = std::max(acc.waves, st.waves);
        acc.kernel_variant = st.kernel_variant;
        // finish times of the last launch, offset by the launches before it
        acc.finish_p50_ms = before_ms + st.finish_p50_ms;
        acc.finish_p90_ms = before_ms + st.finish_p90_ms;
        acc.finish_max_ms = before_ms + st.finish_max_ms;
        ++acc.launches;
    };
    if (per_root) {
        GpRangePlan plan(begin, end, t.slice_ms, t.time_limit_s, t0, 4);
        uint64_t b = 0, e = 0;
        while (plan.next(&b, &e)) {
            gp_tuning sub = t;
            sub.task_begin = (int64_t)b;
            sub.task_end = (int64_t)e;
            uint64_t r = 0;
            gp_stats st{};
            rc = run_plans(g, sched, 1, &sub, &r, nullptr, &st);
            if (rc != GP_OK) return rc;
            raw += r;
            add(st);
            plan.finished(e, st.kernel_ms, wall_seconds());
        }
        cur = plan.cur;
        timed_out = plan.timed_out;
    } else {
        while (cur < end) {
            gp_tuning sub = t;
            sub.task_begin = (int64_t)cur;
            sub.task_end = (int64_t)end;
            uint64_t r = 0, did = 0;
            gp_stats st{};
            double abort_ms = 0;
            if (hard) {
                abort_ms = (t.abort_s - (wall_seconds() - t0)) * 1000.0;
                if (abort_ms <= 0) {
                    timed_out = true;
                    break;
                }
            }
            int32_t ab = 0;
            rc = run_plans(g, sched, 1, &sub, &r, nullptr, &st, nullptr, nullptr, t.slice_ms < 0 ? 0 : slice, &did,
                           hard ? &abort_ms : nullptr, &ab);
            if (rc != GP_OK) return rc;
            if (ab) { // abandoned: the launches before it stand; its time counts
                acc.finish_p50_ms = acc.kernel_ms + st.finish_p50_ms;
                acc.finish_p90_ms = acc.kernel_ms + st.finish_p90_ms;
                acc.finish_max_ms = acc.kernel_ms + st.finish_max_ms;
                acc.kernel_ms += st.kernel_ms;
                ++acc.launches;
                timed_out = true;
                break;
            }
            raw += r;
            add(st);
            cur += did;
            if (cur < end && t.time_limit_s > 0 && wall_seconds() - t0 > t.time_limit_s) {
                timed_out = true;
                break;
            }
        }
    }
    acc.finished = timed_out ? 0 : 1;
    acc.tasks_total = end > begin ? end - begin : 0;
    if (raw_sum) *raw_sum = raw;
    if (stats) *stats = acc;
    if (timed_out)
        return gp_fail(GP_ETIME, "time limit of " + std::to_string(t.time_limit_s) + " s reached after " +
                                     std::to_string(cur - begin) + " of " + std::to_string(end - begin) + " tasks");
    if (count) count[0] = (int64_t)raw / (sched->redundancy ? sched->redundancy : 1);
    return GP_OK;
}

} // namespace

extern "C" int gp_count(const gp_graph* g, const gp_sched_flat* sched, const gp_tuning* tuning, uint64_t* raw_sum,
                        int64_t* count, gp_stats* stats) {
    if (!g || !sched) return gp_fail(GP_EINVAL, "null argument");
    return count_ranged(const_cast<gp_graph*>(g), sched, tuning, raw_sum, count, stats);
}

// ---------------------------------------------------------------------------
// allocation-free counting (gp_count_noalloc)
// ---------------------------------------------------------------------------
// Every device buffer the count needs besides the graph and the caller's
// outputs (kernel plan and work layout, counters, clock stamps, the deferred
// task list and the per-wave prefix sets) lives in one static device array of
// GP_NOALLOC_BYTES, so the call makes no hipMalloc/hipFree.  Each device the
// library runs on holds its own copy of the array for the life of the
// process.  Calls on one device are serialised by a per-device lock, so they
// never overlap on the array inside a process.
#define GP_NOALLOC_BYTES (128ull << 20)   // static scratch per device
#define GP_NOALLOC_HEAVY (1u << 16)       // deferred-task list entries (the rest run in place)
#define GP_NOALLOC_DEVICES 64

namespace {

__device__ __align__(256) unsigned char scratch[GP_NOALLOC_BYTES];

struct NoAllocDevice {
    std::mutex mu;
    unsigned char* ptr = nullptr; // device address of `scratch` on this device
};
NoAllocDevice g_noalloc[GP_NOALLOC_DEVICES];

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

} // namespace

extern "C" int gp_count_noalloc(const gp_graph* gc, const gp_sched_flat* sched, const gp_tuning* tuning,
                                uint64_t* raw_sum, int64_t* count, gp_stats* stats) {
    if (!gc || !sched) return gp_fail(GP_EINVAL, "null argument");
    gp_graph* g = const_cast<gp_graph*>(gc);
    if (g->device < 0 || g->device >= GP_NOALLOC_DEVICES) return gp_fail(GP_ERANGE, "device index out of range");
    HIP_TRY(hipSetDevice(g->device));
    NoAllocDevice& nd = g_noalloc[g->device];
    std::lock_guard<std::mutex> lk(nd.mu);
    if (!nd.ptr) {
        // the array's address on this device (the host-side name is only a stand-in)
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gp_fail(e != hipSuccess ? (int)e : GP_ENODEV,
                           std::string("static scratch address: ") + hipGetErrorString(e));
        nd.ptr = static_cast<unsigned char*>(p);
    }
    // one launch (plus the deferred parts) on the generic kernel; the pair
    // level uses its LDS map, since its dense arrays would not fit, and keeps
    // the ordinary task order (grouping by v1 serves the dense arrays)
    gp_tuning t{};
    if (tuning) t = *tuning;
    t.pair_map = 1;
    t.no_pair_v1 = 1;
    t.use_graph = 0;
    t.no_special = 1;
    Launch ln;
    {
        const int rc = prepare(g, sched, 1, &t, &ln);
        if (rc != GP_OK) return rc;
    }
    const size_t plans_bytes = sizeof(KPlan);
    const size_t meta_bytes = plans_bytes + sizeof(uint64_t) * 4;
    const size_t n_cnt = 8; // fetch, parts fetch, deferred count, the plan's 3, abort flag, start time
    const size_t off_cnt = align256(meta_bytes);
    const size_t off_heavy = off_cnt + align256(n_cnt * sizeof(unsigned long long));
    const size_t off_clk = off_heavy + align256(GP_NOALLOC_HEAVY * sizeof(unsigned long long));
    const size_t wave_bytes = (size_t)ln.wave_stride * sizeof(int32_t);
    const size_t per_wave = 4 * sizeof(unsigned long long) + wave_bytes;
    if (off_clk + 512 + per_wave > GP_NOALLOC_BYTES)
        return gp_fail(GP_ENOMEM, "gp_count_noalloc: one wave needs " + std::to_string(per_wave) +
                                      " bytes of scratch (prefix sets sized by the second-largest degree) beyond " +
                                      std::to_string(off_clk) + " fixed bytes; the static scratch holds " +
                                      std::to_string(GP_NOALLOC_BYTES) + " bytes");
    const int64_t fit = (int64_t)((GP_NOALLOC_BYTES - off_clk - 512) / per_wave);
    const int64_t waves = std::max<int64_t>(1, std::min<int64_t>(ln.waves, fit));
    const size_t off_scr = align256(off_clk + 4 * sizeof(unsigned long long) * (size_t)waves);
    if (off_scr + (size_t)waves * wave_bytes > GP_NOALLOC_BYTES)
        return gp_fail(GP_ENOMEM, "gp_count_noalloc: layout exceeds the static scratch");
    unsigned char* base = nd.ptr;
    unsigned char meta[sizeof(KPlan) + 4 * sizeof(uint64_t)];
    std::memcpy(meta, ln.kps.data(), plans_bytes);
    const uint64_t layout[4] = {ln.task_off[0], ln.task_off[1], ln.task_first[0], ln.task_cnt[0]};
    std::memcpy(meta + plans_bytes, layout, sizeof(layout));
    KArgs a{};
    std::memset(&a, 0, sizeof(a));
    a.plans = reinterpret_cast<const KPlan*>(base);
    a.nplans = 1;
    a.task_off = reinterpret_cast<const uint64_t*>(base + plans_bytes);
    a.task_first = reinterpret_cast<const uint64_t*>(base + plans_bytes + 2 * sizeof(uint64_t));
    a.task_cnt = reinterpret_cast<const uint64_t*>(base + plans_bytes + 3 * sizeof(uint64_t));
    a.vertex = g->d_vertex;
    a.edge = g->d_edge;
    a.src = g->d_src;
    a.v_cnt = g->v_cnt;
    a.chunk = ln.chunk;
    a.fetch_ticks = ln.fetch_ticks;
    a.n_tasks = ln.n_tasks;
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base + off_cnt);
    a.counter = cnt;
    a.counter2 = cnt + 1;
    a.heavy_cnt = cnt + 2;
    a.out = cnt + 3;
    a.aborted = cnt + 6;
    a.origin = cnt + 7;
    a.heavy = reinterpret_cast<unsigned long long*>(base + off_heavy);
    a.heavy_cap = GP_NOALLOC_HEAVY;
    a.scratch = reinterpret_cast<int32_t*>(base + off_scr);
    a.slot_stride = ln.slot_stride;
    a.wave_stride = ln.wave_stride;
    a.tmp_stride = ln.tmp_stride;
    a.bslot_stride = ln.bslot_stride;
    a.hash_slots = ln.hash_slots;
    a.hash_lg = ln.hash_lg;
    a.probe_pct = t.probe_pct > 0 ? t.probe_pct : 100;
    a.bitmap_ratio = t.bitmap_ratio == 0 ? GP_BITMAP_RATIO : std::max(0, t.bitmap_ratio);
    a.merge_ratio = 0; // the plain instantiation
    a.dense = nullptr;
    a.dense_stride = 0;
    unsigned long long* clk = reinterpret_cast<unsigned long long*>(base + off_clk);
    a.clocks = clk;
    const KPlan& kp = ln.kps[0];
    const int ext = (!ln.faithful && kp.batched && kp.bl_base >= 0 && g->v_cnt < GP_FLAG_VMAX) ? GP_EXT_SEARCH
                                                                                              : GP_EXT_NONE;
    const bool parts = ln.any_split && ln.n_tasks > 0;
    unsigned long long host_out[6] = {0, 0, 0, 0, 0, 0};
    hipError_t e = hipMemcpyAsync(base, meta, meta_bytes, hipMemcpyHostToDevice, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, 0, cnt, (int)n_cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess && ln.n_tasks > 0) {
        void* kargs[] = {&a};
        e = hipLaunchKernel(count_kernel_fn(ln.faithful, false, false, ext, false), dim3((unsigned)waves), dim3(WAVE),
                            kargs, ln.lds_bytes, 0);
        if (e == hipSuccess && parts) {
            KArgs a2 = a;
            a2.clocks = clk + 2 * waves;
            void* pargs[] = {&a2};
            e = hipLaunchKernel(count_kernel_fn(false, true, false, ext, false), dim3((unsigned)waves), dim3(WAVE), pargs,
                                ln.lds_bytes, 0);
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, cnt, sizeof(host_out), hipMemcpyDeviceToHost, 0);
    std::vector<unsigned long long> clocks(ln.n_tasks > 0 ? 4 * (size_t)waves : 0);
    if (e == hipSuccess && ln.n_tasks > 0)
        e = hipMemcpyAsync(clocks.data(), clk, sizeof(unsigned long long) * (parts ? 4 : 2) * (size_t)waves,
                           hipMemcpyDeviceToHost, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(0);
    if (e != hipSuccess) return gp_fail((int)e, std::string("gp_count_noalloc: ") + hipGetErrorString(e));
    const uint64_t sum = host_out[3];
    if (raw_sum) *raw_sum = sum;
    if (count) *count = (int64_t)sum / (sched->redundancy ? sched->redundancy : 1);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->balg_bytes = ln.faithful ? host_out[4] + (kp.root_mode ? 8ull * ln.task_cnt[0]
                                                                      : ln.task_first[0] == 0 ? 8ull * (uint64_t)g->v_cnt : 0)
                                        : 0;
        stats->tasks = host_out[5];
        stats->waves = (int32_t)waves;
        stats->kernel_variant = ln.faithful ? 1 : 0;
        stats->split_tasks = parts ? std::min<uint64_t>(host_out[2], GP_NOALLOC_HEAVY) : 0;
        stats->inplace_tasks = parts && host_out[2] > GP_NOALLOC_HEAVY ? host_out[2] - GP_NOALLOC_HEAVY : 0;
        stats->launches = 1;
        stats->finished = 1;
        stats->tasks_total = ln.task_cnt[0];
        if (ln.n_tasks > 0) {
            // kernel time from the waves' wall-clock stamps (no events: creating
            // one is a runtime allocation)
            int rate_khz = 0;
            if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess ||
                rate_khz <= 0)
                rate_khz = 100000;
            unsigned long long t0 = ~0ull, t1 = 0;
            for (int64_t w = 0; w < waves; ++w) {
                t0 = std::min(t0, clocks[2 * w]);
                unsigned long long f = clocks[2 * w + 1];
                if (parts) f = std::max(f, clocks[2 * waves + 2 * w + 1]);
                t1 = std::max(t1, f);
            }
            stats->kernel_ms = t1 > t0 ? (double)(t1 - t0) / rate_khz : 0.0;
            stats->finish_max_ms = stats->kernel_ms;
        }
    }
    return GP_OK;
}

extern "C" int gp_count_batch(const gp_graph* g, const gp_sched_flat* scheds, int32_t n_scheds,
                              const gp_tuning* tuning, uint64_t* raw_sums, int64_t* counts, gp_stats* stats) {
    if (!g || !scheds) return gp_fail(GP_EINVAL, "null argument");
    if (n_scheds < 1 || n_scheds > GP_MAX_BATCH) return gp_fail(GP_ERANGE, "schedule count out of range");
    if (n_scheds > 1 && tuning && (tuning->task_begin != 0 || tuning->task_end > 0))
        return gp_fail(GP_EINVAL, "task ranges apply to a single schedule");
    // One launch per LDS class (pair level / batched level with few or many
    // leaf blocks / neither): plans of one class share the launch's LDS size
    // without shrinking the residency of cheaper plans.
    const bool faithful = tuning && tuning->collect_balg;
    std::vector<std::vector<int> > groups(6);
    for (int i = 0; i < n_scheds; ++i) {
        KPlan kp;
        int rc = gp_make_kplan(&scheds[i], faithful, !(tuning && tuning->no_batch), !(tuning && tuning->no_pair), &kp);
        if (rc != GP_OK) return rc;
        const int cls = kp.pair ? 0 : kp.batched ? (kp.nblocks <= 3 ? 1 : 2) : 3;
        groups[cls].push_back(i);
    }
    if (stats) std::memset(stats, 0, sizeof(*stats));
    for (const std::vector<int>& grp : groups) {
        if (grp.empty()) continue;
        std::vector<gp_sched_flat> sub;
        for (int i : grp) sub.push_back(scheds[i]);
        std::vector<uint64_t> raw(grp.size());
        std::vector<int64_t> cnt(grp.size());
        gp_stats st{};
        int rc = run_plans(const_cast<gp_graph*>(g), sub.data(), (int)sub.size(), tuning, raw.data(), cnt.data(), &st);
        if (rc != GP_OK) return rc;
        for (size_t j = 0; j < grp.size(); ++j) {
            if (raw_sums) raw_sums[grp[j]] = raw[j];
            if (counts) counts[grp[j]] = cnt[j];
        }
        if (stats) {
            stats->kernel_ms += st.kernel_ms;
            stats->balg_bytes += st.balg_bytes;
            stats->tasks += st.tasks;
            stats->split_tasks += st.split_tasks;
            stats->waves = std::max(stats->waves, st.waves);
            stats->kernel_variant = st.kernel_variant;
            // finish times of the last group's launch, offset by the groups before it
            stats->finish_p50_ms = stats->kernel_ms - st.kernel_ms + st.finish_p50_ms;
            stats->finish_p90_ms = stats->kernel_ms - st.kernel_ms + st.finish_p90_ms;
            stats->finish_max_ms = stats->kernel_ms - st.kernel_ms + st.finish_max_ms;
        }
    }
    return GP_OK;
}

// gp_count_batch over part of each schedule's tasks: launches over the
// schedules' remaining tasks laid end to end (one class of LDS size at a
// time, as in gp_count_batch), each with a kernel deadline of slice_ms (or
// what is left of the time limit), so that each counts a prefix of those
// tasks; the time limit is checked between launches.  After a launch every
// schedule's count is known up to a task index: the next launch, or the
// caller, resumes from there, and raw sums of consecutive ranges add up.
extern "C" int gp_count_batch_progress(const gp_graph* gc, const gp_sched_flat* scheds, int32_t n_scheds,
                                       const gp_tuning* tuning, const uint64_t* begin, const double* abort_s,
                                       uint64_t* done, uint64_t* raw_sums, int32_t* aborted, gp_stats* stats) {
    if (!gc || !scheds || !done || !raw_sums) return gp_fail(GP_EINVAL, "null argument");
    if (n_scheds < 1 || n_scheds > GP_MAX_BATCH) return gp_fail(GP_ERANGE, "schedule count out of range");
    gp_graph* g = const_cast<gp_graph*>(gc);
    gp_tuning t{};
    if (tuning) t = *tuning;
    if (t.task_begin != 0 || t.task_end > 0) return gp_fail(GP_EINVAL, "task ranges come from begin[]");
    if (t.collect_balg) return gp_fail(GP_EINVAL, "the faithful loop nest is not run in progress mode");
    const double t0 = wall_seconds();
    const int n = n_scheds;
    std::vector<uint64_t> total((size_t)n), b((size_t)n);
    std::vector<double> limit((size_t)n);     // abort time per schedule, seconds after t0 (<= 0: none)
    std::vector<int32_t> stopped((size_t)n, 0); // 1 once the schedule's abort time has passed
    std::vector<std::vector<int> > groups(4);
    for (int i = 0; i < n; ++i) {
        KPlan kp;
        int rc = gp_make_kplan(&scheds[i], false, !t.no_batch, !t.no_pair, &kp);
        if (rc != GP_OK) return rc;
        total[i] = kp.root_mode ? (uint64_t)g->v_cnt : (uint64_t)g->e_cnt;
        b[i] = begin ? std::min(begin[i], total[i]) : 0;
        done[i] = b[i];
        raw_sums[i] = 0;
        limit[i] = abort_s ? abort_s[i] : t.abort_s;
        groups[kp.pair ? 0 : kp.batched ? (kp.nblocks <= 3 ? 1 : 2) : 3].push_back(i);
    }
    gp_stats acc{};
    bool timed_out = false;
    const double slice = t.slice_ms > 0 ? (double)t.slice_ms : 2000.0;
    for (const std::vector<int>& grp : groups) {
        for (;;) {
            // the class's schedules that still have tasks and time, in order
            const double now = wall_seconds() - t0;
            std::vector<gp_sched_flat> sub;
            std::vector<uint64_t> rb, re;
            std::vector<double> ab_ms;
            std::vector<int> who;
            bool any_limit = false;
            for (int i : grp) {
                if (done[i] >= total[i] || stopped[i]) continue;
                if (limit[i] > 0 && now >= limit[i]) {
                    stopped[i] = 1;
                    continue;
                }
                who.push_back(i);
                sub.push_back(scheds[i]);
                rb.push_back(done[i]);
                re.push_back(total[i]);
                // from the start of the launch about to begin
                ab_ms.push_back(limit[i] > 0 ? (limit[i] - now) * 1000.0 : 0.0);
                any_limit = any_limit || limit[i] > 0;
            }
            if (who.empty()) break;
            double left_ms = slice;
            if (t.time_limit_s > 0) {
                left_ms = (t.time_limit_s - now) * 1000.0;
                if (left_ms <= 0) {
                    timed_out = true;
                    break;
                }
                left_ms = std::min(left_ms, slice);
            }
            std::vector<uint64_t> raw(sub.size()), did(sub.size());
            std::vector<int32_t> ab(sub.size(), 0);
            gp_stats st{};
            int rc = run_plans(g, sub.data(), (int)sub.size(), &t, raw.data(), nullptr, &st, rb.data(), re.data(),
                               left_ms, did.data(), any_limit ? ab_ms.data() : nullptr, ab.data(),
                               t.task_abort_s > 0 ? t.task_abort_s * 1000.0 : 0.0);
            if (rc != GP_OK) return rc;
            for (size_t k = 0; k < who.size(); ++k) {
                if (ab[k]) { // abandoned: its launches before this one stand
                    stopped[who[k]] = 1;
                    continue;
                }
                raw_sums[who[k]] += raw[k];
                done[who[k]] += did[k];
            }
            const double before_ms = acc.kernel_ms;
            acc.kernel_ms += st.kernel_ms;
            acc.tasks += st.tasks;
            acc.split_tasks += st.split_tasks;
            acc.inplace_tasks += st.inplace_tasks;
            acc.waves = std::max(acc.waves, st.waves);
            acc.finish_p50_ms = before_ms + st.finish_p50_ms;
            acc.finish_p90_ms = before_ms + st.finish_p90_ms;
            acc.finish_max_ms = before_ms + st.finish_max_ms;
            ++acc.launches;
        }
        if (timed_out) break;
    }
    bool all_done = true;
    for (int i = 0; i < n; ++i) {
        acc.tasks_total += total[i] - b[i];
        all_done = all_done && done[i] == total[i];
        if (aborted) aborted[i] = stopped[i];
    }
    acc.finished = all_done ? 1 : 0;
    if (stats) *stats = acc;
    if (!all_done)
        return gp_fail(GP_ETIME, "the time limit or an abort time came before every schedule's count was complete");
    return GP_OK;
}

extern "C" int gp_count_footprint(const gp_graph* g, const gp_sched_flat* scheds, int32_t n_scheds,
                                  const gp_tuning* tuning, uint64_t* bytes) {
    if (!g || !scheds || !bytes) return gp_fail(GP_EINVAL, "null argument");
    Launch ln;
    int rc = prepare(g, scheds, n_scheds, tuning, &ln);
    if (rc != GP_OK) return rc;
    *bytes = ln.device_bytes();
    return GP_OK;
}

extern "C" int gp_count_tasks(const gp_graph* g, const gp_sched_flat* sched, uint64_t* tasks) {
    if (!g || !sched || !tasks) return gp_fail(GP_EINVAL, "null argument");
    KPlan kp;
    int rc = gp_make_kplan(sched, false, true, true, &kp);
    if (rc != GP_OK) return rc;
    *tasks = kp.root_mode ? (uint64_t)g->v_cnt : (uint64_t)g->e_cnt;
    return GP_OK;
}

extern "C" int gp_tc(const gp_graph* g, uint64_t* triangles, gp_stats* stats) {
    if (!g || !triangles) return gp_fail(GP_EINVAL, "null argument");
    // triangle with v1 < v0 and v2 < v1: each triangle once
    const int adj[9] = {0, 1, 1, 1, 0, 1, 1, 1, 0};
    Schedule s(adj, 3);
    s.add_restrict({{0, 1}, {1, 2}});
    gp_sched_flat flat;
    int rc = s.flatten(&flat);
    if (rc != GP_OK) return gp_fail(rc, "triangle schedule");
    uint64_t raw = 0;
    rc = run_plans(const_cast<gp_graph*>(g), &flat, 1, nullptr, &raw, nullptr, stats);
    if (rc != GP_OK) return rc;
    *triangles = raw;
    return GP_OK;
}
