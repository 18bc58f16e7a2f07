// Schedule-specialised kernels.  Each computes exactly the raw sum the
// reference's loop nest computes for one schedule shape, by a different
// grouping of the same integer sum:
//   - house_twohop_kernel: House under the GraphPi order, per root with
//     two-hop counts (below);
//   - support_kernel: sup(e) = |N(u) ∩ N(v)| for every edge slot (u, v);
//   - vertex_triangles_kernel and tri_edge_kernel: the closed form of "two
//     triangles joined by an edge" (the host adds 6 x the 4-clique count;
//     see run_tri_edge_tri in count.hip).
//
// House (5 vertices, 6 edges) under the GraphPi schedule's vertex order,
// counted per root with two-hop counts.
//
// The reference evaluates House (src/graph.cpp:247-398) as: v0; v1 in N(v0)
// with v1 < v0; v2 in N(v0); then an IEP leaf over v3 in A = N(v0) ∩ N(v1)
// and v4 in B = N(v1) ∩ N(v2) (src/graph.cpp:306-344), adding per v2
//     val(v2) = |A \ emb| |B \ emb| - |(A ∩ B) \ emb|,   emb = {v0, v1, v2}.
// With no self-loops, v0, v1 are never in A, v0 is always in B, v1 and v2
// never are, so with a = |A|, c(u, w) = |N(u) ∩ N(w)|:
//     val(v2) = (a - [v2 in A]) (c(v1, v2) - 1) - |A ∩ N(v2)|.
// Summed over v2 in N(v0) \ {v1} (d0 = |N(v0)|, d1 = |N(v1)|):
//     sum = a (S1 - d1 - d0 + 3) - SA1 - SA0,
//     S1  = sum_{y in N(v1)} t0[y],          t0[y] = |N(y) ∩ N(v0)|,
//     SA0 = sum_{x in A} t0[x],
//     SA1 = sum_{x in A} sup(v1, x),        sup(u, w) = |N(u) ∩ N(w)| per edge.
// (sum_{v2 in N(v0)} c(v1, v2) = sum_{y in N(v1)} |N(y) ∩ N(v0)|;
// sum_{v2 in N(v0)\{v1}} |A ∩ N(v2)| = SA0 - a; sum_{v2 in A} (c(v1,v2) - 1)
// = SA1 - a.)  The regrouping is an integer identity, so the u64 sum equals
// the reference's long long sum bit for bit (two's-complement wrap-around).
//
// One wave takes one root v0 at a time: it scatters t0 over its private
// dense array (every N(w), w in N(v0): W(v0) increments), flags the members
// of N(v0) in bit 31, then for every task (v0, v1), v1 < v0, streams N(v1)
// and its slice of the edge-support array once, gathering t0.  The array is
// cleared again before the next root.  Per root the work is 2 W(v0) + the
// sum of d1 over its tasks, against d_low(v0) W(v0) for the per-task leaf
// level (on R-MAT scale 22: 4.1e11 against 3.9e13 elements).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "special.h"

#define WAVE 64
#define GP_UNROLL 4
#define T0_FLAG 0x80000000u
#define T0_MASK 0x7fffffffu

namespace {

typedef const int32_t __attribute__((address_space(1)))* gcptr;
typedef int32_t __attribute__((address_space(1)))* gptr;

__device__ __forceinline__ int lane_id() { return (int)threadIdx.x; }

__device__ __forceinline__ uint64_t first_lane_u64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (uint64_t)hi << 32 | lo;
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

// first index with a[i] >= x (wave-uniform result)
__device__ int32_t wave_lower_bound(gcptr a, int32_t n, int32_t x) {
    const int lane = lane_id();
    int32_t lo = 0, hi = n;
    while (hi - lo > WAVE) {
        const int32_t step = (hi - lo + WAVE - 1) / WAVE;
        const int32_t pos = lo + lane * step;
        const bool lt = pos < hi && a[pos] < x;
        const int c = __popcll(__ballot(lt));
        if (c == 0) return lo;
        const int32_t nlo = lo + (c - 1) * step + 1;
        const int32_t nhi = min(hi, lo + c * step);
        lo = nlo;
        hi = nhi;
    }
    const int32_t pos = lo + lane;
    const bool lt = pos < hi && a[pos] < x;
    return lo + __popcll(__ballot(lt));
}

// last row r with vertex[r] <= slot (slot < e_cnt)
__device__ int32_t row_of(const uint32_t* __restrict__ vertex, int32_t v_cnt, uint64_t slot) {
    int32_t lo = 0, hi = v_cnt;
    while (hi - lo > 1) {
        const int32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)vertex[mid] <= slot)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Calls f(x) for every x in N(u) of each lane's vertex u (u < 0: none).
// Lists shorter than a wave are streamed flattened across the lanes; longer
// ones by the whole wave, GP_UNROLL loads in flight per lane.
template <typename F>
__device__ __forceinline__ void stream_lists(const uint32_t* __restrict__ vertex, gcptr E, int32_t u, int32_t* s_off,
                                             F&& f) {
    const int lane = lane_id();
    int32_t us = 0, un = 0;
    if (u >= 0) {
        const uint32_t l = vertex[u];
        us = (int32_t)l;
        un = (int32_t)(vertex[u + 1] - l);
    }
    const bool lng = un >= WAVE;
    int32_t tot = 0;
    const int32_t ex = wave_excl_scan(lng ? 0 : un, &tot);
    s_off[lane] = ex;
    __syncthreads();
    for (int32_t f0 = 0; f0 < tot; f0 += WAVE) {
        const int32_t fi = f0 + lane;
        const bool act = fi < tot;
        int lo = 0, hi = WAVE; // owner: last lane whose offset is <= fi
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= fi) lo = mid; else hi = mid;
        }
        const int32_t os = __shfl(us, lo);
        const int32_t oe = __shfl(ex, lo);
        if (act) f(E[os + (fi - oe)]);
    }
    uint64_t lm = __ballot(lng);
    while (lm) {
        const int j = __ffsll((long long)lm) - 1;
        lm &= lm - 1;
        const gcptr pv = E + __shfl(us, j);
        const int32_t pn = __shfl(un, j);
        for (int32_t f0 = 0; f0 < pn; f0 += GP_UNROLL * WAVE) {
            int32_t x[GP_UNROLL];
#pragma unroll
            for (int q = 0; q < GP_UNROLL; ++q) x[q] = f0 + q * WAVE + lane < pn ? pv[f0 + q * WAVE + lane] : -1;
#pragma unroll
            for (int q = 0; q < GP_UNROLL; ++q)
                if (x[q] >= 0) f(x[q]);
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(WAVE) house_twohop_kernel(GpHouseArgs a) {
    __shared__ int32_t s_off[WAVE];
    const int lane = lane_id();
    if (lane == 0) a.clocks[2 * blockIdx.x] = wall_clock64();
    const gcptr E = (gcptr)a.edge;
    const gcptr SUP = (gcptr)a.sup;
    int32_t* const t0 = a.dense + (int64_t)blockIdx.x * a.dense_stride;
    const uint64_t sb = a.slot_begin, se = a.slot_end;
    uint64_t acc = 0;      // sum over tasks of a (S1 - d1 - d0 + 3), wave-uniform
    uint64_t lane_sub = 0; // per-lane part of SA0 + SA1
    uint64_t taken = 0;
    if (se > sb) {
        const int32_t row_lo = row_of(a.vertex, a.v_cnt, sb);
        const int32_t row_hi = row_of(a.vertex, a.v_cnt, se - 1) + 1;
        for (;;) {
            unsigned long long w = 0;
            if (lane == 0) w = atomicAdd(a.counter, 1ull);
            w = first_lane_u64(w);
            if (w >= (unsigned long long)(row_hi - row_lo)) break;
            const int32_t v0 = row_lo + (int32_t)w;
            const uint32_t l0 = a.vertex[v0];
            const int32_t d0 = (int32_t)(a.vertex[v0 + 1] - l0);
            const uint64_t rb = max((uint64_t)l0, sb), re = min((uint64_t)l0 + (uint64_t)d0, se);
            if (rb >= re) continue;
            taken += re - rb;
            // tasks: slots of N(v0) below v0 (a prefix of the sorted row) within
            // the range, or the whole row without the restriction v1 < v0
            const int32_t low = a.all_v1 ? d0 : wave_lower_bound(E + l0, d0, v0);
            const uint64_t tb = rb, te = min(re, (uint64_t)l0 + (uint64_t)low);
            if (tb >= te) continue;
            // 1. t0[x] = |N(x) ∩ N(v0)|, then flag the members of N(v0)
            for (int32_t i0 = 0; i0 < d0; i0 += WAVE) {
                const int32_t i = i0 + lane;
                stream_lists(a.vertex, E, i < d0 ? E[l0 + i] : -1, s_off, [&](int32_t x) { atomicAdd(&t0[x], 1); });
            }
            __threadfence();
            for (int32_t i = lane; i < d0; i += WAVE) atomicOr((uint32_t*)&t0[E[l0 + i]], T0_FLAG);
            __threadfence();
            // 2. every task (v0, v1): one pass over N(v1) and its supports
            for (uint64_t t = tb; t < te; ++t) {
                const int32_t v1 = E[t];
                const uint32_t l1 = a.vertex[v1];
                const int32_t d1 = (int32_t)(a.vertex[v1 + 1] - l1);
                uint32_t cnt_a = 0;
                uint64_t s1 = 0;
                for (int32_t f0 = 0; f0 < d1; f0 += GP_UNROLL * WAVE) {
                    int32_t y[GP_UNROLL], sp[GP_UNROLL];
#pragma unroll
                    for (int q = 0; q < GP_UNROLL; ++q) {
                        const int32_t i = f0 + q * WAVE + lane;
                        y[q] = i < d1 ? E[l1 + i] : -1;
                        sp[q] = i < d1 ? SUP[l1 + i] : 0;
                    }
#pragma unroll
                    for (int q = 0; q < GP_UNROLL; ++q) {
                        if (y[q] < 0) continue;
                        const uint32_t c = (uint32_t)__hip_atomic_load(&t0[y[q]], __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT);
                        const uint32_t n = c & T0_MASK;
                        s1 += n;
                        if (c & T0_FLAG) {
                            ++cnt_a;
                            lane_sub += (uint64_t)n + (uint64_t)(uint32_t)sp[q];
                        }
                    }
                }
                const uint64_t av = wave_sum_u32(cnt_a);
                const uint64_t S1 = wave_sum_u64(s1);
                acc += av * (S1 - (uint64_t)d1 - (uint64_t)d0 + 3ull);
            }
            __threadfence();
            // 3. clear what was written
            for (int32_t i0 = 0; i0 < d0; i0 += WAVE) {
                const int32_t i = i0 + lane;
                stream_lists(a.vertex, E, i < d0 ? E[l0 + i] : -1, s_off, [&](int32_t x) { t0[x] = 0; });
            }
            __threadfence();
            for (int32_t i = lane; i < d0; i += WAVE) t0[E[l0 + i]] = 0;
            __threadfence();
        }
    }
    const uint64_t sub = wave_sum_u64(lane_sub);
    if (lane == 0) {
        const uint64_t total = acc - sub;
        if (total) atomicAdd(&a.out[0], (unsigned long long)total);
        if (taken) atomicAdd(&a.out[2], (unsigned long long)taken);
        a.clocks[2 * blockIdx.x + 1] = wall_clock64();
    }
}

// sup[e] = |N(src[e]) ∩ N(edge[e])|: one wave per edge slot in turn, lanes
// take 64 elements of the shorter list and binary-search the longer one
__global__ void __launch_bounds__(WAVE) support_kernel(const uint32_t* __restrict__ vertex, const int32_t* __restrict__ edge,
                                                       const int32_t* __restrict__ src, uint64_t e_cnt, int32_t* __restrict__ sup,
                                                       unsigned long long* counter) {
    const int lane = lane_id();
    const gcptr E = (gcptr)edge;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(counter, (unsigned long long)WAVE);
        base = first_lane_u64(base);
        if (base >= e_cnt) break;
        const uint64_t end = min((uint64_t)base + WAVE, e_cnt);
        // one slot per lane for pairs of short lists; the whole wave otherwise
        const uint64_t e = base + (uint64_t)lane;
        int32_t u = -1, v = -1, du = 0, dv = 0;
        uint32_t lu = 0, lv = 0;
        if (e < end) {
            u = src[e];
            v = E[e];
            lu = vertex[u];
            du = (int32_t)(vertex[u + 1] - lu);
            lv = vertex[v];
            dv = (int32_t)(vertex[v + 1] - lv);
        }
        const bool small = e < end && min(du, dv) <= 8;
        if (small) {
            // lane-private: each element of the shorter list searched in the longer
            gcptr S = E + (du <= dv ? lu : lv), L = E + (du <= dv ? lv : lu);
            const int32_t ns = min(du, dv), nl = max(du, dv);
            int32_t c = 0;
            for (int32_t i = 0; i < ns; ++i) {
                const int32_t x = S[i];
                int32_t lo = 0, len = nl;
                while (len > 0) {
                    const int32_t half = len >> 1;
                    const bool right = L[lo + half] < x;
                    lo = right ? lo + half + 1 : lo;
                    len = right ? len - half - 1 : half;
                }
                c += (lo < nl && L[lo] == x) ? 1 : 0;
            }
            sup[e] = c;
        }
        uint64_t big = __ballot(e < end && !small);
        while (big) {
            const int j = __ffsll((long long)big) - 1;
            big &= big - 1;
            const int32_t ju = __shfl(du, j), jv = __shfl(dv, j);
            const uint32_t jlu = (uint32_t)__shfl((int32_t)lu, j), jlv = (uint32_t)__shfl((int32_t)lv, j);
            gcptr S = E + (ju <= jv ? jlu : jlv), L = E + (ju <= jv ? jlv : jlu);
            const int32_t ns = min(ju, jv), nl = max(ju, jv);
            uint32_t c = 0;
            for (int32_t i0 = 0; i0 < ns; i0 += WAVE) {
                const int32_t i = i0 + lane;
                bool hit = false;
                if (i < ns) {
                    const int32_t x = S[i];
                    int32_t lo = 0, len = nl;
                    while (len > 0) {
                        const int32_t half = len >> 1;
                        const bool right = L[lo + half] < x;
                        lo = right ? lo + half + 1 : lo;
                        len = right ? len - half - 1 : half;
                    }
                    hit = lo < nl && L[lo] == x;
                }
                c += (uint32_t)__popcll(__ballot(hit));
            }
            if (lane == j) sup[e] = (int32_t)c;
        }
    }
}

__global__ void vertex_triangles_kernel(const uint32_t* __restrict__ vertex, const int32_t* __restrict__ sup,
                                        int32_t v_cnt, uint64_t* __restrict__ t) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < v_cnt;
         v += (int64_t)gridDim.x * blockDim.x) {
        uint64_t s = 0;
        for (uint32_t e = vertex[v]; e < vertex[v + 1]; ++e) s += (uint32_t)sup[e];
        t[v] = s / 2;
    }
}

// Per edge {a, b} (the slot with a = src > b = dst): (t_a - s)(t_b - s) minus
// the sum over common neighbours v of (sup(a, v) - 1)(sup(b, v) - 1).  The
// common neighbours come with their slots in both lists: the shorter list's
// elements are searched in the longer one, and the search's position is the
// slot.  Lanes take one edge each when the shorter list has at most 8
// elements, the whole wave otherwise.
__global__ void __launch_bounds__(WAVE) tri_edge_kernel(const uint32_t* __restrict__ vertex,
                                                        const int32_t* __restrict__ edge,
                                                        const int32_t* __restrict__ src,
                                                        const int32_t* __restrict__ sup,
                                                        const uint64_t* __restrict__ t, uint64_t e_cnt,
                                                        unsigned long long* counter, unsigned long long* out,
                                                        unsigned long long* out128) {
    __shared__ unsigned long long s_lo[WAVE], s_hi[WAVE];
    const int lane = lane_id();
    const gcptr E = (gcptr)edge;
    const gcptr SUP = (gcptr)sup;
    uint64_t acc = 0;    // per lane, u64 wrap-around
    __int128 exact = 0;  // per lane, the same sum in 128 bits (no wrap-around)
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(counter, (unsigned long long)WAVE);
        base = first_lane_u64(base);
        if (base >= e_cnt) break;
        const uint64_t end = min((uint64_t)base + WAVE, e_cnt);
        const uint64_t e = base + (uint64_t)lane;
        int32_t a = 0, b = 0, da = 0, db = 0;
        uint32_t la = 0, lb = 0;
        bool mine = false;
        if (e < end) {
            a = src[e];
            b = E[e];
            mine = a > b;
            if (mine) {
                la = vertex[a];
                da = (int32_t)(vertex[a + 1] - la);
                lb = vertex[b];
                db = (int32_t)(vertex[b + 1] - lb);
                const uint64_t s = (uint32_t)SUP[e];
                const uint64_t p = (t[a] - s) * (t[b] - s); // each factor < 2^32: exact in 64 bits
                acc += p;
                exact += (__int128)p;
            }
        }
        const bool small = mine && min(da, db) <= 8;
        if (small) {
            const bool a_short = da <= db;
            const uint32_t ls = a_short ? la : lb, ll = a_short ? lb : la;
            const int32_t ns = min(da, db), nl = max(da, db);
            for (int32_t i = 0; i < ns; ++i) {
                const int32_t x = E[ls + i];
                int32_t lo = 0, len = nl;
                while (len > 0) {
                    const int32_t half = len >> 1;
                    const bool right = E[ll + lo + half] < x;
                    lo = right ? lo + half + 1 : lo;
                    len = right ? len - half - 1 : half;
                }
                if (lo < nl && E[ll + lo] == x) {
                    const uint64_t q = (uint64_t)(uint32_t)(SUP[ls + i] - 1) * (uint64_t)(uint32_t)(SUP[ll + lo] - 1);
                    acc -= q;
                    exact -= (__int128)q;
                }
            }
        }
        uint64_t big = __ballot(mine && !small);
        while (big) {
            const int j = __ffsll((long long)big) - 1;
            big &= big - 1;
            const int32_t ja = __shfl(da, j), jb = __shfl(db, j);
            const uint32_t jla = (uint32_t)__shfl((int32_t)la, j), jlb = (uint32_t)__shfl((int32_t)lb, j);
            const bool a_short = ja <= jb;
            const uint32_t ls = a_short ? jla : jlb, ll = a_short ? jlb : jla;
            const int32_t ns = min(ja, jb), nl = max(ja, jb);
            for (int32_t i0 = 0; i0 < ns; i0 += WAVE) {
                const int32_t i = i0 + lane;
                if (i < ns) {
                    const int32_t x = E[ls + i];
                    int32_t lo = 0, len = nl;
                    while (len > 0) {
                        const int32_t half = len >> 1;
                        const bool right = E[ll + lo + half] < x;
                        lo = right ? lo + half + 1 : lo;
                        len = right ? len - half - 1 : half;
                    }
                    if (lo < nl && E[ll + lo] == x) {
                        const uint64_t q =
                            (uint64_t)(uint32_t)(SUP[ls + i] - 1) * (uint64_t)(uint32_t)(SUP[ll + lo] - 1);
                        acc -= q;
                        exact -= (__int128)q;
                    }
                }
            }
        }
    }
    const uint64_t tot = wave_sum_u64(acc);
    if (lane == 0 && tot) atomicAdd(&out[0], (unsigned long long)tot);
    // the 128-bit sum: lane 0 adds the lanes' values, one slot per wave
    s_lo[lane] = (unsigned long long)(unsigned __int128)exact;
    s_hi[lane] = (unsigned long long)((unsigned __int128)exact >> 64);
    __syncthreads();
    if (lane == 0) {
        unsigned __int128 w = 0;
        for (int i = 0; i < WAVE; ++i) w += ((unsigned __int128)s_hi[i] << 64) | s_lo[i];
        out128[2 * blockIdx.x] = (unsigned long long)w;
        out128[2 * blockIdx.x + 1] = (unsigned long long)(w >> 64);
    }
}

} // namespace

hipError_t gp_vertex_triangles_launch(const uint32_t* vertex, const int32_t* sup, int32_t v_cnt, uint64_t* t,
                                      hipStream_t stream) {
    const unsigned blocks = (unsigned)std::min<int64_t>(4096, ((int64_t)v_cnt + 255) / 256 + 1);
    hipLaunchKernelGGL(vertex_triangles_kernel, dim3(blocks), dim3(256), 0, stream, vertex, sup, v_cnt, t);
    return hipGetLastError();
}

hipError_t gp_tri_edge_launch(const uint32_t* vertex, const int32_t* edge, const int32_t* src, const int32_t* sup,
                              const uint64_t* t, uint64_t e_cnt, unsigned long long* counter, unsigned long long* out,
                              unsigned long long* out128, unsigned waves, hipStream_t stream) {
    hipLaunchKernelGGL(tri_edge_kernel, dim3(waves), dim3(WAVE), 0, stream, vertex, edge, src, sup, t, e_cnt, counter,
                       out, out128);
    return hipGetLastError();
}

hipError_t gp_house_launch(const GpHouseArgs& args, unsigned waves, hipStream_t stream) {
    void* kargs[] = {const_cast<GpHouseArgs*>(&args)};
    return hipLaunchKernel(reinterpret_cast<const void*>(&house_twohop_kernel), dim3(waves), dim3(WAVE), kargs, 0,
                           stream);
}

hipError_t gp_support_launch(const uint32_t* vertex, const int32_t* edge, const int32_t* src, uint64_t e_cnt,
                             int32_t* sup, unsigned long long* counter, unsigned waves, hipStream_t stream) {
    hipLaunchKernelGGL(support_kernel, dim3(waves), dim3(WAVE), 0, stream, vertex, edge, src, e_cnt, sup, counter);
    return hipGetLastError();
}
