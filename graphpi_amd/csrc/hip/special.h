// Launch interfaces of the schedule-specialised kernels (special.hip): House
// counted per root with two-hop counts, the edge-support kernel both shapes
// read, and the per-edge sum of the "two triangles joined by an edge" closed
// form.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct GpHouseArgs {
    const uint32_t* vertex;
    const int32_t* edge;
    const int32_t* sup;        // per edge slot: |N(src) ∩ N(dst)|
    int32_t v_cnt;
    uint64_t slot_begin;       // task range: edge slots [slot_begin, slot_end)
    uint64_t slot_end;
    int32_t all_v1;            // 1: every v1 in N(v0) (no restriction v1 < v0)
    unsigned long long* counter; // row fetch counter, zero at launch
    int32_t* dense;            // per-wave count arrays over all vertex ids, zero at launch (left zero)
    int64_t dense_stride;
    unsigned long long* out;   // [0] raw sum, [2] tasks taken
    unsigned long long* clocks; // per wave: start, finish
};

hipError_t gp_house_launch(const GpHouseArgs& args, unsigned waves, hipStream_t stream);
hipError_t gp_support_launch(const uint32_t* vertex, const int32_t* edge, const int32_t* src, uint64_t e_cnt,
                             int32_t* sup, unsigned long long* counter, unsigned waves, hipStream_t stream);

// t[v] = triangles containing v = (sum of sup over v's edge slots) / 2
hipError_t gp_vertex_triangles_launch(const uint32_t* vertex, const int32_t* sup, int32_t v_cnt, uint64_t* t,
                                      hipStream_t stream);
// out[0] += sum over edges {a, b} (a > b) of (t_a - s)(t_b - s) - sum_{v in N(a) ∩ N(b)} (s_av - 1)(s_bv - 1),
// s = sup(a, b), in u64 (two's-complement wrap-around); out128[2w], out128[2w+1] = wave w's share of the
// same sum in 128 bits (low, high word; signed)
hipError_t gp_tri_edge_launch(const uint32_t* vertex, const int32_t* edge, const int32_t* src, const int32_t* sup,
                              const uint64_t* t, uint64_t e_cnt, unsigned long long* counter, unsigned long long* out,
                              unsigned long long* out128, unsigned waves, hipStream_t stream);
