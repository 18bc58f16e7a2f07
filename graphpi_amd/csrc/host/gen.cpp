// Synthetic graph generators (host).  Both produce an undirected simple graph
// and hand it to the CSR builder exactly as if it had been written as a
// GraphPi text file (one line "u v" per edge, u < v, lines sorted by (u, v),
// ids compacted to the vertices that appear) and read back by the loader:
// ids are then renumbered by first appearance in that line order.
//
// Random numbers: counter-based SplitMix64, so every sample depends only on
// (seed, sample index) and the output is identical for any thread count.
//
// R-MAT (gen = 0): 2^scale * edge_factor samples; each sample descends scale
// levels choosing quadrant (0,0),(0,1),(1,0),(1,1) with probabilities
// a, b, c, 1-a-b-c from one uniform double in [0,1) per level.
//
// Chung-Lu (gen = 1): n vertices with expected degrees w_i drawn from a
// Pareto law with tail exponent alpha (w = u^(-1/(alpha-1)), u in (0,1]),
// scaled so that sum(w) = 2 * target_edges and capped at max_expected_degree
// (the scale is recomputed after capping, 8 rounds); then samples, each
// endpoint drawn independently with probability w_i / sum(w) (inverse CDF by
// binary search), self-loops and duplicates dropped, in rounds that continue
// the sample stream until at least target_edges distinct edges exist (the
// first round draws target_edges samples, each later one the shortfall plus
// an eighth).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../gp_internal.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// undirected keys (min<<32|max) -> CSR through the text-file semantics
int finish(std::vector<uint64_t>& und, gp_csr** out) {
    gp_sort_unique(und);
    // compact ids to the vertices that appear, keeping their order
    uint32_t maxv = 0;
    for (uint64_t k : und) maxv = std::max<uint32_t>(maxv, (uint32_t)k);
    std::vector<uint32_t> present((size_t)maxv + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)und.size(); ++i) {
        present[und[i] >> 32] = 1;
        present[(uint32_t)und[i]] = 1;
    }
    std::vector<uint32_t> compact((size_t)maxv + 1, 0);
    uint32_t n = 0;
    for (size_t v = 0; v <= maxv; ++v)
        if (present[v]) compact[v] = n++;
    // first appearance over lines (u, v) in sorted order; compaction keeps
    // the order, so relabel the compacted ids directly
    std::vector<int64_t> first_id((size_t)n, -1);
    uint32_t next_id = 0;
    for (uint64_t k : und) {
        uint32_t u = compact[k >> 32], v = compact[(uint32_t)k];
        if (first_id[u] < 0) first_id[u] = next_id++;
        if (first_id[v] < 0) first_id[v] = next_id++;
    }
    std::vector<uint64_t> pairs(2 * und.size());
    std::vector<int64_t> degree((size_t)n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)und.size(); ++i) {
        uint64_t a = (uint64_t)first_id[compact[und[i] >> 32]];
        uint64_t b = (uint64_t)first_id[compact[(uint32_t)und[i]]];
        pairs[2 * i] = a << 32 | b;
        pairs[2 * i + 1] = b << 32 | a;
    }
    for (uint64_t k : pairs) ++degree[k >> 32];
    und.clear();
    und.shrink_to_fit();
    return gp_build_csr(pairs, (int32_t)n, degree, out);
}

int rmat(int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed, gp_csr** out) {
    if (scale < 1 || scale > 30 || edge_factor < 1) return gp_fail(GP_EINVAL, "bad R-MAT size");
    const int64_t m = edge_factor << scale;
    const uint64_t base = splitmix64(seed ^ 0x5DEECE66Dull);
    std::vector<uint64_t> und((size_t)m);
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        uint64_t state = base + (uint64_t)i * (uint64_t)scale;
        uint32_t u = 0, v = 0;
        for (int lvl = 0; lvl < scale; ++lvl) {
            double r = u01(splitmix64(state + (uint64_t)lvl));
            uint32_t bit = 1u << (scale - 1 - lvl);
            if (r < a) {
            } else if (r < ab) {
                v |= bit;
            } else if (r < abc) {
                u |= bit;
            } else {
                u |= bit;
                v |= bit;
            }
        }
        if (u == v)
            und[i] = ~0ull;
        else
            und[i] = (uint64_t)std::min(u, v) << 32 | std::max(u, v);
    }
    und.erase(std::remove(und.begin(), und.end(), ~0ull), und.end());
    return finish(und, out);
}

int chung_lu(int64_t n, int64_t target_edges, double alpha, double max_w, uint64_t seed, gp_csr** out) {
    if (n < 2 || target_edges < 1 || alpha <= 1.0 || max_w < 1) return gp_fail(GP_EINVAL, "bad Chung-Lu parameters");
    const uint64_t base = splitmix64(seed ^ 0xC0FFEEull);
    std::vector<double> w((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        double u = 1.0 - u01(splitmix64(base + (uint64_t)i)); // (0, 1]
        w[i] = std::pow(u, -1.0 / (alpha - 1.0));
    }
    // scale to the target total, capping at max_w
    double target = 2.0 * (double)target_edges;
    for (int round = 0; round < 8; ++round) {
        double s = 0;
        for (double x : w) s += x;
        double f = target / s;
        for (double& x : w) x = std::min(x * f, max_w);
    }
    std::vector<double> cdf((size_t)n);
    double s = 0;
    for (int64_t i = 0; i < n; ++i) {
        s += w[i];
        cdf[i] = s;
    }
    const uint64_t base2 = splitmix64(seed ^ 0xFEEDull);
    std::vector<uint64_t> und;
    uint64_t drawn = 0;  // samples taken from the stream so far
    for (int round = 0; round < 64 && und.size() < (size_t)target_edges; ++round) {
        const uint64_t short_by = (uint64_t)target_edges - und.size();
        const uint64_t m = round == 0 ? short_by : short_by + short_by / 8 + 16;
        const size_t old = und.size();
        und.resize(old + (size_t)m);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < (int64_t)m; ++j) {
            const uint64_t i = drawn + (uint64_t)j;
            double r1 = u01(splitmix64(base2 + 2 * i)) * s;
            double r2 = u01(splitmix64(base2 + 2 * i + 1)) * s;
            uint32_t u = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), r1) - cdf.begin());
            uint32_t v = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), r2) - cdf.begin());
            if (u >= (uint32_t)n) u = (uint32_t)n - 1;
            if (v >= (uint32_t)n) v = (uint32_t)n - 1;
            und[old + (size_t)j] = u == v ? ~0ull : ((uint64_t)std::min(u, v) << 32 | std::max(u, v));
        }
        drawn += m;
        gp_sort_unique(und);
        if (!und.empty() && und.back() == ~0ull) und.pop_back(); // the self-loop marker sorts last
    }
    return finish(und, out);
}

} // namespace

extern "C" int gp_csr_generate(int32_t gen, int64_t p0, int64_t p1, double a, double b, double c,
                               uint64_t seed, gp_csr** out) {
    if (!out) return gp_fail(GP_EINVAL, "null argument");
    if (gen == 0) return rmat((int)p0, p1, a, b, c, seed, out);
    if (gen == 1) return chung_lu(p0, p1, a, b, seed, out);
    return gp_fail(GP_EINVAL, "unknown generator");
}
