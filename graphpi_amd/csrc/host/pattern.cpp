// Pattern and MotifGenerator (interfaces of the reference's include/pattern.h
// and include/motif_generator.h; behaviour of src/pattern.cpp:9-263 and
// src/motif_generator.cpp:5-44).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <unordered_set>
#include <vector>

#include "motif_generator.h"
#include "pattern.h"

Pattern::Pattern(int _size, bool clique) : adj_mat(nullptr), size(_size) {
    adj_mat = new int[size * size]();
    if (clique)
        for (int i = 0; i < size; ++i)
            for (int j = 0; j < i; ++j) add_edge(i, j);
}

Pattern::Pattern(int _size, const char* buffer) : adj_mat(nullptr), size(_size) {
    adj_mat = new int[size * size]();
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            if (buffer[INDEX(i, j, size)] == '1') add_edge(i, j);
}

Pattern::~Pattern() { delete[] adj_mat; }

Pattern::Pattern(const Pattern& p) : adj_mat(nullptr), size(p.size) {
    adj_mat = new int[size * size];
    std::memcpy(adj_mat, p.adj_mat, sizeof(int) * size * size);
}

// Preset edge lists (src/pattern.cpp:45-133).  The vertex numbering matters:
// it is the order the planner starts from, so it is kept as the reference's.
Pattern::Pattern(PatternType type) : adj_mat(nullptr), size(0) {
    static const int rect[][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 3}};
    static const int qg3[][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}};
    static const int penta[][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 4}, {3, 4}};
    static const int house[][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 3}, {1, 4}, {2, 4}};
    static const int hourglass[][2] = {{0, 1}, {0, 2}, {0, 4}, {1, 2}, {1, 5},
                                       {2, 3}, {3, 4}, {3, 5}, {4, 5}};
    static const int c6tri[][2] = {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {1, 4},
                                   {2, 3}, {2, 5}, {0, 4}, {0, 5}};
    const int(*edges)[2] = nullptr;
    int m = 0;
    switch (type) {
        case Rectangle: size = 4; edges = rect; m = 4; break;
        case QG3: size = 4; edges = qg3; m = 5; break;
        case Pentagon: size = 5; edges = penta; m = 5; break;
        case House: size = 5; edges = house; m = 6; break;
        case Hourglass: size = 6; edges = hourglass; m = 9; break;
        case Cycle_6_Tri: size = 6; edges = c6tri; m = 9; break;
        case Clique_7_Minus: size = 7; break;
        default:
            std::printf("invalid pattern type!\n");
            std::abort();
    }
    adj_mat = new int[size * size]();
    if (type == Clique_7_Minus) {
        // K7 without the edge (5,6)
        for (int i = 0; i < size; ++i)
            for (int j = 0; j < i; ++j)
                if (!(i == size - 1 && j == size - 2)) add_edge(j, i);
        return;
    }
    for (int e = 0; e < m; ++e) add_edge(edges[e][0], edges[e][1]);
}

void Pattern::add_edge(int x, int y) {
    adj_mat[INDEX(x, y, size)] = 1;
    adj_mat[INDEX(y, x, size)] = 1;
}

void Pattern::del_edge(int x, int y) {
    adj_mat[INDEX(x, y, size)] = 0;
    adj_mat[INDEX(y, x, size)] = 0;
}

bool Pattern::check_connected() const {
    if (size <= 0) return true;
    std::vector<char> seen(size, 0);
    std::vector<int> stack{0};
    seen[0] = 1;
    int reached = 1;
    while (!stack.empty()) {
        int x = stack.back();
        stack.pop_back();
        for (int y = 0; y < size; ++y)
            if (adj_mat[INDEX(x, y, size)] && !seen[y]) {
                seen[y] = 1;
                ++reached;
                stack.push_back(y);
            }
    }
    return reached == size;
}

void Pattern::count_all_isomorphism(std::set<std::set<int> >& s) const {
    std::vector<int> perm(size);
    for (int i = 0; i < size; ++i) perm[i] = i;
    do {
        std::set<int> edges;
        for (int i = 0; i < size; ++i)
            for (int j = i + 1; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)]) {
                    int a = std::min(perm[i], perm[j]), b = std::max(perm[i], perm[j]);
                    edges.insert(a * size + b);
                }
        s.insert(edges);
    } while (std::next_permutation(perm.begin(), perm.end()));
}

void Pattern::print() const {
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            if (adj_mat[INDEX(i, j, size)]) std::printf("(%d,%d) ", i, j);
    std::printf("\n");
}

bool Pattern::is_dag() const {
    // peel vertices with no incoming arc; a cycle leaves some unpeeled
    std::vector<int> indeg(size, 0);
    for (int x = 0; x < size; ++x)
        for (int y = 0; y < size; ++y)
            if (adj_mat[INDEX(x, y, size)] > 0) ++indeg[y];
    std::vector<int> ready;
    for (int v = 0; v < size; ++v)
        if (indeg[v] == 0) ready.push_back(v);
    int peeled = 0;
    while (!ready.empty()) {
        int x = ready.back();
        ready.pop_back();
        ++peeled;
        for (int y = 0; y < size; ++y)
            if (adj_mat[INDEX(x, y, size)] > 0 && --indeg[y] == 0) ready.push_back(y);
    }
    return peeled == size;
}

// ---------------------------------------------------------------------------
// MotifGenerator: the reference walks the vertex pairs (0,1),(0,2),...,(n-2,n-1)
// deciding "edge present" before "edge absent" and keeps the first connected
// edge set of each isomorphism class (src/motif_generator.cpp:5-37).  That walk
// visits edge subsets in decreasing order of the bitmask whose most significant
// bit is pair (0,1), which is how it is enumerated here; each new class marks
// all of its relabellings as seen.
// ---------------------------------------------------------------------------
std::vector<Pattern> MotifGenerator::generate() {
    std::vector<Pattern> out;
    const int n = size;
    if (n < 1) return out;
    std::vector<std::pair<int, int> > pairs;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) pairs.push_back({i, j});
    const int m = (int)pairs.size();
    if (m == 0) {
        out.push_back(Pattern(n));
        return out;
    }
    // pair index k <-> bit (m-1-k); masks over at most 28 pairs (n <= 8)
    std::unordered_set<uint32_t> classes;
    std::vector<int> perm(n);
    for (int64_t mask = (1ll << m) - 1; mask >= 0; --mask) {
        uint32_t code = (uint32_t)mask;
        if (classes.count(code)) continue;
        Pattern p(n);
        for (int k = 0; k < m; ++k)
            if (mask >> (m - 1 - k) & 1) p.add_edge(pairs[k].first, pairs[k].second);
        if (!p.check_connected()) continue;
        out.push_back(p);
        for (int i = 0; i < n; ++i) perm[i] = i;
        do {
            uint32_t c = 0;
            for (int k = 0; k < m; ++k)
                if (mask >> (m - 1 - k) & 1) {
                    int a = perm[pairs[k].first], b = perm[pairs[k].second];
                    if (a > b) std::swap(a, b);
                    // index of pair (a,b) in the walk order
                    int idx = a * n - a * (a + 1) / 2 + (b - a - 1);
                    c |= 1u << (m - 1 - idx);
                }
            classes.insert(c);
        } while (std::next_permutation(perm.begin(), perm.end()));
    }
    return out;
}
