// Graph input: GraphPi text files, the binary u32 CSR, complete graphs and
// in-memory edge lists.  Semantics follow the reference's DataLoader
// (src/dataloader.cpp:8-262): first-appearance id remap, self-loops dropped,
// header counts checked, optional degree reordering, duplicates removed,
// adjacency sorted ascending.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <parallel/algorithm>
#include <string>
#include <unordered_map>
#include <vector>

#include "../gp_internal.h"
#include "dataloader.h"
#include "graph.h"

void gp_sort_unique(std::vector<uint64_t>& keys) {
    __gnu_parallel::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
}

int gp_build_csr(std::vector<uint64_t>& pairs, int32_t v_cnt, const std::vector<int64_t>& degree,
                 gp_csr** out) {
    gp_sort_unique(pairs);
    if (pairs.size() > 0xFFFFFFFFull) return gp_fail(GP_ERANGE, "more than 2^32-1 directed edges");
    gp_csr* g = new (std::nothrow) gp_csr();
    if (!g) return gp_fail(GP_ENOMEM, "out of host memory");
    g->v_cnt = v_cnt;
    g->e_cnt = (uint32_t)pairs.size();
    g->edge = new (std::nothrow) int32_t[pairs.size() ? pairs.size() : 1];
    g->vertex = new (std::nothrow) uint32_t[(size_t)v_cnt + 1];
    if (!g->edge || !g->vertex) {
        gp_csr_free(g);
        return gp_fail(GP_ENOMEM, "out of host memory");
    }
    // vertex[v] = first position of v; vertices without edges point at the
    // next vertex's start
    std::vector<int64_t> first((size_t)v_cnt + 1, -1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)pairs.size(); ++i) {
        uint32_t u = (uint32_t)(pairs[i] >> 32);
        g->edge[i] = (int32_t)(uint32_t)pairs[i];
        if (i == 0 || (uint32_t)(pairs[i - 1] >> 32) != u) first[u] = i;
    }
    g->vertex[v_cnt] = g->e_cnt;
    for (int64_t v = (int64_t)v_cnt - 1; v >= 0; --v)
        g->vertex[v] = first[v] >= 0 ? (uint32_t)first[v] : g->vertex[v + 1];
    // degree statistics of the input lines (duplicates included)
    int64_t d1 = 0, d2 = 0;
    for (int64_t d : degree) {
        if (d > d1) {
            d2 = d1;
            d1 = d;
        } else if (d > d2) {
            d2 = d;
        }
    }
    g->max_degree = (int32_t)d1;
    g->max_intersection_size = (int32_t)d2;
    g->tri_cnt = -1;
    *out = g;
    return GP_OK;
}

// Checks that a CSR is safe to count on: vertex[0] == 0, offsets
// non-decreasing and ending at e_cnt, every neighbour id in [0, v_cnt), each
// list strictly ascending (sorted, no duplicates) and free of self-loops.
// The kernels binary-search the lists and index by neighbour ids, so a CSR
// that breaks any of these would read out of bounds or miscount.
int gp_csr_validate(const int32_t* edge, const uint32_t* vertex, int32_t v_cnt, uint32_t e_cnt) {
    if (v_cnt < 0) return gp_fail(GP_EINVAL, "negative vertex count");
    if (!vertex || (!edge && e_cnt)) return gp_fail(GP_EINVAL, "bad CSR arguments");
    if (vertex[0] != 0) return gp_fail(GP_EINVAL, "vertex[0] != 0");
    if (vertex[v_cnt] != e_cnt) return gp_fail(GP_EINVAL, "vertex[v_cnt] != e_cnt");
    int64_t bad_v = -1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : bad_v)
    for (int64_t v = 0; v < (int64_t)v_cnt; ++v) {
        const uint32_t l = vertex[v], r = vertex[v + 1];
        if (r < l || r > e_cnt) {
            bad_v = std::max(bad_v, v);
            continue;
        }
        int64_t prev = -1;
        for (uint32_t i = l; i < r; ++i) {
            const int64_t x = edge[i];
            if (x <= prev || x >= v_cnt || x == v) {
                bad_v = std::max(bad_v, v);
                break;
            }
            prev = x;
        }
    }
    if (bad_v >= 0)
        return gp_fail(GP_EINVAL, "malformed adjacency of vertex " + std::to_string(bad_v) +
                                      " (offsets must not decrease, ids must lie in [0, v_cnt), each list "
                                      "strictly ascending without self-loops)");
    return GP_OK;
}

namespace {

// whitespace-separated signed decimal integers
struct IntReader {
    const char* p;
    const char* end;
    bool next(long long* v) {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
        if (p >= end) return false;
        bool neg = false;
        if (*p == '-' || *p == '+') {
            neg = *p == '-';
            ++p;
        }
        if (p >= end || *p < '0' || *p > '9') {
            bad = true;
            return false;
        }
        long long x = 0;
        while (p < end && *p >= '0' && *p <= '9') x = x * 10 + (*p++ - '0');
        *v = neg ? -x : x;
        return true;
    }
    bool bad = false;
};

struct MappedFile {
    const char* data = nullptr;
    size_t size = 0;
    int fd = -1;
    bool open(const char* path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0) return false;
        size = (size_t)st.st_size;
        if (size == 0) return true;
        void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) return false;
        data = (const char*)m;
        return true;
    }
    ~MappedFile() {
        if (data) munmap((void*)data, size);
        if (fd >= 0) ::close(fd);
    }
};

// degree reordering (reference src/dataloader.cpp:96-109): rank vertices by
// degree with std::sort on (id, degree) pairs and renumber by rank
void reorder_by_degree(std::vector<uint64_t>& pairs, std::vector<int64_t>& degree, int32_t v_cnt,
                       int oriented_type) {
    std::vector<std::pair<int, int> > rank((size_t)v_cnt);
    for (int32_t i = 0; i < v_cnt; ++i) rank[i] = std::make_pair((int)i, (int)degree[i]);
    if (oriented_type == 1)
        std::sort(rank.begin(), rank.end(),
                  [](std::pair<int, int> a, std::pair<int, int> b) { return a.second > b.second; });
    else
        std::sort(rank.begin(), rank.end(),
                  [](std::pair<int, int> a, std::pair<int, int> b) { return a.second < b.second; });
    std::vector<uint32_t> new_id((size_t)v_cnt);
    for (int32_t i = 0; i < v_cnt; ++i) new_id[rank[i].first] = (uint32_t)i;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)pairs.size(); ++i) {
        uint64_t x = new_id[pairs[i] >> 32], y = new_id[(uint32_t)pairs[i]];
        pairs[i] = x << 32 | y;
    }
    std::vector<int64_t> nd((size_t)v_cnt);
    for (int32_t i = 0; i < v_cnt; ++i) nd[new_id[i]] = degree[i];
    degree.swap(nd);
}

} // namespace

extern "C" int gp_csr_load(const char* path, int32_t binary, int32_t oriented_type, gp_csr** out) {
    if (!path || !out) return gp_fail(GP_EINVAL, "null argument");
    MappedFile f;
    if (!f.open(path)) return gp_fail(GP_EIO, std::string("File not found. ") + path);
    if (binary) {
        // [v_cnt, e_cnt, max_deg, vertex[v_cnt+1], edge[e_cnt]] (src/dataloader.cpp:165-190)
        const uint32_t* w = (const uint32_t*)f.data;
        size_t words = f.size / 4;
        if (words < 3) return gp_fail(GP_EIO, "binary CSR too short");
        uint32_t v_cnt = w[0], e_cnt = w[1];
        if (v_cnt > (uint32_t)INT32_MAX) return gp_fail(GP_EIO, "binary CSR: vertex count above 2^31-1");
        if (words < 4 + (size_t)v_cnt + e_cnt) return gp_fail(GP_EIO, "binary CSR truncated");
        gp_csr* g = new gp_csr();
        g->v_cnt = (int32_t)v_cnt;
        g->e_cnt = e_cnt;
        g->max_degree = (int32_t)w[2];
        g->vertex = new uint32_t[(size_t)v_cnt + 1];
        g->edge = new int32_t[e_cnt ? e_cnt : 1];
        std::memcpy(g->vertex, w + 3, sizeof(uint32_t) * ((size_t)v_cnt + 1));
        std::memcpy(g->edge, w + 4 + v_cnt, sizeof(int32_t) * (size_t)e_cnt);
        // the reference keeps max_deg for both bounds (src/dataloader.cpp:179-181)
        int64_t d2 = 0, d1 = 0;
        for (uint32_t v = 0; v < v_cnt; ++v) {
            int64_t d = (int64_t)g->vertex[v + 1] - g->vertex[v];
            if (d > d1) {
                d2 = d1;
                d1 = d;
            } else if (d > d2) {
                d2 = d;
            }
        }
        g->max_intersection_size = (int32_t)std::max<int64_t>(d2, 0);
        g->tri_cnt = -1;
        if (gp_csr_validate(g->edge, g->vertex, (int32_t)v_cnt, e_cnt) != GP_OK) {
            const std::string why = gp_last_error();
            gp_csr_free(g);
            return gp_fail(GP_EIO, "binary CSR: " + why);
        }
        *out = g;
        return GP_OK;
    }
    IntReader rd{f.data, f.data + f.size};
    long long hv, he;
    if (!rd.next(&hv) || !rd.next(&he) || hv < 0 || he < 0)
        return gp_fail(GP_EIO, "missing or malformed header");
    // v_cnt is an int and e_cnt (2 x the lines) an unsigned int in the CSR
    if (hv > INT32_MAX || he > (long long)(UINT32_MAX / 2))
        return gp_fail(GP_EIO, "header counts exceed the CSR's int32 vertex / uint32 edge range");
    const int32_t v_cnt = (int32_t)hv;
    const uint64_t want = (uint64_t)he; // header counts every line, self-loops included
    std::vector<uint64_t> pairs;
    pairs.reserve(2 * want);
    std::vector<int64_t> degree((size_t)std::max<int32_t>(v_cnt, 0), 0);
    std::unordered_map<long long, uint32_t> ids;
    ids.reserve((size_t)v_cnt * 2 + 16);
    uint64_t lines = 0;
    uint32_t next_id = 0;
    bool overflow = false;
    auto id_of = [&](long long x) -> uint32_t {
        auto it = ids.find(x);
        if (it != ids.end()) return it->second;
        uint32_t id = next_id++;
        ids.emplace(x, id);
        return id;
    };
    long long x, y;
    while (rd.next(&x)) {
        if (!rd.next(&y)) return gp_fail(GP_EIO, "odd number of endpoints");
        ++lines;
        if (x == y) continue;
        uint32_t a = id_of(x), b = id_of(y);
        if (a >= (uint32_t)v_cnt || b >= (uint32_t)v_cnt) {
            overflow = true;
            continue;
        }
        pairs.push_back((uint64_t)a << 32 | b);
        pairs.push_back((uint64_t)b << 32 | a);
        ++degree[a];
        ++degree[b];
    }
    if (rd.bad) return gp_fail(GP_EIO, "non-numeric token");
    if (overflow || next_id != (uint32_t)v_cnt) return gp_fail(GP_EIO, "vertex number error!");
    if (lines != want) return gp_fail(GP_EIO, "edge number error!");
    if (oriented_type == 1 || oriented_type == 2) reorder_by_degree(pairs, degree, v_cnt, oriented_type);
    return gp_build_csr(pairs, v_cnt, degree, out);
}

extern "C" int gp_csr_complete(int32_t n, gp_csr** out) {
    if (n < 2 || !out) return gp_fail(GP_EINVAL, "complete graph needs n >= 2");
    // lines (i, j) for j < i, ids by first appearance: (1,0) gives 1->0, 0->1, ...
    std::vector<uint64_t> pairs;
    std::vector<int64_t> degree(n, 0);
    std::vector<int64_t> id(n, -1);
    uint32_t next_id = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            if (id[i] < 0) id[i] = next_id++;
            if (id[j] < 0) id[j] = next_id++;
            uint64_t a = (uint64_t)id[i], b = (uint64_t)id[j];
            pairs.push_back(a << 32 | b);
            pairs.push_back(b << 32 | a);
            ++degree[a];
            ++degree[b];
        }
    return gp_build_csr(pairs, n, degree, out);
}

extern "C" int gp_csr_reorder(const gp_csr* in, int32_t oriented_type, gp_csr** out) {
    if (!in || !out) return gp_fail(GP_EINVAL, "null argument");
    if (oriented_type < 0 || oriented_type > 2) return gp_fail(GP_EINVAL, "oriented_type must be 0, 1 or 2");
    const int32_t v_cnt = in->v_cnt;
    std::vector<uint64_t> pairs((size_t)in->e_cnt);
    std::vector<int64_t> degree((size_t)std::max<int32_t>(v_cnt, 0));
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)v_cnt; ++v) {
        degree[v] = (int64_t)in->vertex[v + 1] - in->vertex[v];
        for (uint32_t i = in->vertex[v]; i < in->vertex[v + 1]; ++i)
            pairs[i] = (uint64_t)v << 32 | (uint32_t)in->edge[i];
    }
    if (oriented_type != 0) reorder_by_degree(pairs, degree, v_cnt, oriented_type);
    int rc = gp_build_csr(pairs, v_cnt, degree, out);
    if (rc == GP_OK) (*out)->tri_cnt = in->tri_cnt; // relabelling keeps the triangle count
    return rc;
}

extern "C" int gp_csr_from_edges(const int32_t* edges, uint64_t n_edges, gp_csr** out) {
    if ((!edges && n_edges) || !out) return gp_fail(GP_EINVAL, "null argument");
    std::vector<uint64_t> pairs;
    pairs.reserve(2 * n_edges);
    std::unordered_map<int32_t, uint32_t> ids;
    std::vector<int64_t> degree;
    uint32_t next_id = 0;
    auto id_of = [&](int32_t x) -> uint32_t {
        auto it = ids.find(x);
        if (it != ids.end()) return it->second;
        ids.emplace(x, next_id);
        degree.push_back(0);
        return next_id++;
    };
    for (uint64_t i = 0; i < n_edges; ++i) {
        int32_t x = edges[2 * i], y = edges[2 * i + 1];
        if (x == y) continue;
        uint32_t a = id_of(x), b = id_of(y);
        pairs.push_back((uint64_t)a << 32 | b);
        pairs.push_back((uint64_t)b << 32 | a);
        ++degree[a];
        ++degree[b];
    }
    return gp_build_csr(pairs, (int32_t)next_id, degree, out);
}

extern "C" int gp_csr_free(gp_csr* g) {
    if (!g) return GP_OK;
    delete[] g->edge;
    delete[] g->vertex;
    delete g;
    return GP_OK;
}

extern "C" int gp_csr_hash(const gp_csr* g, uint64_t* hash) {
    if (!g || !hash) return gp_fail(GP_EINVAL, "null argument");
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    mix(g->vertex, sizeof(uint32_t) * ((size_t)g->v_cnt + 1));
    mix(g->edge, sizeof(int32_t) * (size_t)g->e_cnt);
    *hash = h;
    return GP_OK;
}

extern "C" int gp_csr_write_text(const gp_csr* g, const char* path) {
    if (!g || !path) return gp_fail(GP_EINVAL, "null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return gp_fail(GP_EIO, std::string("cannot write ") + path);
    std::fprintf(f, "%d %u\n", g->v_cnt, g->e_cnt / 2);
    for (int32_t u = 0; u < g->v_cnt; ++u)
        for (uint32_t e = g->vertex[u]; e < g->vertex[u + 1]; ++e)
            if (u < g->edge[e]) std::fprintf(f, "%d %d\n", u, g->edge[e]);
    std::fclose(f);
    return GP_OK;
}

extern "C" int gp_csr_write_binary(const gp_csr* g, const char* path) {
    if (!g || !path) return gp_fail(GP_EINVAL, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return gp_fail(GP_EIO, std::string("cannot write ") + path);
    uint32_t hdr[3] = {(uint32_t)g->v_cnt, g->e_cnt, (uint32_t)g->max_degree};
    std::fwrite(hdr, sizeof(uint32_t), 3, f);
    std::fwrite(g->vertex, sizeof(uint32_t), (size_t)g->v_cnt + 1, f);
    std::fwrite(g->edge, sizeof(int32_t), g->e_cnt, f);
    std::fclose(f);
    return GP_OK;
}

// ---------------------------------------------------------------------------
// DataLoader (C++ interface)
// ---------------------------------------------------------------------------

Graph* DataLoader::adopt(gp_csr* csr) {
    Graph* g = new Graph();
    g->v_cnt = csr->v_cnt;
    g->e_cnt = csr->e_cnt;
    g->edge = csr->edge;
    g->vertex = csr->vertex;
    g->max_degree = csr->max_degree;
    g->max_intersection_size = csr->max_intersection_size;
    g->tri_cnt = csr->tri_cnt;
    csr->edge = nullptr;
    csr->vertex = nullptr;
    gp_csr_free(csr);
    return g;
}

bool DataLoader::load_text(Graph*& g, const char* path, int oriented_type) {
    gp_csr* c = nullptr;
    int rc = gp_csr_load(path, 0, oriented_type, &c);
    if (rc != GP_OK) {
        std::printf("%s\n", gp_last_error());
        return false;
    }
    g = adopt(c);
    return true;
}

bool DataLoader::load_binary(Graph*& g, const char* path) {
    gp_csr* c = nullptr;
    int rc = gp_csr_load(path, 1, 0, &c);
    if (rc != GP_OK) {
        std::printf("%s\n", gp_last_error());
        return false;
    }
    g = adopt(c);
    return true;
}

bool DataLoader::load_data(Graph*& g, DataType type, const char* path, int oriented_type) {
    long long tri = -1;
    switch (type) {
        case Patents: tri = Patents_tri_cnt; break;
        case LiveJournal: tri = LiveJournal_tri_cnt; break;
        case MiCo: tri = MiCo_tri_cnt; break;
        case CiteSeer: tri = CiteSeer_tri_cnt; break;
        case Wiki_Vote: tri = Wiki_Vote_tri_cnt; break;
        case Orkut: tri = Orkut_tri_cnt; break;
        case Twitter: tri = Twitter_tri_cnt; break;
        case complete8: tri = -1; break;
        default:
            std::printf("invalid DataType!\n");
            return false;
    }
    bool ok = type == Twitter ? load_binary(g, path) : load_text(g, path, oriented_type);
    if (ok) g->tri_cnt = tri;
    return ok;
}

bool DataLoader::load_complete(Graph*& g, int clique_size) {
    gp_csr* c = nullptr;
    if (gp_csr_complete(clique_size, &c) != GP_OK) return false;
    g = adopt(c);
    return true;
}
