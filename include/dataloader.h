// dataloader.h — graph input (same interface as the reference's
// include/dataloader.h:7-48).  The triangle counts below are dataset facts the
// reference hard-codes for its performance model; they are kept so that
// load_data(g, Wiki_Vote, ...) yields the same tri_cnt.
#pragma once
#include <map>
#include <utility>

#include "graph.h"

enum DataType { Patents, Orkut, complete8, LiveJournal, MiCo, Twitter, CiteSeer, Wiki_Vote, Invalid };

const long long Patents_tri_cnt = 7515023LL;
const long long LiveJournal_tri_cnt = 177820130LL;
const long long MiCo_tri_cnt = 12534960LL;
const long long CiteSeer_tri_cnt = 1166LL;
const long long Wiki_Vote_tri_cnt = 608389LL;
const long long Orkut_tri_cnt = 627584181LL;
const long long Twitter_tri_cnt = 34824916864LL;

class DataLoader {
public:
    // oriented_type: 0 keep first-appearance ids, 1 high degree first, 2 low degree first
    bool load_data(Graph*& g, DataType type, const char* path, int oriented_type = 0);
    bool load_complete(Graph*& g, int clique_size);

    // ---- additions ----
    // GraphPi text file with tri_cnt left at -1
    bool load_text(Graph*& g, const char* path, int oriented_type = 0);
    // binary u32 CSR [v_cnt, e_cnt, max_deg, vertex[v_cnt+1], edge[e_cnt]]
    bool load_binary(Graph*& g, const char* path);
    // take ownership of a library CSR (from gp_csr_generate / gp_csr_from_edges)
    static Graph* adopt(gp_csr* csr);
};
