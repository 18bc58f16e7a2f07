// schedule.h — host-side planner: vertex order, symmetry-breaking
// restrictions, inclusion-exclusion (IEP) tables and the runtime tables the
// counting kernels read.  Public interface kept identical to the reference's
// include/schedule.h:9-58 so that `Schedule s(p, valid, pm, rt, iep, v, e, t);
// g->pattern_matching(s, n)` compiles unchanged; the implementation is this
// project's own (graphpi_amd/csrc/host/schedule.cpp).
#pragma once

#include <utility>
#include <vector>

#include "graphpi_amd.h"
#include "pattern.h"

class Schedule {
public:
    // performance_modeling_type: 0 none, 1 GraphPi model, 2 GraphZero model, 3 naive model
    // restricts_type: 0 none, 1 GraphPi restrictions, 2 GraphZero restrictions
    Schedule(const Pattern& pattern, bool& is_pattern_valid, int performance_modeling_type,
             int restricts_type, bool use_in_exclusion_optimize, int v_cnt, unsigned int e_cnt,
             long long tri_cnt = 0);
    // fixed order, no restrictions, no IEP (aborts on a disconnected prefix)
    Schedule(const int* _adj_mat, int _size);
    ~Schedule();

    inline int get_total_prefix_num() const { return total_prefix_num; }
    inline int get_father_prefix_id(int prefix_id) const { return father_prefix_id[prefix_id]; }
    inline int get_loop_set_prefix_id(int loop) const { return loop_set_prefix_id[loop]; }
    inline int get_size() const { return size; }
    inline int get_last(int i) const { return last[i]; }
    inline int get_next(int i) const { return next[i]; }
    inline int get_in_exclusion_optimize_num() const { return in_exclusion_optimize_num; }
    inline void set_in_exclusion_optimize_num(int num) { in_exclusion_optimize_num = num; }
    int get_in_exclusion_optimize_num_when_not_optimize();
    void add_restrict(const std::vector<std::pair<int, int> >& restricts);
    inline int get_total_restrict_num() const { return total_restrict_num; }
    inline int get_restrict_last(int i) const { return restrict_last[i]; }
    inline int get_restrict_next(int i) const { return restrict_next[i]; }
    inline int get_restrict_index(int i) const { return restrict_index[i]; }
    inline long long get_in_exclusion_optimize_redundancy() const {
        return in_exclusion_optimize_redundancy;
    }
    int get_max_degree() const;
    int get_multiplicity() const;
    void aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const;
    void aggressive_optimize_get_all_pairs(
        std::vector<std::vector<std::pair<int, int> > >& ordered_pairs_vector);
    void aggressive_optimize_dfs(Pattern base_dag, std::vector<std::vector<int> > isomorphism_vec,
                                 std::vector<std::vector<std::vector<int> > > permutation_groups,
                                 std::vector<std::pair<int, int> > ordered_pairs,
                                 std::vector<std::vector<std::pair<int, int> > >& ordered_pairs_vector);
    void restrict_selection(int v_cnt, unsigned int e_cnt, long long tri_cnt,
                            std::vector<std::vector<std::pair<int, int> > > ordered_pairs_vector,
                            std::vector<std::pair<int, int> >& best_restricts) const;
    void restricts_generate(const int* cur_adj_mat,
                            std::vector<std::vector<std::pair<int, int> > >& restricts);

    void GraphZero_aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const;
    void GraphZero_get_automorphisms(std::vector<std::vector<int> >& Aut) const;

    std::vector<std::vector<int> > get_isomorphism_vec() const;
    static std::vector<std::vector<int> > calc_permutation_group(const std::vector<int> vec, int size);
    inline const int* get_adj_mat_ptr() const { return adj_mat.data(); }

    void print_schedule() const;

    // Copy the runtime tables into the POD the C ABI and the kernels consume.
    // Returns GP_OK or GP_ERANGE when the schedule exceeds the ABI limits.
    int flatten(gp_sched_flat* out) const;

    std::vector<std::vector<std::vector<int> > > in_exclusion_optimize_group;
    std::vector<int> in_exclusion_optimize_val;
    std::vector<std::pair<int, int> > restrict_pair;

private:
    Schedule(const Schedule&);
    Schedule& operator=(const Schedule&);

    std::vector<int> adj_mat;
    std::vector<int> father_prefix_id, last, next, loop_set_prefix_id;
    std::vector<std::vector<int> > prefix_data;
    std::vector<int> restrict_last, restrict_next, restrict_index;
    int size = 0;
    int total_prefix_num = 0;
    int total_restrict_num = 0;
    int in_exclusion_optimize_num = 0;
    long long in_exclusion_optimize_redundancy = 1;

    void init_tables();
    void build_loop_invariant();
    int find_father_prefix(const std::vector<int>& data);
    double our_estimate_schedule_restrict(const std::vector<int>& order,
                                          const std::vector<std::pair<int, int> >& pairs, int v_cnt,
                                          unsigned int e_cnt, long long tri_cnt);
    double GraphZero_estimate_schedule_restrict(const std::vector<int>& order,
                                                const std::vector<std::pair<int, int> >& pairs,
                                                int v_cnt, unsigned int e_cnt);
    double Naive_estimate_schedule_restrict(const std::vector<int>& order,
                                            const std::vector<std::pair<int, int> >& pairs,
                                            int v_cnt, unsigned int e_cnt);
    void init_in_exclusion_optimize();
    int get_vec_optimize_num(const std::vector<int>& vec);
    void remove_invalid_permutation(std::vector<std::vector<int> >& candidate_permutations);
    void set_in_exclusion_optimize_redundancy();
};
