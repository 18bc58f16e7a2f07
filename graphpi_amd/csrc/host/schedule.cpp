// Host planner (interface: reference include/schedule.h; behaviour:
// reference src/schedule.cpp).  Produces the same order, restriction set, IEP
// tables and runtime tables as the reference for the same inputs; the tests
// compare every field against schedules dumped by the reference build.
//
// One deliberate difference in mechanism: the reference validates restriction
// sets and derives the IEP redundancy by running Graph::pattern_matching on
// complete graphs (src/schedule.cpp:1433-1456, :1746-1764).  On K_m every
// candidate set is "all vertices not yet embedded", so that count is the number
// of injective maps of the n pattern vertices into m graph vertices that obey
// the active restrictions:  C(m, n) * L(P), where L(P) is the number of
// orderings of the n vertices consistent with the restriction pairs P (a pair
// (a, b) asks value(b) < value(a)).  With IEP width k, the loop levels apply
// only pairs keyed at depths < n - k.  complete_graph_count() evaluates exactly
// that, so the planner never runs the counting engine on the CPU.
#include "schedule.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace {

// number of permutations 'val' of 0..n-1 with val[b] < val[a] for each (a, b)
long long linear_extensions(int n, const std::vector<std::pair<int, int> >& pairs) {
    std::vector<unsigned> below(n, 0); // vertices that must take a smaller value
    for (const auto& p : pairs) below[p.first] |= 1u << p.second;
    const unsigned full = (1u << n) - 1;
    std::vector<long long> ways(full + 1, 0);
    ways[0] = 1;
    // 'placed' = the vertices holding the |placed| smallest values
    for (unsigned placed = 0; placed < full; ++placed) {
        if (!ways[placed]) continue;
        for (int x = 0; x < n; ++x)
            if (!(placed >> x & 1) && (below[x] & ~placed) == 0) ways[placed | 1u << x] += ways[placed];
    }
    return ways[full];
}

long long binom(int m, int n) {
    if (n < 0 || n > m) return 0;
    long long r = 1;
    for (int i = 1; i <= n; ++i) r = r * (m - n + i) / i;
    return r;
}

long long factorial(int n) {
    long long r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

// What Graph::pattern_matching(s, 1) returns on the complete graph K_m.
long long complete_graph_count(int m, const Schedule& s) {
    const int n = s.get_size();
    const int k = s.get_in_exclusion_optimize_num();
    std::vector<std::pair<int, int> > active;
    for (const auto& p : s.restrict_pair)
        if (k <= 1 || p.second < n - k) active.push_back(p);
    long long raw = binom(m, n) * linear_extensions(n, active);
    return raw / s.get_in_exclusion_optimize_redundancy();
}

// counts[i] = orderings satisfying the first i+1 of the (sorted) pairs
std::vector<long long> prefix_extension_counts(int n, const std::vector<std::pair<int, int> >& sorted) {
    std::vector<long long> counts;
    std::vector<std::pair<int, int> > cur;
    for (const auto& p : sorted) {
        cur.push_back(p);
        counts.push_back(linear_extensions(n, cur));
    }
    return counts;
}

// conditional survival fractions of the restriction chain, in the same
// floating-point steps as the reference (src/schedule.cpp:1526-1545)
std::vector<double> restriction_fractions(int n, const std::vector<std::pair<int, int> >& sorted) {
    std::vector<long long> c = prefix_extension_counts(n, sorted);
    std::vector<double> sum(c.size());
    double total = 1;
    for (int i = 2; i <= n; ++i) total *= i;
    for (size_t i = 0; i < c.size(); ++i) sum[i] = (double)c[i] / total;
    for (int i = (int)sum.size() - 1; i > 0; --i) sum[i] /= sum[i - 1];
    return sum;
}

std::vector<int> relabel(const std::vector<int>& adj, int n, const std::vector<int>& order) {
    // position p of the new order holds original vertex order[p]
    std::vector<int> rank(n), out(n * n);
    for (int i = 0; i < n; ++i) rank[order[i]] = i;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[INDEX(rank[i], rank[j], n)] = adj[INDEX(i, j, n)];
    return out;
}

} // namespace

// ---------------------------------------------------------------------------

void Schedule::init_tables() {
    int max_prefix_num = size * (size - 1) / 2;
    if (max_prefix_num < 1) max_prefix_num = 1;
    father_prefix_id.assign(max_prefix_num, -1);
    next.assign(max_prefix_num, -1);
    last.assign(size, -1);
    loop_set_prefix_id.assign(size, -1);
    prefix_data.assign(max_prefix_num, std::vector<int>());
    restrict_last.assign(size, -1);
    restrict_next.assign(max_prefix_num, -1);
    restrict_index.assign(max_prefix_num, -1);
    total_prefix_num = 0;
    total_restrict_num = 0;
}

Schedule::Schedule(const Pattern& pattern, bool& is_pattern_valid, int performance_modeling_type,
                   int restricts_type, bool use_in_exclusion_optimize, int v_cnt, unsigned int e_cnt,
                   long long tri_cnt) {
    if (performance_modeling_type != 0 && tri_cnt == -1) {
        std::printf("Fatal: Can not use performance modeling if not have triangle number of this dataset.\n");
        std::fflush(stdout);
        std::abort();
    }
    is_pattern_valid = true;
    size = pattern.get_size();
    adj_mat.assign(pattern.get_adj_mat_ptr(), pattern.get_adj_mat_ptr() + size * size);
    const std::vector<int> original = adj_mat;

    std::vector<std::pair<int, int> > best_pairs;
    if (performance_modeling_type != 0) {
        std::vector<std::vector<int> > candidates;
        std::vector<int> perm(size);
        for (int i = 0; i < size; ++i) perm[i] = i;
        do candidates.push_back(perm);
        while (std::next_permutation(perm.begin(), perm.end()));
        remove_invalid_permutation(candidates);
        if (performance_modeling_type == 1) {
            int widest = 0;
            for (const auto& vec : candidates) widest = std::max(widest, get_vec_optimize_num(vec));
            std::vector<std::vector<int> > keep;
            for (const auto& vec : candidates)
                if (get_vec_optimize_num(vec) == widest) keep.push_back(vec);
            candidates.swap(keep);
        }
        std::vector<int> best_order(size);
        bool have_best = false;
        double min_val = 0;
        // restriction candidates depend only on the relabelled matrix
        std::map<std::vector<int>, std::vector<std::vector<std::pair<int, int> > > > memo;
        auto model = [&](const std::vector<int>& vec, const std::vector<std::pair<int, int> >& pairs) {
            if (performance_modeling_type == 1) return our_estimate_schedule_restrict(vec, pairs, v_cnt, e_cnt, tri_cnt);
            if (performance_modeling_type == 2) return GraphZero_estimate_schedule_restrict(vec, pairs, v_cnt, e_cnt);
            return Naive_estimate_schedule_restrict(vec, pairs, v_cnt, e_cnt);
        };
        for (const auto& vec : candidates) {
            std::vector<int> cur = relabel(original, size, vec);
            auto it = memo.find(cur);
            if (it == memo.end()) {
                std::vector<std::vector<std::pair<int, int> > > rv;
                if (restricts_type == 1) {
                    restricts_generate(cur.data(), rv);
                } else {
                    Schedule plain(cur.data(), size);
                    std::vector<std::pair<int, int> > pairs;
                    plain.GraphZero_aggressive_optimize(pairs);
                    rv.push_back(pairs);
                }
                it = memo.emplace(cur, rv).first;
            }
            const auto& rv = it->second;
            if (rv.empty()) {
                std::vector<std::pair<int, int> > none;
                double val = model(vec, none);
                if (!have_best || val < min_val) {
                    have_best = true;
                    min_val = val;
                    best_order = vec;
                    best_pairs = none;
                }
            }
            for (const auto& pairs : rv) {
                double val = model(vec, pairs);
                if (!have_best || val < min_val) {
                    have_best = true;
                    min_val = val;
                    best_order = vec;
                    best_pairs = pairs;
                }
            }
        }
        adj_mat = relabel(original, size, best_order);
    } else {
        std::vector<int> identity(size);
        for (int i = 0; i < size; ++i) identity[i] = i;
        std::vector<std::vector<std::pair<int, int> > > rv;
        if (restricts_type != 0) {
            if (restricts_type == 1) {
                restricts_generate(adj_mat.data(), rv);
            } else {
                std::vector<std::pair<int, int> > pairs;
                GraphZero_aggressive_optimize(pairs);
                rv.push_back(pairs);
            }
        }
        bool have_best = false;
        double min_val = 0;
        for (const auto& pairs : rv) {
            double val = restricts_type == 1 ? our_estimate_schedule_restrict(identity, pairs, v_cnt, e_cnt, tri_cnt)
                                             : GraphZero_estimate_schedule_restrict(identity, pairs, v_cnt, e_cnt);
            if (!have_best || val < min_val) {
                have_best = true;
                min_val = val;
                best_pairs = pairs;
            }
        }
    }

    if (use_in_exclusion_optimize) {
        std::vector<int> identity(size);
        for (int i = 0; i < size; ++i) identity[i] = i;
        in_exclusion_optimize_num = get_vec_optimize_num(identity);
        if (in_exclusion_optimize_num <= 1)
            in_exclusion_optimize_num = 0;
        else
            init_in_exclusion_optimize();
    } else {
        in_exclusion_optimize_num = 0;
    }

    init_tables();
    for (int i = 1; i < size; ++i) {
        bool linked = false;
        for (int j = 0; j < i && !linked; ++j) linked = adj_mat[INDEX(i, j, size)] != 0;
        if (!linked) {
            is_pattern_valid = false;
            return;
        }
    }
    build_loop_invariant();
    if (restricts_type != 0) add_restrict(best_pairs);
    set_in_exclusion_optimize_redundancy();
}

Schedule::Schedule(const int* _adj_mat, int _size) {
    size = _size;
    adj_mat.assign(_adj_mat, _adj_mat + size * size);
    init_tables();
    in_exclusion_optimize_num = 0;
    for (int i = 1; i < size; ++i) {
        bool linked = false;
        for (int j = 0; j < i && !linked; ++j) linked = adj_mat[INDEX(i, j, size)] != 0;
        if (!linked) {
            std::printf("invalid schedule!\n");
            std::fflush(stdout);
            std::abort();
        }
    }
    build_loop_invariant();
    set_in_exclusion_optimize_redundancy();
}

Schedule::~Schedule() {}

int Schedule::get_max_degree() const {
    int mx = 0;
    for (int i = 0; i < size; ++i) {
        int d = 0;
        for (int j = 0; j < size; ++j) d += adj_mat[INDEX(i, j, size)];
        mx = std::max(mx, d);
    }
    return mx;
}

// Prefix p = the sorted list of earlier pattern vertices whose neighbourhoods
// are intersected.  Prefixes are shared: each is created once, keyed by its
// last element, and points at the prefix without that element (its father).
void Schedule::build_loop_invariant() {
    loop_set_prefix_id[0] = -1;
    for (int i = 1; i < size; ++i) {
        std::vector<int> data;
        for (int j = 0; j < i; ++j)
            if (adj_mat[INDEX(i, j, size)]) data.push_back(j);
        loop_set_prefix_id[i] = find_father_prefix(data);
    }
}

int Schedule::find_father_prefix(const std::vector<int>& data) {
    if (data.empty()) return -1;
    int key = data.back();
    for (int p = last[key]; p != -1; p = next[p])
        if (prefix_data[p] == data) return p;
    int father = find_father_prefix(std::vector<int>(data.begin(), data.end() - 1));
    int id = total_prefix_num++;
    father_prefix_id[id] = father;
    next[id] = last[key];
    last[key] = id;
    prefix_data[id] = data;
    return id;
}

// Drops a pair (a, c) implied by (a, b) and (b, c), then links each pair into
// the list of the depth it is checked at (its second element).
void Schedule::add_restrict(const std::vector<std::pair<int, int> >& restricts) {
    restrict_pair = restricts;
    for (size_t i = 0; i < restrict_pair.size();) {
        bool implied = false;
        for (size_t j = 0; j < restrict_pair.size() && !implied; ++j) {
            if (i == j || restrict_pair[j].first != restrict_pair[i].first) continue;
            for (size_t k = 0; k < restrict_pair.size(); ++k)
                if (k != i && k != j && restrict_pair[k].second == restrict_pair[i].second &&
                    restrict_pair[j].second == restrict_pair[k].first) {
                    implied = true;
                    break;
                }
        }
        if (implied)
            restrict_pair.erase(restrict_pair.begin() + i);
        else
            ++i;
    }
    std::fill(restrict_last.begin(), restrict_last.end(), -1);
    if (restrict_next.size() < restrict_pair.size()) {
        restrict_next.resize(restrict_pair.size(), -1);
        restrict_index.resize(restrict_pair.size(), -1);
    }
    std::fill(restrict_next.begin(), restrict_next.end(), -1);
    total_restrict_num = 0;
    for (const auto& p : restrict_pair) {
        restrict_index[total_restrict_num] = p.first;
        restrict_next[total_restrict_num] = restrict_last[p.second];
        restrict_last[p.second] = total_restrict_num;
        ++total_restrict_num;
    }
}

int Schedule::get_multiplicity() const { return (int)get_isomorphism_vec().size(); }

std::vector<std::vector<int> > Schedule::get_isomorphism_vec() const {
    std::vector<std::vector<int> > out;
    std::vector<int> v(size);
    for (int i = 0; i < size; ++i) v[i] = i;
    do {
        bool keeps_edges = true;
        for (int i = 0; i < size && keeps_edges; ++i)
            for (int j = i + 1; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)] && !adj_mat[INDEX(v[i], v[j], size)]) {
                    keeps_edges = false;
                    break;
                }
        if (keeps_edges) out.push_back(v);
    } while (std::next_permutation(v.begin(), v.end()));
    return out;
}

std::vector<std::vector<int> > Schedule::calc_permutation_group(const std::vector<int> vec, int size) {
    std::vector<char> used(size, 0);
    std::vector<std::vector<int> > cycles;
    for (int i = 0; i < (int)vec.size(); ++i) {
        if (used[i]) continue;
        std::vector<int> cyc{i};
        used[i] = 1;
        for (int x = vec[i]; !used[x]; x = vec[x]) {
            used[x] = 1;
            cyc.push_back(x);
        }
        cycles.push_back(cyc);
    }
    return cycles;
}

namespace {
// A group that is exactly one transposition (other points fixed) yields the
// pair directly; returns 1 and sets *pair in that case, 0 if it has no 2-cycle
// or several, -1 when a longer cycle is present.
int lone_transposition(const std::vector<std::vector<int> >& group, std::pair<int, int>* pair) {
    int twos = 0;
    for (const auto& c : group) {
        if (c.size() == 2) {
            ++twos;
            *pair = std::make_pair(c[0], c[1]);
        } else if (c.size() != 1) {
            return -1;
        }
    }
    return twos;
}
} // namespace

void Schedule::aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const {
    std::vector<std::vector<int> > iso = get_isomorphism_vec();
    std::vector<std::vector<std::vector<int> > > groups;
    for (const auto& v : iso) groups.push_back(calc_permutation_group(v, size));
    ordered_pairs.clear();
    for (size_t i = 0; i < groups.size();) {
        std::pair<int, int> p;
        if (lone_transposition(groups[i], &p) == 1) {
            groups.erase(groups.begin() + i);
            iso.erase(iso.begin() + i);
            ordered_pairs.push_back(p);
        } else {
            ++i;
        }
    }
    Pattern base(size);
    for (const auto& p : ordered_pairs) base.add_ordered_edge(p.first, p.second);
    bool changed = true;
    while (changed && iso.size() != 1) {
        for (size_t i = 0; i < iso.size();) {
            Pattern t(base);
            for (const auto& p : ordered_pairs) t.add_ordered_edge(iso[i][p.first], iso[i][p.second]);
            if (!t.is_dag()) {
                groups.erase(groups.begin() + i);
                iso.erase(iso.begin() + i);
            } else {
                ++i;
            }
        }
        changed = false;
        for (size_t i = 0; i < groups.size(); ++i) {
            const std::vector<int>* two = nullptr;
            for (const auto& c : groups[i])
                if (c.size() == 2) {
                    two = &c;
                    break;
                }
            if (two) {
                std::pair<int, int> p((*two)[0], (*two)[1]);
                groups.erase(groups.begin() + i);
                iso.erase(iso.begin() + i);
                ordered_pairs.push_back(p);
                base.add_ordered_edge(p.first, p.second);
                changed = true;
                break;
            }
        }
    }
}

void Schedule::aggressive_optimize_get_all_pairs(
    std::vector<std::vector<std::pair<int, int> > >& ordered_pairs_vector) {
    std::vector<std::vector<int> > iso = get_isomorphism_vec();
    std::vector<std::vector<std::vector<int> > > groups;
    for (const auto& v : iso) groups.push_back(calc_permutation_group(v, size));
    ordered_pairs_vector.clear();
    std::vector<std::pair<int, int> > pairs;
    for (size_t i = 0; i < groups.size();) {
        std::pair<int, int> p;
        if (lone_transposition(groups[i], &p) == 1) {
            groups.erase(groups.begin() + i);
            iso.erase(iso.begin() + i);
            pairs.push_back(p);
        } else {
            ++i;
        }
    }
    Pattern base(size);
    for (const auto& p : pairs) base.add_ordered_edge(p.first, p.second);
    aggressive_optimize_dfs(base, iso, groups, pairs, ordered_pairs_vector);
}

// Branch on every transposition of the first automorphism (in order) that
// still has one; an automorphism is eliminated once the chosen pairs and their
// image under it cannot be ordered consistently (they form a directed cycle).
void Schedule::aggressive_optimize_dfs(Pattern base_dag, std::vector<std::vector<int> > isomorphism_vec,
                                       std::vector<std::vector<std::vector<int> > > permutation_groups,
                                       std::vector<std::pair<int, int> > ordered_pairs,
                                       std::vector<std::vector<std::pair<int, int> > >& ordered_pairs_vector) {
    for (size_t i = 0; i < isomorphism_vec.size();) {
        Pattern t(base_dag);
        const std::vector<int>& iso = isomorphism_vec[i];
        for (const auto& p : ordered_pairs) t.add_ordered_edge(iso[p.first], iso[p.second]);
        if (!t.is_dag()) {
            permutation_groups.erase(permutation_groups.begin() + i);
            isomorphism_vec.erase(isomorphism_vec.begin() + i);
        } else {
            ++i;
        }
    }
    if (isomorphism_vec.size() == 1) {
        ordered_pairs_vector.push_back(ordered_pairs);
        return;
    }
    for (size_t i = 0; i < permutation_groups.size(); ++i) {
        bool branched = false;
        for (const auto& c : permutation_groups[i]) {
            if (c.size() != 2) continue;
            branched = true;
            std::pair<int, int> p(c[0], c[1]);
            auto groups2 = permutation_groups;
            auto iso2 = isomorphism_vec;
            auto pairs2 = ordered_pairs;
            Pattern base2(base_dag);
            groups2.erase(groups2.begin() + i);
            iso2.erase(iso2.begin() + i);
            pairs2.push_back(p);
            base2.add_ordered_edge(p.first, p.second);
            aggressive_optimize_dfs(base2, iso2, groups2, pairs2, ordered_pairs_vector);
        }
        if (branched) break;
    }
}

void Schedule::GraphZero_get_automorphisms(std::vector<std::vector<int> >& Aut) const {
    Aut.clear();
    std::vector<int> p(size);
    for (int i = 0; i < size; ++i) p[i] = i;
    do {
        bool same = true;
        for (int i = 0; i < size && same; ++i)
            for (int j = 0; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)] != adj_mat[INDEX(p[i], p[j], size)]) {
                    same = false;
                    break;
                }
        if (same) Aut.push_back(p);
    } while (std::next_permutation(p.begin(), p.end()));
}

// GraphZero's symmetry breaking: walk the vertices, record (v, x[v]) for every
// automorphism moving v, keep only automorphisms fixing v; then keep one pair
// per second element, with the largest first element.
void Schedule::GraphZero_aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const {
    std::vector<std::vector<int> > Aut;
    GraphZero_get_automorphisms(Aut);
    std::vector<std::pair<int, int> > L;
    for (int v = 0; v < size; ++v) {
        std::vector<std::vector<int> > fixing;
        for (const auto& x : Aut) {
            if (x[v] == v) {
                fixing.push_back(x);
                continue;
            }
            std::pair<int, int> cur(std::min(v, x[v]), std::max(v, x[v]));
            if (std::find(L.begin(), L.end(), cur) == L.end()) L.push_back(cur);
        }
        Aut.swap(fixing);
    }
    ordered_pairs.clear();
    for (const auto& l : L) {
        bool fresh = true;
        for (auto& o : ordered_pairs)
            if (l.second == o.second) {
                fresh = false;
                if (l.first > o.first) o.first = l.first;
                break;
            }
        if (fresh) ordered_pairs.push_back(l);
    }
}

void Schedule::restricts_generate(const int* cur_adj_mat,
                                  std::vector<std::vector<std::pair<int, int> > >& restricts) {
    Schedule schedule(cur_adj_mat, get_size());
    schedule.aggressive_optimize_get_all_pairs(restricts);
    const int n = schedule.get_size();
    long long ans = complete_graph_count(n + 1, schedule) / schedule.get_multiplicity();
    for (size_t i = 0; i < restricts.size();) {
        Schedule cur(schedule.get_adj_mat_ptr(), n);
        cur.add_restrict(restricts[i]);
        if (complete_graph_count(n + 1, cur) != ans)
            restricts.erase(restricts.begin() + i);
        else
            ++i;
    }
}

// Width of the trailing independent set of order vec (or -1 when some vertex
// has no earlier neighbour).
int Schedule::get_vec_optimize_num(const std::vector<int>& vec) {
    for (int i = 1; i < size; ++i) {
        bool linked = false;
        for (int j = 0; j < i && !linked; ++j) linked = adj_mat[INDEX(vec[i], vec[j], size)] != 0;
        if (!linked) return -1;
    }
    for (int k = 2; k <= size; ++k)
        for (int i = size - k + 1; i < size; ++i)
            if (adj_mat[INDEX(vec[size - k], vec[i], size)]) return k - 1;
    return -1;
}

void Schedule::remove_invalid_permutation(std::vector<std::vector<int> >& candidate_permutations) {
    std::vector<std::vector<int> > keep;
    for (const auto& vec : candidate_permutations) {
        bool ok = true;
        for (int x = 1; x < size && ok; ++x) {
            bool linked = false;
            for (int y = 0; y < x && !linked; ++y) linked = adj_mat[INDEX(vec[x], vec[y], size)] != 0;
            ok = linked;
        }
        if (ok) keep.push_back(vec);
    }
    candidate_permutations.swap(keep);
}

// GraphPi cost model (src/schedule.cpp:1487-1574), evaluated with the same
// sequence of double operations so that ties resolve identically.
double Schedule::our_estimate_schedule_restrict(const std::vector<int>& order,
                                                const std::vector<std::pair<int, int> >& pairs,
                                                int v_cnt, unsigned int e_cnt, long long tri_cnt) {
    const int md = std::max(get_max_degree(), 1);
    std::vector<double> p_size(md), pp_size(md);
    double p0 = e_cnt * 1.0 / v_cnt / v_cnt;
    double p1 = tri_cnt * 1.0 * v_cnt / e_cnt / e_cnt;
    p_size[0] = v_cnt;
    for (int i = 1; i < md; ++i) p_size[i] = p_size[i - 1] * p0;
    pp_size[0] = 1;
    for (int i = 1; i < md; ++i) pp_size[i] = pp_size[i - 1] * p1;

    std::vector<int> cur = relabel(adj_mat, size, order);
    std::vector<std::pair<int, int> > restricts = pairs;
    std::sort(restricts.begin(), restricts.end());
    std::vector<double> sum = restriction_fractions(size, restricts);

    std::vector<std::vector<int> > invariant_size(size);
    double val = 1;
    for (int i = size - 1; i >= 0; --i) {
        int cnt_forward = 0;
        for (int j = 0; j < i; ++j)
            if (cur[INDEX(j, i, size)]) ++cnt_forward;
        int c = cnt_forward;
        for (int j = i - 1; j >= 0; --j)
            if (cur[INDEX(j, i, size)]) invariant_size[j].push_back(c--);
        for (int s : invariant_size[i])
            if (s > 1) {
                double t = p_size[1] * pp_size[s - 2];
                t = t + p_size[1];
                val = val + t;
            }
        val += 1;
        for (size_t j = 0; j < restricts.size(); ++j)
            if (restricts[j].second == i) val *= sum[j];
        if (i) {
            double t = p_size[1] * pp_size[cnt_forward - 1];
            val *= t;
        } else {
            val *= p_size[0];
        }
    }
    return val;
}

// GraphZero cost model (src/schedule.cpp:1576-1653).  The reference indexes
// its size-max_degree table at cnt_forward, one past the end for a vertex whose
// earlier neighbours are all the pattern's max-degree neighbours; the table is
// extended by one term of the same geometric sequence here.
double Schedule::GraphZero_estimate_schedule_restrict(const std::vector<int>& order,
                                                      const std::vector<std::pair<int, int> >& pairs,
                                                      int v_cnt, unsigned int e_cnt) {
    const int md = std::max(get_max_degree(), 1);
    std::vector<double> p_size(md + 1);
    double p = e_cnt * 1.0 / v_cnt / v_cnt;
    p_size[0] = v_cnt;
    for (int i = 1; i <= md; ++i) p_size[i] = p_size[i - 1] * p;

    std::vector<int> cur = relabel(adj_mat, size, order);
    std::vector<std::pair<int, int> > restricts = pairs;
    std::sort(restricts.begin(), restricts.end());
    std::vector<double> sum = restriction_fractions(size, restricts);

    std::vector<std::vector<int> > invariant_size(size);
    double val = 1;
    for (int i = size - 1; i >= 0; --i) {
        int cnt_forward = 0;
        for (int j = 0; j < i; ++j)
            if (cur[INDEX(j, i, size)]) ++cnt_forward;
        int c = cnt_forward;
        for (int j = i - 1; j >= 0; --j)
            if (cur[INDEX(j, i, size)]) invariant_size[j].push_back(c--);
        for (int s : invariant_size[i])
            if (s > 1) {
                double t = p_size[s - 1] + p_size[1];
                val = val + t;
            }
        for (size_t j = 0; j < restricts.size(); ++j)
            if (restricts[j].second == i) val *= sum[j];
        val *= p_size[cnt_forward];
    }
    return val;
}

// Naive cost model (src/schedule.cpp:1655-1713).
double Schedule::Naive_estimate_schedule_restrict(const std::vector<int>& order,
                                                  const std::vector<std::pair<int, int> >& pairs,
                                                  int v_cnt, unsigned int e_cnt) {
    double p = e_cnt * 2.0 / v_cnt / v_cnt;
    std::vector<int> cur = relabel(adj_mat, size, order);
    std::vector<std::pair<int, int> > restricts = pairs;
    std::sort(restricts.begin(), restricts.end());
    std::vector<double> sum = restriction_fractions(size, restricts);
    double val = 1;
    for (int i = size - 1; i >= 0; --i) {
        int cnt_forward = 0;
        for (int j = 0; j < i; ++j)
            if (cur[INDEX(j, i, size)]) ++cnt_forward;
        for (size_t j = 0; j < restricts.size(); ++j)
            if (restricts[j].second == i) val *= sum[j];
        val *= v_cnt;
        for (int j = 0; j < i - cnt_forward; ++j) val *= (1 - p);
        for (int j = 0; j < cnt_forward; ++j) val *= p;
    }
    return val;
}

// Picks the restriction set with the lowest GraphPi cost for the current
// order (the reference's variant, src/schedule.cpp:1335-1431, reads one entry
// before its table for vertex 0; the root term uses v_cnt here, as
// our_estimate_schedule_restrict does).
void Schedule::restrict_selection(int v_cnt, unsigned int e_cnt, long long tri_cnt,
                                  std::vector<std::vector<std::pair<int, int> > > ordered_pairs_vector,
                                  std::vector<std::pair<int, int> >& best_restricts) const {
    std::vector<int> identity(size);
    for (int i = 0; i < size; ++i) identity[i] = i;
    Schedule& self = const_cast<Schedule&>(*this);
    bool have_best = false;
    double min_val = 0;
    for (auto& restricts : ordered_pairs_vector) {
        std::sort(restricts.begin(), restricts.end());
        double val = self.our_estimate_schedule_restrict(identity, restricts, v_cnt, e_cnt, tri_cnt);
        if (!have_best || val < min_val) {
            have_best = true;
            min_val = val;
            best_restricts = restricts;
        }
    }
}

// IEP tables: every set partition of the k trailing vertices, enumerated as a
// restricted growth string (open a new block first, then join blocks 0..c-1),
// weighted by the Moebius function of the partition lattice,
// prod over blocks of (-1)^(|B|-1) (|B|-1)!.  The reference derives the same
// weights by counting even/odd connected spanning subgraphs of each block
// (src/schedule.cpp:1114-1210).
void Schedule::init_in_exclusion_optimize() {
    const int k = in_exclusion_optimize_num;
    in_exclusion_optimize_group.clear();
    in_exclusion_optimize_val.clear();
    std::vector<int> id(k, 0);
    std::vector<int> weight(k + 1, 1); // weight[s] = (-1)^(s-1) (s-1)!
    for (int s = 1; s <= k; ++s) {
        int f = 1;
        for (int i = 2; i < s; ++i) f *= i;
        weight[s] = (s % 2 == 1) ? f : -f;
    }
    // explicit stack instead of recursion: (position, blocks opened so far)
    struct Frame {
        int pos, blocks, choice;
    };
    std::vector<Frame> st{{0, 0, -1}};
    while (!st.empty()) {
        Frame& f = st.back();
        if (f.pos == k) {
            std::vector<std::vector<int> > group(f.blocks);
            for (int j = 0; j < k; ++j) group[id[j]].push_back(j);
            int val = 1;
            for (const auto& b : group) val *= weight[b.size()];
            in_exclusion_optimize_group.push_back(group);
            in_exclusion_optimize_val.push_back(val);
            st.pop_back();
            continue;
        }
        // choice -1 = new block, then 0..blocks-1
        if (f.choice >= f.blocks) {
            st.pop_back();
            continue;
        }
        int c = f.choice++;
        int pos = f.pos, blocks = f.blocks;
        if (c == -1) {
            id[pos] = blocks;
            st.push_back({pos + 1, blocks + 1, -1});
        } else {
            id[pos] = c;
            st.push_back({pos + 1, blocks, -1});
        }
    }
}

void Schedule::set_in_exclusion_optimize_redundancy() {
    int k = get_in_exclusion_optimize_num();
    if (k <= 1) {
        in_exclusion_optimize_redundancy = 1;
        return;
    }
    in_exclusion_optimize_redundancy = 1;
    long long ans = complete_graph_count(size, *this);
    set_in_exclusion_optimize_num(0);
    long long true_ans = complete_graph_count(size, *this);
    set_in_exclusion_optimize_num(k);
    in_exclusion_optimize_redundancy = ans / true_ans;
}

int Schedule::get_in_exclusion_optimize_num_when_not_optimize() {
    std::vector<int> identity(size);
    for (int i = 0; i < size; ++i) identity[i] = i;
    return get_vec_optimize_num(identity);
}

void Schedule::print_schedule() const {
    std::printf("Schedule:\n");
    for (int i = 0; i < size; ++i) {
        for (int j = 0; j < size; ++j) std::printf("%d", adj_mat[INDEX(i, j, size)]);
        std::puts("");
    }
}

int Schedule::flatten(gp_sched_flat* out) const {
    if (!out) return GP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    if (size < 2 || size > GP_MAX_SIZE) return GP_ERANGE;
    if (total_prefix_num > GP_MAX_PREFIX || total_restrict_num > GP_MAX_RESTRICT) return GP_ERANGE;
    const int k = in_exclusion_optimize_num;
    if (k > GP_MAX_IEP_K || (int)in_exclusion_optimize_group.size() > GP_MAX_IEP_GROUPS) return GP_ERANGE;
    out->size = size;
    out->total_prefix_num = total_prefix_num;
    for (int p = 0; p < GP_MAX_PREFIX; ++p) {
        out->father[p] = p < total_prefix_num ? father_prefix_id[p] : -1;
        out->next[p] = p < total_prefix_num ? next[p] : -1;
    }
    for (int d = 0; d < GP_MAX_SIZE; ++d) {
        out->last[d] = d < size ? last[d] : -1;
        out->loop_set[d] = d < size ? loop_set_prefix_id[d] : -1;
        out->restrict_last[d] = d < size ? restrict_last[d] : -1;
    }
    out->total_restrict_num = total_restrict_num;
    for (int r = 0; r < GP_MAX_RESTRICT; ++r) {
        out->restrict_next[r] = r < total_restrict_num ? restrict_next[r] : -1;
        out->restrict_index[r] = r < total_restrict_num ? restrict_index[r] : -1;
    }
    out->iep_num = k;
    out->iep_group_num = k > 1 ? (int)in_exclusion_optimize_group.size() : 0;
    for (int g = 0; g < out->iep_group_num; ++g) {
        for (int i = 0; i < GP_MAX_SIZE; ++i) out->iep_block[g][i] = -1;
        const auto& grp = in_exclusion_optimize_group[g];
        for (size_t b = 0; b < grp.size(); ++b)
            for (int j : grp[b]) out->iep_block[g][j] = (int8_t)b;
        out->iep_val[g] = in_exclusion_optimize_val[g];
    }
    out->redundancy = in_exclusion_optimize_redundancy;
    for (int i = 0; i < size * size; ++i) out->adj[i] = adj_mat[i];
    out->restrict_pair_num = (int)restrict_pair.size();
    if (out->restrict_pair_num > GP_MAX_RESTRICT) return GP_ERANGE;
    for (size_t r = 0; r < restrict_pair.size(); ++r) {
        out->restrict_pair[r][0] = restrict_pair[r].first;
        out->restrict_pair[r][1] = restrict_pair[r].second;
    }
    out->multiplicity = get_multiplicity();
    return GP_OK;
}
