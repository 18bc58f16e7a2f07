// pattern.h — pattern graph (same class interface as the reference's
// include/pattern.h:10-43 so existing driver code compiles unchanged).
#pragma once
#include <set>
#include <vector>

#ifndef INDEX
#define INDEX(x, y, n) ((x) * (n) + (y))
#endif

// Preset patterns (reference include/pattern.h:10-18, src/pattern.cpp:45-133).
enum PatternType { Rectangle, QG3, Pentagon, House, Hourglass, Cycle_6_Tri, Clique_7_Minus };

class Pattern {
public:
    explicit Pattern(int _size, bool clique = false);
    // buffer: size*size characters, '1' marks an edge (row-major)
    Pattern(int _size, const char* buffer);
    Pattern(int _size, char* buffer) : Pattern(_size, (const char*)buffer) {}
    ~Pattern();
    Pattern(const Pattern& p);
    Pattern(PatternType type);
    void add_edge(int x, int y);
    void del_edge(int x, int y);
    inline void add_ordered_edge(int x, int y) { adj_mat[INDEX(x, y, size)] = 1; }
    inline int get_size() const { return size; }
    inline const int* get_adj_mat_ptr() const { return adj_mat; }
    bool check_connected() const;
    // inserts the edge set of every relabelling of this pattern into s
    void count_all_isomorphism(std::set<std::set<int> >& s) const;
    void print() const;
    // true when the directed graph adj_mat[x][y] > 0 (x -> y) has no cycle
    bool is_dag() const;

private:
    Pattern& operator=(const Pattern&);
    int* adj_mat;
    int size;
};
