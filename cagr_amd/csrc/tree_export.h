// Host-side conversion of the native tree to the reference's serialised
// layout (CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_treeutils.cpp:202-258):
// parent [M], children [M,8] (-1 = none), is_leaf [M], half_w [M] = 2^-depth,
// num_points [M], node2point_indexstart [M], node2point_index [sum num_points]
// where every node lists its points in ascending original index and the lists
// are concatenated in preorder.  Used by cagr_tree_export (parity tests and
// external callers of the reference layout), never by a hot path.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <vector>
#include "tree_logic.h"

namespace cagr {

inline int64_t export_reference_layout(const TopoRec* topo, int64_t m, const int32_t* perm,
                                       const uint64_t* sorted_keys, int64_t* parent,
                                       int64_t* children, uint8_t* is_leaf, double* half_w,
                                       int64_t* num_points, int64_t* n2p_index,
                                       int64_t* n2p_start) {
    for (int64_t i = 0; i < m * 8; ++i) children[i] = -1;
    int64_t run = 0;
    std::vector<int64_t> buf;
    for (int64_t i = 0; i < m; ++i) {
        const TopoRec& t = topo[i];
        const int d = topo_depth(t);
        parent[i] = t.parent;
        if (t.parent >= 0) children[(int64_t)t.parent * 8 + octant_at(sorted_keys[t.begin], d)] = i;
        is_leaf[i] = topo_leaf(t) ? 1 : 0;
        half_w[i] = std::ldexp(1.0, -d);
        const int64_t c = (int64_t)t.end - t.begin;
        num_points[i] = c;
        n2p_start[i] = run;
        if (n2p_index) {
            buf.assign(perm + t.begin, perm + t.end);
            std::sort(buf.begin(), buf.end());
            std::copy(buf.begin(), buf.end(), n2p_index + run);
        }
        run += c;
    }
    return run;
}

}  // namespace cagr
