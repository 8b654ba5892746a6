// Per-element octree logic shared by the HIP tree-build kernels.
//
// Every function here is a pure function of its arguments, marked
// CAGR_HD (= __host__ __device__ under hipcc).  The kernels in tree_build.hip
// call them once per thread; tests/host_logic/ compiles this header with g++
// and replays the same build serially on the CPU to check the logic against
// the oracle tree before anything runs on a GPU.
//
// Tree semantics follow the reference's recursive builder
// (CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_treeutils.cpp:90-177):
// root centre 0, half-width 1; at a node of depth d the octant of x is
// (x>=cx) + 2(y>=cy) + 4(z>=cz); a node splits iff it holds more than
// leaf_size points and d < max_depth; ids are DFS preorder, children in
// octant order.
//
// Integer restatement used here.  Along one axis the 15 split decisions a
// point meets (depth 0..14, centres at odd multiples of 2^-d) are exactly the
// bits, MSB first, of the cell index
//      ix = clamp(floor(x * 2^14) + 2^14, 0, 2^15 - 1)
// (floor and the power-of-two scale are exact; NaN compares false everywhere,
// i.e. ix = 0).  Interleaving the three cell indices, octant bits (z,y,x) per
// level from the root down, gives a 45-bit key whose lexicographic order is
// the preorder of leaves.  With the points stably sorted by key:
//   * cpl(a,b)   = number of leading levels on which two keys agree (0..15;
//     -1 for keys of different clouds of a forest);
//   * a point b's depth-d cell holds > S points iff some window of S+1
//     consecutive sorted points containing b has cpl >= d, so
//     leaf_depth(b) = min(1 + max_window_cpl(b), max_depth)   (0 if N <= S);
//   * the nodes that start at sorted position b are the chain of depths
//     start_depth(b) .. leaf_depth(b), start_depth(b) = cpl(key[b-1], key[b])+1
//     (0 for b = 0);
//   * preorder = order by (begin, depth), so node (b, d) gets id
//     node_offset[b] + d - start_depth(b), node_offset = exclusive scan of the
//     chain lengths, and a node's escape (the id after its subtree) is
//     node_offset[end].
#pragma once
#include <cstdint>
#include <cmath>

#ifndef CAGR_HD
#define CAGR_HD __host__ __device__
#endif

namespace cagr {

constexpr int kLevels = 15;          // split levels encoded in a key (ALLOWED_MAX_DEPTH)
constexpr int kKeyBits = 3 * kLevels;

// floor(x * 2^14) + 2^14 clamped to [0, 2^15 - 1]; NaN -> 0.
CAGR_HD inline uint32_t cell_index(double x) {
    const double s = x * 16384.0;             // exact (power of two)
    if (!(s >= -16384.0)) return 0u;          // also catches NaN
    if (s >= 16383.0) return 32767u;
    return (uint32_t)((int32_t)floor(s) + 16384);
}

CAGR_HD inline uint32_t cell_index(float x) {
    const float s = x * 16384.0f;
    if (!(s >= -16384.0f)) return 0u;
    if (s >= 16383.0f) return 32767u;
    return (uint32_t)((int32_t)floorf(s) + 16384);
}

// spread the low 15 bits of v so that bit i lands on bit 3i
CAGR_HD inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x7fffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

CAGR_HD inline uint64_t morton_key(uint32_t ix, uint32_t iy, uint32_t iz) {
    return spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
}

// Keys of a forest (several independent clouds sorted together) carry the
// cloud id above the 45 Morton bits; points of different clouds share no
// cell, not even a root.
CAGR_HD inline uint64_t forest_key(uint64_t cloud, uint64_t morton) {
    return (cloud << kKeyBits) | morton;
}

// number of leading levels (3-bit groups from the top) on which a and b agree;
// -1 if they belong to different clouds
CAGR_HD inline int cpl(uint64_t a, uint64_t b) {
    if ((a ^ b) >> kKeyBits) return -1;
    const uint64_t x = (a ^ b) & ((1ull << kKeyBits) - 1);
    if (x == 0) return kLevels;
#if defined(__HIP_DEVICE_COMPILE__)
    const int h = 63 - __clzll((long long)x);
#else
    const int h = 63 - __builtin_clzll(x);
#endif
    return (kKeyBits - 1 - h) / 3;
}

// octant of the child at depth d (>= 1) on the path of key
CAGR_HD inline int octant_at(uint64_t key, int d) {
    return (int)((key >> (3 * (kLevels - d))) & 7u);
}

// leaf depth of sorted position b (keys sorted, n points, leaf size S)
CAGR_HD inline int leaf_depth_of(const uint64_t* keys, int64_t n, int64_t b, int S,
                                 int max_depth) {
    if (n <= S) return 0;
    int64_t lo = b - S;
    if (lo < 0) lo = 0;
    int64_t hi = b;
    if (hi > n - 1 - S) hi = n - 1 - S;
    int m = -1;
    for (int64_t i = lo; i <= hi; ++i) {
        const int c = cpl(keys[i], keys[i + S]);
        m = c > m ? c : m;
    }
    const int ld = m + 1;
    return ld < max_depth ? ld : max_depth;
}

CAGR_HD inline int start_depth_of(const uint64_t* keys, int64_t b) {
    return b == 0 ? 0 : cpl(keys[b - 1], keys[b]) + 1;
}

// one past the last sorted position sharing key[b]'s depth-d cell (d >= 0;
// depth 0 is the cloud's root cell)
CAGR_HD inline int64_t cell_end(const uint64_t* keys, int64_t n, int64_t b, int d) {
    int64_t lo = b + 1, hi = n;     // answer in [b+1, n]
    const uint64_t kb = keys[b];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (cpl(kb, keys[mid]) >= d) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// first sorted position sharing key[b]'s depth-d cell
CAGR_HD inline int64_t cell_begin(const uint64_t* keys, int64_t b, int d) {
    int64_t lo = 0, hi = b;         // answer in [0, b]
    const uint64_t kb = keys[b];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (cpl(kb, keys[mid]) >= d) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Static per-node traversal record (32 bytes, read with wave-uniform loads).
struct alignas(16) TopoRec {
    double thr2;      // (open_factor * half_w)^2; far iff dist^2 > thr2
    int32_t escape;   // preorder id of the first node after this subtree
    int32_t begin;    // first sorted point of the node
    int32_t end;      // one past its last sorted point
    int32_t info;     // depth | (is_leaf << 8) | (child count << 12)
    int32_t parent;   // -1 for a root
    int32_t kids;     // position of the first child in the depth order (-1: none)
};

CAGR_HD inline int topo_depth(const TopoRec& t) { return t.info & 0xff; }
CAGR_HD inline bool topo_leaf(const TopoRec& t) { return (t.info >> 8) & 1; }
CAGR_HD inline int topo_nkids(const TopoRec& t) { return (t.info >> 12) & 15; }

}  // namespace cagr

namespace cagr {

// ---- per-element bodies of the build kernels (one call per thread) ----

// chain of nodes starting at sorted position b: leaf depth, start depth, length
CAGR_HD inline void chain_body(int64_t b, const uint64_t* keys, int64_t n, int S, int max_depth,
                               int8_t* ld, int8_t* sd, int32_t* cnt) {
    const int l = leaf_depth_of(keys, n, b, S, max_depth);
    const int s = start_depth_of(keys, b);
    ld[b] = (int8_t)l;
    sd[b] = (int8_t)s;
    cnt[b] = l >= s ? l - s + 1 : 0;
}

// topology records of the nodes starting at b; off = exclusive scan of cnt
// (off[n] = number of nodes)
CAGR_HD inline void node_body(int64_t b, const uint64_t* keys, int64_t n, const int8_t* ld,
                              const int8_t* sd, const int32_t* off, double open_factor,
                              TopoRec* topo) {
    const int s = sd[b], l = ld[b];
    for (int d = s; d <= l; ++d) {
        const int32_t id = off[b] + (d - s);
        const int64_t e = cell_end(keys, n, b, d);
        TopoRec t;
        const double thr = open_factor * ldexp(1.0, -d);
        t.thr2 = thr * thr;
        t.escape = off[e];
        t.begin = (int32_t)b;
        t.end = (int32_t)e;
        t.info = d | ((d == l ? 1 : 0) << 8);
        if (d == 0) {
            t.parent = -1;
        } else if (d - 1 >= s) {
            t.parent = id - 1;
        } else {
            const int64_t b2 = cell_begin(keys, b, d - 1);
            t.parent = off[b2] + (d - 1 - sd[b2]);
        }
        t.kids = -1;  // set once the depth order exists
        topo[id] = t;
    }
}

// First child's position in the depth order (node ids sorted by depth,
// ascending id within a depth) and child count of every internal node, from
// the position i >= 1 of each node of depth >= 1.  The children of p are
// consecutive there: a node whose id lies between two children of p is inside
// p's subtree, so at p's depth + 1 it is a child of p too.
CAGR_HD inline void kids_body(int64_t i, TopoRec* topo, const int32_t* order, int64_t m) {
    const int32_t p = topo[order[i]].parent;
    if (topo[order[i - 1]].parent == p) return;  // not the first child
    int64_t j = i + 1;
    while (j < m && topo[order[j]].parent == p) ++j;
    topo[p].kids = (int32_t)i;
    topo[p].info = topo[p].info | ((int32_t)(j - i) << 12);
}

// leaves list their points in ascending original index: insertion-sort the
// leaf's slice of the permutation (already sorted within equal keys)
CAGR_HD inline void leaf_sort_body(const TopoRec& t, int32_t* perm) {
    for (int32_t i = t.begin + 1; i < t.end; ++i) {
        const int32_t v = perm[i];
        int32_t j = i - 1;
        while (j >= t.begin && perm[j] > v) {
            perm[j + 1] = perm[j];
            --j;
        }
        perm[j + 1] = v;
    }
}

}  // namespace cagr
