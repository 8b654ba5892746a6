// Marching cubes on the PGR octree, per leaf (GaussRecon's
// Octree::GetMCIsoTriangles with fullDepthIso 0, nonLinearFit 1, no
// barycentres, triangles out: gaussrecon_src/Octree.cpp:518-1257 with
// MarchingCubes.cpp, Cube.cpp and Geometry.cpp:74-220).
//
// Every function is a pure function of the tree view and its arguments,
// marked CAGR_HD (= __host__ __device__ under hipcc): the kernels of
// pgr_mesh.hip call them once per leaf, and tests/host_logic/ compiles this
// header with g++ to replay the mesh serially against the reference's own
// build before anything runs on a GPU.
//
// The reference walks a pointer octree.  Here a node is (depth, offsets) and
// the tree is the sorted list of its internal nodes:
//   * a node's marching-cubes index is GetIndex of the values at its own 8
//     grid corners -- the reference propagates leaf corner bits upwards
//     (SetIsoSurfaceCorners, :518-611), and the bit of corner c of a node
//     always comes from the leaf reached by descending into child c, which
//     holds the same grid point;
//   * a face or edge neighbour "exists and has children" (the only case the
//     root search and the face walk act on) iff the same-depth neighbour
//     cell is an internal node: coarser neighbours have no children by
//     construction, and ReconOctNode::__edgeNeighbor returns either that
//     cell, NULL or a childless coarser node.
#pragma once
#include <cstddef>
#include <cstdint>

#include "mc_tables.h"

// every float expression rounds each operation, as the reference's x86 build
#if defined(__clang__)
#pragma clang fp contract(off)
#endif

#ifndef CAGR_HD
#define CAGR_HD __host__ __device__
#endif

namespace cagr {
namespace mc {

// limits of the per-leaf work (checked; exceeding them is reported)
// (the test clouds peak at 19 edges and 19 roots per leaf; the 1M-point
// ellipsoid cloud needs more than 64 and fits 128)
constexpr int kMaxEdges = 128;     // iso edges of one leaf incl. finer neighbours
constexpr int kMaxVerts = 128;     // distinct roots met by one leaf
constexpr int kMaxWalk = 48;       // explicit stacks of the descents (depth <= 13)

struct Node {
    int d;
    int32_t o[3];
};

// the tree and the field, all arrays in the caller's memory space
struct View {
    int md;                        // Octree::maxDepth (tree height + 1)
    const uint64_t* in_keys;       // internal nodes: d<<39 | z<<26 | y<<13 | x, ascending
    int64_t n_in;
    const int64_t* ck;             // grid-corner keys (ReconOctNode::getCornerIndex), ascending
    const int32_t* ck_idx;         // their gridDataVector positions
    int64_t n_ck;
    const float* values;           // GridData::value per gridDataVector position
    float iso;
    float smooth_width;
    float grid;                    // 1 / 2^(md + 1): coordinate of one key unit
    // optional: marching-cubes indices computed once per node (null: from
    // the corner values at every use)
    const uint8_t* in_mc;          // per internal node, in in_keys order
    const uint64_t* leaf_mk;       // leaves' left-aligned Morton keys (preorder = ascending)
    const uint8_t* leaf_mc;        // per leaf
    int64_t n_leaves;
    // optional open-addressing tables (slot = key + 1, 0 empty; linear
    // probing) answering the same lookups as the sorted arrays in a probe or two
    const uint64_t* in_slot;       // internal node key -> position in in_keys
    const int32_t* in_val;
    uint64_t in_mask;
    const uint64_t* lf_slot;       // leaf Morton key -> leaf position
    const int32_t* lf_val;
    uint64_t lf_mask;
};

CAGR_HD inline uint64_t hash_mix(uint64_t k) {      // splitmix64 finaliser
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ull;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebull;
    return k ^ (k >> 31);
}
CAGR_HD inline int64_t hash_find(const uint64_t* slot, const int32_t* val, uint64_t mask,
                                 uint64_t key) {
    for (uint64_t h = hash_mix(key) & mask;; h = (h + 1) & mask) {
        const uint64_t s = slot[h];
        if (s == key + 1) return val[h];
        if (s == 0) return -1;
    }
}

struct RootInfo {
    Node node;
    int edge;
    int64_t key;
};

// ------------------------------------------------------------ cube indexing
// (Cube.cpp: corner c = x | y<<1 | z<<2, edge e = orientation<<2 | i | j<<1,
// face f = 2*axis + side)
CAGR_HD inline int corner_bit(int c) {               // MarchingCubes::cornerMap
    const int m[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    return m[c];
}
CAGR_HD inline void edge_corners(int e, int& c1, int& c2) {
    const int o = e >> 2, i = e & 1, j = (e >> 1) & 1;
    if (o == 0) { c1 = (i << 1) | (j << 2); c2 = c1 | 1; }
    else if (o == 1) { c1 = i | (j << 2); c2 = c1 | 2; }
    else { c1 = i | (j << 1); c2 = c1 | 4; }
}
CAGR_HD inline void faces_of_edge(int e, int& f1, int& f2) {     // FacesAdjacentToEdge
    const int o = e >> 2, i = e & 1, j = (e >> 1) & 1;
    if (o == 0) { f1 = 2 + i; f2 = 4 + j; }
    else if (o == 1) { f1 = 0 + i; f2 = 4 + j; }
    else { f1 = 0 + i; f2 = 2 + j; }
}
CAGR_HD inline int face_reflect_edge(int e, int f) {             // FaceReflectEdgeIndex
    const int axis = f >> 1, o = e >> 2;
    if (o == axis) return e;
    // the edge's coordinate along `axis` is i (first free axis) or j (second)
    const int first = (o == 0) ? 1 : 0;
    return (axis == first) ? (e ^ 1) : (e ^ 2);
}
CAGR_HD inline int edge_reflect_edge(int e) { return e ^ 3; }   // EdgeReflectEdgeIndex
CAGR_HD inline int face_of_edges(int e1, int e2) {               // FaceAdjacentToEdges
    int f1, f2, g1, g2;
    faces_of_edge(e1, f1, f2);
    faces_of_edge(e2, g1, g2);
    if (f1 == g1 || f1 == g2) return f1;
    if (f2 == g1 || f2 == g2) return f2;
    return -1;
}
CAGR_HD inline void face_corners(int f, int c[4]) {              // FaceCorners
    const int i = f & 1;
    switch (f >> 1) {
        case 0: c[0] = i; c[1] = i | 2; c[2] = i | 4; c[3] = i | 6; return;
        case 1: c[0] = i << 1; c[1] = 1 | (i << 1); c[2] = (i << 1) | 4; c[3] = 1 | (i << 1) | 4; return;
        default: c[0] = i << 2; c[1] = 1 | (i << 2); c[2] = 2 | (i << 2); c[3] = 3 | (i << 2); return;
    }
}
CAGR_HD inline int edge_mask(int mc) {           // edges whose corners differ in sign
    int m = 0;
    for (int e = 0; e < 12; ++e) {
        int c1, c2;
        edge_corners(e, c1, c2);
        if (((mc >> corner_bit(c1)) & 1) != ((mc >> corner_bit(c2)) & 1)) m |= 1 << e;
    }
    return m;
}
CAGR_HD inline bool has_edge_roots(int mc, int e) { return (edge_mask(mc) >> e) & 1; }
CAGR_HD inline bool has_roots(int mc) { return mc != 0 && mc != 255; }

// ------------------------------------------------------------ tree lookups
CAGR_HD inline uint64_t node_key(int d, int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)d << 39) | ((uint64_t)z << 26) | ((uint64_t)y << 13) | (uint64_t)x;
}
template <class K>
CAGR_HD inline int64_t lower_bound(const K* a, int64_t n, K v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
CAGR_HD inline bool is_internal(const View& t, int d, int64_t x, int64_t y, int64_t z) {
    const int64_t lim = (int64_t)1 << d;
    if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) return false;
    const uint64_t k = node_key(d, x, y, z);
    if (t.in_slot) return hash_find(t.in_slot, t.in_val, t.in_mask, k) >= 0;
    const int64_t i = lower_bound(t.in_keys, t.n_in, k);
    return i < t.n_in && t.in_keys[i] == k;
}
CAGR_HD inline bool is_internal(const View& t, const Node& n) {
    return is_internal(t, n.d, n.o[0], n.o[1], n.o[2]);
}
CAGR_HD inline Node child(const Node& n, int c) {
    return Node{n.d + 1, {2 * n.o[0] + (c & 1), 2 * n.o[1] + ((c >> 1) & 1), 2 * n.o[2] + ((c >> 2) & 1)}};
}
CAGR_HD inline int child_index(const Node& n) {
    return (n.o[0] & 1) | ((n.o[1] & 1) << 1) | ((n.o[2] & 1) << 2);
}
CAGR_HD inline Node parent(const Node& n) {
    return Node{n.d - 1, {n.o[0] >> 1, n.o[1] >> 1, n.o[2] >> 1}};
}
// ReconOctNode::getCornerIndex (ReconOctNode.cpp:95-105)
CAGR_HD inline int64_t corner_key(const View& t, const Node& n, int c) {
    int64_t off[3];
    for (int k = 0; k < 3; ++k) off[k] = ((int64_t)n.o[k] + ((c >> k) & 1)) << (t.md - n.d + 1);
    return off[0] | (off[1] << 15) | (off[2] << 30);
}
CAGR_HD inline int32_t corner_slot(const View& t, const Node& n, int c) {
    const int64_t k = corner_key(t, n, c);
    const int64_t i = lower_bound(t.ck, t.n_ck, k);
    return (i < t.n_ck && t.ck[i] == k) ? t.ck_idx[i] : -1;
}
CAGR_HD inline uint64_t morton_left(const Node& n, int H) {
    uint64_t m = 0;
    for (int b = 0; b < n.d; ++b)
        for (int a = 0; a < 3; ++a)
            m |= (uint64_t)((n.o[a] >> (n.d - 1 - b)) & 1) << (3 * (H - 1 - b) + a);
    return m;
}

// MarchingCubes::GetIndex (MarchingCubes.cpp:472-483) at the node's corners
CAGR_HD inline int mc_from_corners(const View& t, const Node& n);
CAGR_HD inline int mc_index(const View& t, const Node& n) {
    if (t.in_mc) {
        const uint64_t k = node_key(n.d, n.o[0], n.o[1], n.o[2]);
        const uint64_t m = morton_left(n, t.md - 1);
        if (t.in_slot) {
            const int64_t i = hash_find(t.in_slot, t.in_val, t.in_mask, k);
            if (i >= 0) return t.in_mc[i];
            const int64_t j = hash_find(t.lf_slot, t.lf_val, t.lf_mask, m);
            if (j >= 0) return t.leaf_mc[j];
        } else {
            const int64_t i = lower_bound(t.in_keys, t.n_in, k);
            if (i < t.n_in && t.in_keys[i] == k) return t.in_mc[i];
            const int64_t j = lower_bound(t.leaf_mk, t.n_leaves, m);
            if (j < t.n_leaves && t.leaf_mk[j] == m) return t.leaf_mc[j];
        }
    }
    return mc_from_corners(t, n);
}
CAGR_HD inline int mc_from_corners(const View& t, const Node& n) {
    int idx = 0;
    for (int c = 0; c < 8; ++c) {
        const int32_t s = corner_slot(t, n, c);
        if (s >= 0 && t.values[s] < t.iso) idx |= 1 << corner_bit(c);
    }
    return idx;
}

// same-depth neighbour across face f / edge e, when it is an internal node
CAGR_HD inline bool face_nbr_internal(const View& t, const Node& n, int f, Node& nb) {
    nb = n;
    nb.o[f >> 1] += (f & 1) ? 1 : -1;
    return is_internal(t, nb);
}
CAGR_HD inline bool edge_nbr_internal(const View& t, const Node& n, int e, Node& nb) {
    const int o = e >> 2, i = e & 1, j = (e >> 1) & 1;
    const int a = (o == 0) ? 1 : 0, b = (o == 2) ? 1 : 2;
    nb = n;
    nb.o[a] += i ? 1 : -1;
    nb.o[b] += j ? 1 : -1;
    return is_internal(t, nb);
}

// Octree::GetRootIndex (Octree.cpp:614-685): the finest edge segment holding
// the root of edge e of node n.  The reference recursion (neighbour with
// children first, then child c1 before c2) runs here on an explicit stack.
CAGR_HD inline bool root_index(const View& t, const Node& n0, int e0, RootInfo& ri) {
    Node sn[kMaxWalk];
    int se[kMaxWalk];
    int top = 0;
    sn[top] = n0;
    se[top++] = e0;
    while (top > 0) {
        --top;
        const Node n = sn[top];
        const int e = se[top];
        if (!has_edge_roots(mc_index(t, n), e)) continue;
        int f1, f2;
        faces_of_edge(e, f1, f2);
        Node fin = n, nb;
        int fe = e;
        bool finer = false;
        if (n.d < t.md) {
            if (face_nbr_internal(t, n, f1, nb)) { fin = nb; fe = face_reflect_edge(e, f1); finer = true; }
            else if (face_nbr_internal(t, n, f2, nb)) { fin = nb; fe = face_reflect_edge(e, f2); finer = true; }
            else if (edge_nbr_internal(t, n, e, nb)) { fin = nb; fe = edge_reflect_edge(e); finer = true; }
        }
        if (!finer) finer = is_internal(t, n);
        int c1, c2;
        edge_corners(fe, c1, c2);
        if (finer) {
            if (top + 2 > kMaxWalk) return false;
            sn[top] = child(fin, c2); se[top++] = fe;
            sn[top] = child(fin, c1); se[top++] = fe;
            continue;
        }
        const int o = fe >> 2, i1 = fe & 1, i2 = (fe >> 1) & 1;
        const int a = (o == 0) ? 1 : 0, b = (o == 2) ? 1 : 2;
        const int sh = t.md - fin.d + 1;
        const int64_t offset = ((int64_t)1 << fin.d) + fin.o[o] - 1;
        const int64_t ea = (int64_t)(fin.o[a] + i1) << sh, eb = (int64_t)(fin.o[b] + i2) << sh;
        ri.node = fin;
        ri.edge = fe;
        ri.key = (int64_t)o | (ea << 5) | (eb << 25) | (offset << 45);
        return true;
    }
    return false;
}

// Octree::GetRoot (Octree.cpp:750-804): the root on the edge, interpolated
// with the smoothing widths, in [0,1]^3 coordinates
CAGR_HD inline void root_position(const View& t, const RootInfo& ri, float p[3]) {
    int c1, c2;
    edge_corners(ri.edge, c1, c2);
    const int32_t s1 = corner_slot(t, ri.node, c1), s2 = corner_slot(t, ri.node, c2);
    float value1 = t.values[s1], value2 = t.values[s2];
    const float width1 = t.smooth_width, width2 = t.smooth_width;
    const int64_t k1 = corner_key(t, ri.node, c1), k2 = corner_key(t, ri.node, c2);
    const float a[3] = {(float)(k1 & 0xFFFF) * t.grid, (float)((k1 >> 15) & 0xFFFF) * t.grid,
                        (float)(k1 >> 30) * t.grid};
    const float b[3] = {(float)(k2 & 0xFFFF) * t.grid, (float)((k2 >> 15) & 0xFFFF) * t.grid,
                        (float)(k2 >> 30) * t.grid};
    const float isov = -t.iso;
    value1 *= -1;
    value2 *= -1;
    float ratio = (value1 - isov) * width1 / ((value1 - isov) * width1 - (value2 - isov) * width2);
    if (ratio > 1.0) ratio = 1.0;
    if (ratio < 0.0) ratio = 0.0;
    for (int k = 0; k < 3; ++k) p[k] = a[k] - ratio * (a[k] - b[k]);
}

// the vertex candidates of a leaf in the reference's order
// (SetMCRootPositions, Octree.cpp:814-850: orientation, then i, then j)
CAGR_HD inline int leaf_roots(const View& t, const Node& leaf, int64_t keys[12], RootInfo* infos) {
    if (!has_roots(mc_index(t, leaf))) return 0;
    int n = 0;
    for (int o = 0; o < 3; ++o)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                RootInfo ri;
                const int e = i | (j << 1) | (o << 2);
                keys[n] = root_index(t, leaf, e, ri) ? ri.key : -1;
                if (infos) infos[n] = ri;
                ++n;
            }
    return n;
}

// Octree::GetRootPair (Octree.cpp:881-898)
CAGR_HD inline bool root_pair(const View& t, const RootInfo& ri, RootInfo& pair) {
    Node n = ri.node;
    int c1, c2;
    edge_corners(ri.edge, c1, c2);
    while (n.d > 0) {
        const int c = child_index(n);
        if (c != c1 && c != c2) return false;
        const Node p = parent(n);
        if (!has_edge_roots(mc_index(t, p), ri.edge))
            return root_index(t, child(p, c == c1 ? c2 : c1), ri.edge, pair);
        n = p;
    }
    return false;
}

// ------------------------------------------------------------ per-leaf polygons
// workspace of one leaf.  Its arrays are lane-strided views into
// caller-provided memory: element i of a lane's array lies kStride elements
// after its element i - 1.  The kernels interleave the workspaces of one
// wave (kStride = 64, lane l at offset l), so that the lanes of a wave that
// touch the same index read one contiguous run instead of 64 cache lines;
// the host replay uses kStride = 1.  The counters are plain members.
#ifndef CAGR_MC_STRIDE
#define CAGR_MC_STRIDE 64
#endif
constexpr int kStride = CAGR_MC_STRIDE;
// index check of every workspace access: a no-op in the kernels; the host
// replay defines it to abort on an index outside the array's length
#ifndef CAGR_MC_CHECK
#define CAGR_MC_CHECK(cond) ((void)0)
#endif

template <class T, int N>   // N = the array's length in elements
struct Lanes {
    T* p;
    CAGR_HD T& operator[](int i) const {
        CAGR_MC_CHECK(i >= 0 && i < N);
        return p[(int64_t)i * kStride];
    }
};
template <class T, int N>
struct Vec {   // a small fixed array kept together as one element
    T v[N];
    CAGR_HD T& operator[](int i) {
        CAGR_MC_CHECK(i >= 0 && i < N);
        return v[i];
    }
    CAGR_HD const T& operator[](int i) const {
        CAGR_MC_CHECK(i >= 0 && i < N);
        return v[i];
    }
};

constexpr size_t align8(size_t x) { return (x + 7) & ~size_t(7); }

// byte offsets of the arrays in one lane's share (multiples of 8), and its size
struct WorkLayout {
    static constexpr size_t ea = 0;
    static constexpr size_t eb = align8(ea + sizeof(int64_t) * kMaxEdges);
    static constexpr size_t vkey = align8(eb + sizeof(int64_t) * kMaxEdges);
    static constexpr size_t vcount = align8(vkey + sizeof(int64_t) * kMaxVerts);
    static constexpr size_t vinfo = align8(vcount + sizeof(int) * kMaxVerts);
    static constexpr size_t loop = align8(vinfo + sizeof(RootInfo) * kMaxVerts);
    static constexpr size_t back = align8(loop + sizeof(int64_t) * kMaxEdges);
    static constexpr size_t vid = align8(back + sizeof(int64_t) * kMaxEdges);
    static constexpr size_t tp = align8(vid + sizeof(int) * kMaxEdges);
    static constexpr size_t te_p = align8(tp + sizeof(Vec<float, 3>) * kMaxEdges);
    static constexpr size_t te_t = align8(te_p + sizeof(Vec<int, 2>) * 3 * kMaxEdges);
    static constexpr size_t tt_e = align8(te_t + sizeof(Vec<int, 2>) * 3 * kMaxEdges);
    static constexpr size_t own_state = align8(tt_e + sizeof(Vec<int, 3>) * kMaxEdges);
    static constexpr size_t own = align8(own_state + sizeof(int8_t) * 12);
    static constexpr size_t out = align8(own + sizeof(RootInfo) * 12);
    static constexpr size_t bytes = align8(out + sizeof(int32_t) * 3 * kMaxEdges);
};

template <class T, int N>
CAGR_HD inline Lanes<T, N> lanes_at(unsigned char* blk, size_t off, int lane) {
    return Lanes<T, N>{reinterpret_cast<T*>(blk + off * kStride) + lane};
}

struct Work {
    Lanes<int64_t, kMaxEdges> ea, eb;          // iso edges (first, second) by root key
    int ne = 0;
    Lanes<int64_t, kMaxVerts> vkey;            // vertexCount of the reference
    Lanes<int, kMaxVerts> vcount;
    Lanes<RootInfo, kMaxVerts> vinfo;
    int nv = 0;
    Lanes<int64_t, kMaxEdges> loop;            // one loop's vertex keys
    Lanes<int64_t, kMaxEdges> back;            // its back chain while it is built
    Lanes<int, kMaxEdges> vid;                 // its vertex numbers
    // triangulation (Geometry.cpp:74-220)
    Lanes<Vec<float, 3>, kMaxEdges> tp;
    Lanes<Vec<int, 2>, 3 * kMaxEdges> te_p, te_t;
    Lanes<Vec<int, 3>, kMaxEdges> tt_e;
    int nte = 0, ntt = 0;
    bool overflow = false;
    // the leaf's own 12 edge roots, found on first use (0 unknown, 1 found, 2 none)
    Lanes<int8_t, 12> own_state;
    Lanes<RootInfo, 12> own;
    Lanes<int32_t, 3 * kMaxEdges> out;         // the leaf's triangles, before they are placed

    // lane `lane` of the kStride interleaved workspaces at `blk`
    // (WorkLayout::bytes * kStride bytes, 8-byte aligned)
    CAGR_HD Work(unsigned char* blk, int lane)
        : ea(lanes_at<int64_t, kMaxEdges>(blk, WorkLayout::ea, lane)),
          eb(lanes_at<int64_t, kMaxEdges>(blk, WorkLayout::eb, lane)),
          vkey(lanes_at<int64_t, kMaxVerts>(blk, WorkLayout::vkey, lane)),
          vcount(lanes_at<int, kMaxVerts>(blk, WorkLayout::vcount, lane)),
          vinfo(lanes_at<RootInfo, kMaxVerts>(blk, WorkLayout::vinfo, lane)),
          loop(lanes_at<int64_t, kMaxEdges>(blk, WorkLayout::loop, lane)),
          back(lanes_at<int64_t, kMaxEdges>(blk, WorkLayout::back, lane)),
          vid(lanes_at<int, kMaxEdges>(blk, WorkLayout::vid, lane)),
          tp(lanes_at<Vec<float, 3>, kMaxEdges>(blk, WorkLayout::tp, lane)),
          te_p(lanes_at<Vec<int, 2>, 3 * kMaxEdges>(blk, WorkLayout::te_p, lane)),
          te_t(lanes_at<Vec<int, 2>, 3 * kMaxEdges>(blk, WorkLayout::te_t, lane)),
          tt_e(lanes_at<Vec<int, 3>, kMaxEdges>(blk, WorkLayout::tt_e, lane)),
          own_state(lanes_at<int8_t, 12>(blk, WorkLayout::own_state, lane)),
          own(lanes_at<RootInfo, 12>(blk, WorkLayout::own, lane)),
          out(lanes_at<int32_t, 3 * kMaxEdges>(blk, WorkLayout::out, lane)) {}
};

CAGR_HD inline int vfind(const Work& w, int64_t key) {
    for (int i = 0; i < w.nv; ++i)
        if (w.vkey[i] == key) return i;
    return -1;
}
CAGR_HD inline int vget(Work& w, const RootInfo& ri) {
    int i = vfind(w, ri.key);
    if (i >= 0) return i;
    if (w.nv >= kMaxVerts) { w.overflow = true; return -1; }
    i = w.nv++;
    w.vkey[i] = ri.key;
    w.vcount[i] = 0;
    w.vinfo[i] = ri;
    return i;
}
CAGR_HD inline void push_edge(Work& w, int64_t a, int64_t b) {
    if (w.ne >= kMaxEdges) { w.overflow = true; return; }
    w.ea[w.ne] = a;
    w.eb[w.ne] = b;
    ++w.ne;
}

// iso edges of leaf n1 on face f (FaceEdgesFunction::Function, Octree.cpp:1203-1248)
CAGR_HD inline void face_edges_of(const View& t, const Node& n1, int f, Work& w) {
    if (is_internal(t, n1)) return;
    const int mc = mc_index(t, n1);
    if (!has_roots(mc)) return;
    const int8_t* tri = kTriTable[mc];
    for (int j = 0; j < 15 && tri[j] >= 0; j += 3)
        for (int k = 0; k < 3; ++k) {
            const int e1 = tri[j + k], e2 = tri[j + (k + 1) % 3];
            if (face_of_edges(e1, e2) != f) continue;
            RootInfo r1, r2;
            if (root_index(t, n1, e1, r1) && root_index(t, n1, e2, r2)) {
                push_edge(w, r2.key, r1.key);
                const int i1 = vget(w, r1), i2 = vget(w, r2);
                if (i1 < 0 || i2 < 0) return;
                w.vcount[i1]--;
                w.vcount[i2]++;
            }
        }
}

CAGR_HD inline bool own_root(const View& t, const Node& n, int e, Work& w, RootInfo& ri) {
    if (w.own_state[e] == 0) w.own_state[e] = root_index(t, n, e, w.own[e]) ? 1 : 2;
    ri = w.own[e];
    return w.own_state[e] == 1;
}

// Octree::GetMCIsoEdges (Octree.cpp:900-1006)
CAGR_HD inline void iso_edges(const View& t, const Node& n, Work& w) {
    w.ne = 0;
    w.nv = 0;
    for (int e = 0; e < 12; ++e) w.own_state[e] = 0;
    const int mc = mc_index(t, n);
    const int8_t* tri = kTriTable[mc];
    for (int f = 0; f < 6; ++f) {
        const int ref = f ^ 1;
        Node nb;
        const bool boundary = n.d <= 0 ||
                              (((n.o[f >> 1] << 1) + ((f & 1) << 1)) % (2 << n.d)) == 0;
        if (face_nbr_internal(t, n, f, nb) && !boundary) {
            // ReconOctNode::processNodeFaces (ReconOctNode.cpp:303-329): the
            // 4 children on face `ref` level by level, then each one's own
            int fc[4];
            face_corners(ref, fc);
            Node st[kMaxWalk];
            int top = 0;
            st[top++] = nb;
            while (top > 0) {
                const Node p = st[--top];
                Node ch[4];
                for (int q = 0; q < 4; ++q) {
                    ch[q] = child(p, fc[q]);
                    face_edges_of(t, ch[q], ref, w);
                }
                for (int q = 3; q >= 0; --q)
                    if (is_internal(t, ch[q])) {
                        if (top >= kMaxWalk) { w.overflow = true; return; }
                        st[top++] = ch[q];
                    }
            }
        } else {
            for (int j = 0; j < 15 && tri[j] >= 0; j += 3)
                for (int k = 0; k < 3; ++k) {
                    const int e1 = tri[j + k], e2 = tri[j + (k + 1) % 3];
                    if (face_of_edges(e1, e2) != f) continue;
                    RootInfo r1, r2;
                    if (own_root(t, n, e1, w, r1) && own_root(t, n, e2, w, r2)) {
                        push_edge(w, r1.key, r2.key);
                        const int i1 = vget(w, r1), i2 = vget(w, r2);
                        if (i1 < 0 || i2 < 0) return;
                        w.vcount[i1]++;
                        w.vcount[i2]--;
                    }
                }
        }
    }
    // close the loops across finer neighbours (Octree.cpp:966-1005)
    for (int i = 0; i < w.ne; ++i) {
        int v = vfind(w, w.ea[i]);
        if (v >= 0 && w.vcount[v]) {
            RootInfo pr;
            if (root_pair(t, w.vinfo[v], pr)) {
                const int pv = vfind(w, pr.key);
                if (pv >= 0) {
                    push_edge(w, pr.key, w.ea[i]);
                    w.vcount[pv]++;
                    w.vcount[v]--;
                }
            }
        }
        v = vfind(w, w.eb[i]);
        if (v >= 0 && w.vcount[v]) {
            RootInfo pr;
            if (root_pair(t, w.vinfo[v], pr)) {
                const int pv = vfind(w, pr.key);
                if (pv >= 0) {
                    push_edge(w, w.eb[i], pr.key);
                    w.vcount[v]++;
                    w.vcount[pv]--;
                }
            }
        }
        if (w.overflow) return;
    }
}

// Octree::GetEdgeLoops (Octree.cpp:1008-1062): removes the next loop from
// the edge list (swap-with-last removal, as the reference) into w.loop;
// returns its length, 0 when no edge is left
CAGR_HD inline int next_loop(Work& w) {
    if (w.ne == 0) return 0;
    const int64_t e0a = w.ea[0], e0b = w.eb[0];
    w.ea[0] = w.ea[w.ne - 1];
    w.eb[0] = w.eb[w.ne - 1];
    --w.ne;
    int64_t front_idx = e0b, back_idx = e0a;
    // front / back chains kept in the loop buffer from both ends
    int nf = 0, nb = 0;
    const Lanes<int64_t, kMaxEdges> fr = w.loop;              // front[0..nf) firsts
    const Lanes<int64_t, kMaxEdges> back = w.back;
    for (int j = w.ne - 1; j >= 0; --j) {
        if (w.ea[j] == front_idx || w.eb[j] == front_idx) {
            int64_t first, second;
            if (w.ea[j] == front_idx) { first = w.ea[j]; second = w.eb[j]; }
            else { first = w.eb[j]; second = w.ea[j]; }
            front_idx = second;
            fr[nf++] = first;
            w.ea[j] = w.ea[w.ne - 1];
            w.eb[j] = w.eb[w.ne - 1];
            --w.ne;
            j = w.ne;
        } else if (w.ea[j] == back_idx || w.eb[j] == back_idx) {
            int64_t first, second;
            if (w.eb[j] == back_idx) { first = w.ea[j]; second = w.eb[j]; }
            else { first = w.eb[j]; second = w.ea[j]; }
            (void)second;
            back_idx = first;
            back[nb++] = first;
            w.ea[j] = w.ea[w.ne - 1];
            w.eb[j] = w.eb[w.ne - 1];
            --w.ne;
            j = w.ne;
        }
    }
    // loop = reversed back, e, front (firsts of each edge)
    const int len = nb + 1 + nf;
    for (int q = nf - 1; q >= 0; --q) w.loop[nb + 1 + q] = fr[q];
    for (int q = 0; q < nb; ++q) w.loop[q] = back[nb - 1 - q];
    w.loop[nb] = e0a;
    return len;
}

// ------------------------------------------------------------ triangulation
CAGR_HD inline double tri_area(const Work& w, int p1, int p2, int p3) {   // Triangulation::area
    float q1[3], q2[3];
    for (int k = 0; k < 3; ++k) {
        q1[k] = w.tp[p2][k] - w.tp[p1][k];
        q2[k] = w.tp[p3][k] - w.tp[p1][k];
    }
    const float x = q1[1] * q2[2] - q1[2] * q2[1];
    const float y = -q1[0] * q2[2] + q1[2] * q2[0];
    const float z = q1[0] * q2[1] - q1[1] * q2[0];
    const double s = (double)x * (double)x + (double)y * (double)y + (double)z * (double)z;
    return (double)(float)sqrt(s);
}
CAGR_HD inline int te_find(const Work& w, int p1, int p2) {
    for (int i = 0; i < w.nte; ++i)
        if ((w.te_p[i][0] == p1 && w.te_p[i][1] == p2) || (w.te_p[i][0] == p2 && w.te_p[i][1] == p1))
            return i;
    return -1;
}
CAGR_HD inline bool tri_factor(const Work& w, int t, int& p1, int& p2, int& p3) {
    int p[3];
    for (int k = 0; k < 3; ++k) {
        const int e = w.tt_e[t][k];
        if (e < 0) return false;
        p[k] = (w.te_t[e][0] == t) ? w.te_p[e][0] : w.te_p[e][1];
    }
    p1 = p[0];
    p2 = p[1];
    p3 = p[2];
    return true;
}
CAGR_HD inline void tri_add(Work& w, int p1, int p2, int p3) {           // Triangulation::addTriangle
    const int p[3] = {p1, p2, p3};
    const int t = w.ntt++;
    w.tt_e[t][0] = w.tt_e[t][1] = w.tt_e[t][2] = -1;
    for (int i = 0; i < 3; ++i) {
        const int a = p[i], b = p[(i + 1) % 3];
        int e = te_find(w, a, b);
        if (e < 0) {
            e = w.nte++;
            w.te_p[e][0] = a;
            w.te_p[e][1] = b;
            w.te_t[e][0] = t;
            w.te_t[e][1] = -1;
        } else if (w.te_p[e][0] == a) {
            if (w.te_t[e][0] < 0) w.te_t[e][0] = t;
            else return;
        } else {
            if (w.te_t[e][1] < 0) w.te_t[e][1] = t;
            else return;
        }
        w.tt_e[t][i] = e;
    }
}
CAGR_HD inline bool tri_flip(Work& w, int e) {                           // Triangulation::flipMinimize
    if (w.te_t[e][0] < 0 || w.te_t[e][1] < 0) return false;
    int P[3], Q[3];
    if (!tri_factor(w, w.te_t[e][0], P[0], P[1], P[2])) return false;
    if (!tri_factor(w, w.te_t[e][1], Q[0], Q[1], Q[2])) return false;
    const double old_area = tri_area(w, P[0], P[1], P[2]) + tri_area(w, Q[0], Q[1], Q[2]);
    int ip, iq;
    for (ip = 0; ip < 3; ++ip) {
        int i;
        for (i = 0; i < 3; ++i) if (P[ip] == Q[i]) break;
        if (i == 3) break;
    }
    for (iq = 0; iq < 3; ++iq) {
        int i;
        for (i = 0; i < 3; ++i) if (P[i] == Q[iq]) break;
        if (i == 3) break;
    }
    if (ip == 3 || iq == 3) return false;
    const int nP[3] = {P[ip], P[(ip + 1) % 3], Q[iq]};
    const int nQ[3] = {Q[iq], P[(ip + 2) % 3], P[ip]};
    const double new_area = tri_area(w, nP[0], nP[1], nP[2]) + tri_area(w, nQ[0], nQ[1], nQ[2]);
    if (old_area <= new_area) return false;
    w.te_p[e][0] = nP[0];
    w.te_p[e][1] = nQ[0];
    const int t0 = w.te_t[e][0], t1 = w.te_t[e][1];
    for (int i = 0; i < 3; ++i) {
        int idx = te_find(w, nQ[i], nQ[(i + 1) % 3]);
        w.tt_e[t0][i] = idx;
        if (idx != e) {
            if (w.te_t[idx][0] == t1) w.te_t[idx][0] = t0;
            if (w.te_t[idx][1] == t1) w.te_t[idx][1] = t0;
        }
        idx = te_find(w, nP[i], nP[(i + 1) % 3]);
        w.tt_e[t1][i] = idx;
        if (idx != e) {
            if (w.te_t[idx][0] == t0) w.te_t[idx][0] = t1;
            if (w.te_t[idx][1] == t0) w.te_t[idx][1] = t1;
        }
    }
    return true;
}

}  // namespace mc
}  // namespace cagr

namespace cagr {
namespace mc {

// the triangles of one leaf (Octree::GetMCIsoTriangles(node, ...) with
// AddTriangles, Octree.cpp:1064-1200): each loop of iso edges becomes a fan
// from its first vertex, flipped to minimum area; `vindex` maps a root key to
// its vertex number (-1 if none: the reference skips such a vertex).  Emits
// (a, b, c) through `emit` in the reference's order; returns the count.
template <class VIndex, class Emit>
CAGR_HD inline int leaf_triangles(const View& t, const Node& leaf, Work& w, VIndex vindex,
                                  Emit emit) {
    w.overflow = false;
    iso_edges(t, leaf, w);
    if (w.overflow) return -1;
    int count = 0;
    int len;
    while ((len = next_loop(w)) > 0) {
        const Lanes<int, kMaxEdges> vid = w.vid;
        int nv = 0;
        for (int j = 0; j < len; ++j) {
            const int v = vindex(w.loop[j]);
            if (v >= 0) vid[nv++] = v;
        }
        if (nv > 3) {
            w.nte = w.ntt = 0;
            for (int j = 0; j < nv; ++j) vindex.position(vid[j], w.tp[j].v);
            for (int j = 1; j < nv - 1; ++j) tri_add(w, 0, j, j + 1);
            for (;;) {
                int e;
                for (e = 0; e < w.nte; ++e)
                    if (tri_flip(w, e)) break;
                if (e == w.nte) break;
            }
            for (int q = 0; q < w.ntt; ++q) {
                int a = 0, b = 0, c = 0;
                tri_factor(w, q, a, b, c);
                emit(vid[a], vid[b], vid[c]);
                ++count;
            }
        } else if (nv == 3) {
            emit(vid[0], vid[1], vid[2]);
            ++count;
        }
    }
    return count;
}

}  // namespace mc
}  // namespace cagr
