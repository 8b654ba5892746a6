// CPU replay of the HIP tree build (test only).  Compiles the shared
// per-element logic of cagr_amd/csrc/tree_logic.h with g++ and runs the same
// sequence the kernels run -- keys, stable sort, chains, scan, node records,
// leaf sorts, the depth order and the child positions -- serially, then writes
// the reference serialised layout so the test can compare it with the oracle
// tree.
#define CAGR_HD
#include "../../cagr_amd/csrc/tree_logic.h"
#include "../../cagr_amd/csrc/tree_export.h"
#include <algorithm>
#include <numeric>
#include <vector>

using namespace cagr;

extern "C" int64_t hl_build(const double* pts, int64_t n, int max_depth, int S, double theta,
                            int64_t* m_out, TopoRec* topo_out /* may be null */,
                            int32_t* perm_out /* may be null */) {
    std::vector<uint64_t> key(n);
    for (int64_t i = 0; i < n; ++i)
        key[i] = morton_key(cell_index(pts[3 * i]), cell_index(pts[3 * i + 1]),
                            cell_index(pts[3 * i + 2]));
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::vector<uint64_t> sk(n);
    for (int64_t i = 0; i < n; ++i) sk[i] = key[perm[i]];
    std::vector<int8_t> ld(n), sd(n);
    std::vector<int32_t> cnt(n), off(n + 1, 0);
    for (int64_t b = 0; b < n; ++b) chain_body(b, sk.data(), n, S, max_depth, ld.data(), sd.data(), cnt.data());
    for (int64_t b = 0; b < n; ++b) off[b + 1] = off[b] + cnt[b];
    const int64_t m = off[n];
    std::vector<TopoRec> topo(m);
    for (int64_t b = 0; b < n; ++b) node_body(b, sk.data(), n, ld.data(), sd.data(), off.data(), 2.0 / theta, topo.data());
    for (int64_t i = 0; i < m; ++i)
        if (topo_leaf(topo[i])) leaf_sort_body(topo[i], perm.data());
    *m_out = m;
    if (topo_out) std::copy(topo.begin(), topo.end(), topo_out);
    if (perm_out) std::copy(perm.begin(), perm.end(), perm_out);
    return m;
}

// keys of the sorted points are needed by the export; recompute them here
extern "C" int64_t hl_export(const double* pts, int64_t n, const TopoRec* topo, int64_t m,
                             const int32_t* perm, int64_t* parent, int64_t* children,
                             uint8_t* is_leaf, double* half_w, int64_t* num_points,
                             int64_t* n2p_index, int64_t* n2p_start) {
    std::vector<uint64_t> sk(n);
    for (int64_t i = 0; i < n; ++i) {
        const double* p = pts + 3 * perm[i];
        sk[i] = morton_key(cell_index(p[0]), cell_index(p[1]), cell_index(p[2]));
    }
    return export_reference_layout(topo, m, perm, sk.data(), parent, children, is_leaf, half_w,
                                   num_points, n2p_index, n2p_start);
}

// Forest replay: clouds [off[c], off[c+1]) of pts sorted together with the
// cloud id above the Morton bits.  Writes the forest's topology records and
// permutation (global indices) and, per cloud, its root and stop node ids.
extern "C" int64_t hl_build_forest(const double* pts, const int64_t* coff, int64_t nclouds,
                                   int max_depth, int S, double theta, TopoRec* topo_out,
                                   int32_t* perm_out, int32_t* roots_out) {
    const int64_t n = coff[nclouds];
    std::vector<uint64_t> key(n);
    for (int64_t c = 0; c < nclouds; ++c)
        for (int64_t i = coff[c]; i < coff[c + 1]; ++i)
            key[i] = forest_key((uint64_t)c, morton_key(cell_index(pts[3 * i]), cell_index(pts[3 * i + 1]),
                                                        cell_index(pts[3 * i + 2])));
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::vector<uint64_t> sk(n);
    for (int64_t i = 0; i < n; ++i) sk[i] = key[perm[i]];
    std::vector<int8_t> ld(n), sd(n);
    std::vector<int32_t> cnt(n), off(n + 1, 0);
    for (int64_t b = 0; b < n; ++b) chain_body(b, sk.data(), n, S, max_depth, ld.data(), sd.data(), cnt.data());
    for (int64_t b = 0; b < n; ++b) off[b + 1] = off[b] + cnt[b];
    const int64_t m = off[n];
    std::vector<TopoRec> topo(m);
    for (int64_t b = 0; b < n; ++b) node_body(b, sk.data(), n, ld.data(), sd.data(), off.data(), 2.0 / theta, topo.data());
    for (int64_t i = 0; i < m; ++i)
        if (topo_leaf(topo[i])) leaf_sort_body(topo[i], perm.data());
    if (topo_out) std::copy(topo.begin(), topo.end(), topo_out);
    if (perm_out) std::copy(perm.begin(), perm.end(), perm_out);
    if (roots_out)
        for (int64_t c = 0; c <= nclouds; ++c) roots_out[c] = off[coff[c]];
    return m;
}

// Depth order (node ids stably sorted by depth) and the child positions of
// every internal node (kids_body over the positions of depth >= 1) of the
// records topo[m] of hl_build / hl_build_forest; writes order[m],
// kids[m] (-1: none) and nkids[m].
extern "C" void hl_kids(TopoRec* topo, int64_t m, int32_t* order, int32_t* kids, int32_t* nkids) {
    std::iota(order, order + m, 0);
    std::stable_sort(order, order + m,
                     [&](int32_t a, int32_t b) { return topo_depth(topo[a]) < topo_depth(topo[b]); });
    int64_t i0 = 0;
    while (i0 < m && topo_depth(topo[order[i0]]) == 0) ++i0;
    for (int64_t i = i0; i < m; ++i) kids_body(i, topo, order, m);
    for (int64_t i = 0; i < m; ++i) {
        kids[i] = topo[i].kids;
        nkids[i] = topo_nkids(topo[i]);
    }
}
