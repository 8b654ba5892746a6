// CPU replay of the HIP PGR marching cubes (test only).  Compiles the shared
// per-leaf logic of cagr_amd/csrc/mc_logic.h with g++ and runs it serially in
// the order the kernels emit: vertices by first occurrence over (leaf
// preorder, edge slot), then triangles leaf by leaf.  The test compares the
// result with the reference's own mesh.
#define CAGR_HD
#define CAGR_MC_STRIDE 1
#include <cstdio>
#include <cstdlib>
// every workspace index is checked against its array's length here
#define CAGR_MC_CHECK(cond)                                                          \
    ((cond) ? (void)0                                                                \
            : (std::fprintf(stderr, "mc_replay: index out of range: %s (%s:%d)\n",  \
                            #cond, __FILE__, __LINE__),                              \
               std::abort()))
#include <cmath>
using std::sqrt;
#include "../../cagr_amd/csrc/mc_logic.h"

#include <algorithm>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

using namespace cagr::mc;

namespace {
std::vector<float> g_pts;
std::vector<int32_t> g_tri;
}

// leaves [nl,4] (depth, x, y, z) in preorder; values [G] in gridDataVector order
extern "C" int hl_mesh(const int32_t* leaves, int64_t nl, int md, const float* values,
                       float iso, float smooth_width, int64_t* nv, int64_t* nt) {
    // internal nodes = strict ancestors of the leaves
    std::vector<uint64_t> in;
    for (int64_t r = 0; r < nl; ++r) {
        Node n{leaves[4 * r], {leaves[4 * r + 1], leaves[4 * r + 2], leaves[4 * r + 3]}};
        while (n.d > 0) {
            n = parent(n);
            in.push_back(node_key(n.d, n.o[0], n.o[1], n.o[2]));
        }
    }
    std::sort(in.begin(), in.end());
    in.erase(std::unique(in.begin(), in.end()), in.end());
    View t{};
    t.md = md;
    t.in_keys = in.data();
    t.n_in = (int64_t)in.size();
    // corner keys in gridDataVector order (first occurrence)
    std::vector<std::pair<int64_t, int32_t>> ck;
    std::map<int64_t, int32_t> seen;
    for (int64_t r = 0; r < nl; ++r) {
        Node n{leaves[4 * r], {leaves[4 * r + 1], leaves[4 * r + 2], leaves[4 * r + 3]}};
        for (int c = 0; c < 8; ++c) {
            const int64_t k = corner_key(t, n, c);
            if (!seen.count(k)) seen.emplace(k, (int32_t)seen.size());
        }
    }
    std::vector<int64_t> keys;
    std::vector<int32_t> idx;
    for (const auto& kv : seen) {
        keys.push_back(kv.first);
        idx.push_back(kv.second);
    }
    t.ck = keys.data();
    t.ck_idx = idx.data();
    t.n_ck = (int64_t)keys.size();
    t.values = values;
    t.iso = iso;
    t.smooth_width = smooth_width;
    t.grid = (float)(1.0f / std::pow(2.0, md + 1));
    // the per-node marching-cubes indices the kernels precompute
    std::vector<uint64_t> mk(nl);
    std::vector<uint8_t> leaf_mc(nl), in_mc(in.size());
    for (int64_t r = 0; r < nl; ++r) {
        Node n{leaves[4 * r], {leaves[4 * r + 1], leaves[4 * r + 2], leaves[4 * r + 3]}};
        mk[r] = morton_left(n, md - 1);
        leaf_mc[r] = (uint8_t)mc_from_corners(t, n);
    }
    for (size_t i = 0; i < in.size(); ++i) {
        const uint64_t k = in[i];
        Node n{(int)(k >> 39), {(int32_t)(k & 0x1FFF), (int32_t)((k >> 13) & 0x1FFF),
                                (int32_t)((k >> 26) & 0x1FFF)}};
        in_mc[i] = (uint8_t)mc_from_corners(t, n);
    }
    t.in_mc = in_mc.data();
    t.leaf_mk = mk.data();
    t.leaf_mc = leaf_mc.data();
    t.n_leaves = nl;

    // vertices
    std::map<int64_t, int32_t> vid;
    g_pts.clear();
    g_tri.clear();
    for (int64_t r = 0; r < nl; ++r) {
        Node n{leaves[4 * r], {leaves[4 * r + 1], leaves[4 * r + 2], leaves[4 * r + 3]}};
        int64_t rk[12];
        RootInfo ri[12];
        const int cnt = leaf_roots(t, n, rk, ri);
        for (int s = 0; s < cnt; ++s) {
            if (rk[s] < 0 || vid.count(rk[s])) continue;
            vid.emplace(rk[s], (int32_t)(g_pts.size() / 3));
            float p[3];
            root_position(t, ri[s], p);
            g_pts.insert(g_pts.end(), p, p + 3);
        }
    }
    // triangles
    struct VI {
        std::map<int64_t, int32_t>* m;
        const std::vector<float>* pts;
        int operator()(int64_t k) const {
            auto it = m->find(k);
            return it == m->end() ? -1 : it->second;
        }
        void position(int v, float p[3]) const {
            for (int k = 0; k < 3; ++k) p[k] = (*pts)[3 * v + k];
        }
    } vi{&vid, &g_pts};
    std::vector<uint64_t> wbuf(WorkLayout::bytes / 8 + 1);
    Work w(reinterpret_cast<unsigned char*>(wbuf.data()), 0);
    for (int64_t r = 0; r < nl; ++r) {
        Node n{leaves[4 * r], {leaves[4 * r + 1], leaves[4 * r + 2], leaves[4 * r + 3]}};
        const int c = leaf_triangles(t, n, w, vi, [&](int a, int b, int cc) {
            g_tri.push_back(a);
            g_tri.push_back(b);
            g_tri.push_back(cc);
        });
        if (c < 0) return 2;
    }
    *nv = (int64_t)g_pts.size() / 3;
    *nt = (int64_t)g_tri.size() / 3;
    return 0;
}

extern "C" void hl_mesh_fetch(float* pts, int32_t* tri) {
    std::copy(g_pts.begin(), g_pts.end(), pts);
    std::copy(g_tri.begin(), g_tri.end(), tri);
}
