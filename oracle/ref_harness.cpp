// C-ABI harness around the reference's torch-free CPU core -- TEST
// INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with the
// reference's own sources, where they lie, into oracle/_ref/libcagr_ref.so:
//   CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_treeutils.cpp
//   CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_kernels.cpp
// Nothing here re-implements the reference: every function only drives the
// reference's templates the way its pybind layer does
// (wn_treecode_cpu_torch_interface.cpp:41-210 for build + upward pass,
// :212-380 for the A / AT / G launchers), with plain buffers instead of torch
// tensors.  The checked results pin the oracle restatement in
// oracle/cagr_oracle.cpp.
#include "wn_treecode_cpu.h"
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
template <typename T>
struct RefTree {
    int64_t num_nodes = 0, num_leaves = 0, depth = 0;
    std::vector<signedindex_t> parent, children, num_points, n2p_start, n2p_index;
    std::vector<bool> is_leaf_dummy;
    std::vector<uint8_t> is_leaf;
    std::vector<T> half_w;
};

template <typename T>
RefTree<T>* build(const T* pts, int64_t n, int64_t max_depth, int64_t leaf_size) {
    auto* rt = new RefTree<T>;
    std::vector<signedindex_t> ids((size_t)n);
    for (int64_t i = 0; i < n; ++i) ids[i] = i;
    signedindex_t cur = 0;
    const signedindex_t md = max_depth, ls = leaf_size;
    OctNode<T>* root = build_tree_cpu_recursive<T>(pts, ids, nullptr, T(0), T(0), T(0), T(1), 0,
                                                   cur, md, ls);
    signedindex_t nn = 0, nl = 0, dp = 0;
    compute_tree_attributes<T>(root, nn, nl, dp);
    rt->num_nodes = nn;
    rt->num_leaves = nl;
    rt->depth = dp;
    rt->parent.assign(nn, 0);
    rt->children.assign(nn * NUM_OCT_CHILDREN, 0);
    rt->num_points.assign(nn, 0);
    rt->n2p_start.assign(nn, 0);
    rt->half_w.assign(nn, T(0));
    bool* leaf = new bool[nn > 0 ? nn : 1];
    std::vector<signedindex_t> list;
    serialize_tree_recursive<T>(root, rt->parent.data(), rt->children.data(), leaf,
                                rt->half_w.data(), rt->num_points.data(), rt->n2p_start.data(),
                                list);
    rt->is_leaf.assign(leaf, leaf + nn);
    delete[] leaf;
    rt->n2p_index = list;
    free_tree_recursive<T>(root);
    return rt;
}

template <typename T>
void export_tree(const RefTree<T>* t, int64_t* parent, int64_t* children, uint8_t* is_leaf,
                 double* half_w, int64_t* num_points, int64_t* n2p_index, int64_t* n2p_start) {
    const size_t m = (size_t)t->num_nodes;
    for (size_t i = 0; i < m; ++i) {
        parent[i] = t->parent[i];
        is_leaf[i] = t->is_leaf[i];
        half_w[i] = (double)t->half_w[i];
        num_points[i] = t->num_points[i];
        n2p_start[i] = t->n2p_start[i];
    }
    for (size_t i = 0; i < m * NUM_OCT_CHILDREN; ++i) children[i] = t->children[i];
    for (size_t i = 0; i < t->n2p_index.size(); ++i) n2p_index[i] = t->n2p_index[i];
}
}  // namespace

extern "C" {

void* ref_tree_build_f64(const double* pts, int64_t n, int64_t max_depth, int64_t leaf_size) {
    return build<double>(pts, n, max_depth, leaf_size);
}
void* ref_tree_build_f32(const float* pts, int64_t n, int64_t max_depth, int64_t leaf_size) {
    return build<float>(pts, n, max_depth, leaf_size);
}
void ref_tree_info(const void* h, int64_t is_f32, int64_t* num_nodes, int64_t* num_leaves,
                   int64_t* depth, int64_t* list_len) {
    if (is_f32) {
        auto* t = (const RefTree<float>*)h;
        *num_nodes = t->num_nodes; *num_leaves = t->num_leaves; *depth = t->depth;
        *list_len = (int64_t)t->n2p_index.size();
    } else {
        auto* t = (const RefTree<double>*)h;
        *num_nodes = t->num_nodes; *num_leaves = t->num_leaves; *depth = t->depth;
        *list_len = (int64_t)t->n2p_index.size();
    }
}
void ref_tree_export(const void* h, int64_t is_f32, int64_t* parent, int64_t* children,
                     uint8_t* is_leaf, double* half_w, int64_t* num_points, int64_t* n2p_index,
                     int64_t* n2p_start) {
    if (is_f32)
        export_tree((const RefTree<float>*)h, parent, children, is_leaf, half_w, num_points,
                    n2p_index, n2p_start);
    else
        export_tree((const RefTree<double>*)h, parent, children, is_leaf, half_w, num_points,
                    n2p_index, n2p_start);
}
void ref_tree_free(void* h, int64_t is_f32) {
    if (is_f32) delete (RefTree<float>*)h;
    else delete (RefTree<double>*)h;
}

// The reference's upward pass in FP64: leaf launcher, then `tree_depth`
// rounds of find_next + nonleaf (wn_treecode_cpu_torch_interface.cpp:120-210).
void ref_upward_f64(int64_t m, const int64_t* parent, const int64_t* children,
                    const uint8_t* is_leaf_u8, const int64_t* num_points,
                    const int64_t* n2p_index, const int64_t* n2p_start, const double* pts,
                    const double* weights, const double* attrs, int64_t c_dim,
                    int64_t tree_depth, double* node_attrs, double* node_rep, double* node_w) {
    std::vector<uint8_t> leaf_store(is_leaf_u8, is_leaf_u8 + m);
    bool* is_leaf = new bool[m > 0 ? m : 1];
    bool* scattered = new bool[m > 0 ? m : 1];
    bool* next = new bool[m > 0 ? m : 1];
    for (int64_t i = 0; i < m; ++i) { is_leaf[i] = leaf_store[i] != 0; scattered[i] = false; next[i] = false; }
    std::memset(node_attrs, 0, sizeof(double) * m * c_dim);
    std::memset(node_rep, 0, sizeof(double) * m * 3);
    std::memset(node_w, 0, sizeof(double) * m);
    const signedindex_t* par = (const signedindex_t*)parent;
    const signedindex_t* ch = (const signedindex_t*)children;
    const signedindex_t* np = (const signedindex_t*)num_points;
    const signedindex_t* idx = (const signedindex_t*)n2p_index;
    const signedindex_t* st = (const signedindex_t*)n2p_start;
    scatter_point_attrs_to_nodes_leaf_cpu_kernel_launcher<double>(
        par, pts, weights, attrs, idx, st, np, is_leaf, scattered, node_attrs, node_rep, node_w,
        c_dim, m);
    for (int64_t d = tree_depth - 1; d >= 0; --d) {
        find_next_to_scatter_cpu_kernel_launcher<double>(ch, is_leaf, scattered, next, idx, m);
        scatter_point_attrs_to_nodes_nonleaf_cpu_kernel_launcher<double>(
            par, ch, pts, weights, attrs, idx, st, np, is_leaf, scattered, next, node_attrs,
            node_rep, node_w, c_dim, m);
    }
    delete[] is_leaf;
    delete[] scattered;
    delete[] next;
}

// op: 0 = multiply_by_A (CPU sign), 1 = multiply_by_AT, 2 = multiply_by_G
void ref_apply_f64(int64_t op, int64_t nq, const double* q, const double* qw, const double* pts,
                   const double* attrs, int64_t m, const int64_t* children,
                   const uint8_t* is_leaf_u8, const double* half_w, const int64_t* num_points,
                   const int64_t* n2p_index, const int64_t* n2p_start, const double* node_attrs,
                   const double* node_rep, int64_t continuous, double* out) {
    bool* is_leaf = new bool[m > 0 ? m : 1];
    for (int64_t i = 0; i < m; ++i) is_leaf[i] = is_leaf_u8[i] != 0;
    const signedindex_t* ch = (const signedindex_t*)children;
    const signedindex_t* np = (const signedindex_t*)num_points;
    const signedindex_t* idx = (const signedindex_t*)n2p_index;
    const signedindex_t* st = (const signedindex_t*)n2p_start;
    if (op == 0)
        multiply_by_A_cpu_kernel_launcher<double>(q, qw, pts, attrs, idx, st, ch, node_attrs,
                                                  is_leaf, half_w, node_rep, np, out, nq,
                                                  continuous != 0);
    else if (op == 1)
        multiply_by_AT_cpu_kernel_launcher<double>(q, qw, pts, attrs, idx, st, ch, node_attrs,
                                                   is_leaf, half_w, node_rep, np, out, nq);
    else
        multiply_by_G_cpu_kernel_launcher<double>(q, qw, pts, attrs, idx, st, ch, node_attrs,
                                                  is_leaf, half_w, node_rep, np, out, nq);
    delete[] is_leaf;
}

}  // extern "C"
