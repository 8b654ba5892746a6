// CAGR treecode oracle -- TEST INFRASTRUCTURE ONLY.
//
// A from-scratch FP64 CPU restatement of the reference's treecode Gauss
// operator, used by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg as the checker.  Nothing in the cagr_amd product path
// links, loads or calls this file.
//
// What it restates (reference paths relative to the reference repo root):
//   * octree build + serialisation
//       CUDA-based/ext/wn_treecode/wn_treecode_cpu/wn_treecode_cpu_treeutils.cpp:90-177, 202-258
//       root box and leaf rule: .../wn_treecode_cpu_torch_interface.cpp:41-102
//   * upward pass ("scatter_point_attrs_to_nodes")
//       .../wn_treecode_cpu/wn_treecode_cpu_kernels.cpp:173-433
//       .../wn_treecode_cuda/wn_treecode_cuda_kernels.cu:419-553
//   * traversal + opening test, shared by every op
//       .../wn_treecode_cuda/wn_treecode_cuda_kernels.cu:556-621 (and its twins)
//   * kernel math: A (both sign conventions), AT, G, A2_{1,2,3}, A2T{1,2,3}, A3
//       .../wn_treecode_cuda/wn_treecode_cuda_kernels.cu:79-414, 694-763, 843-1257
//       .../wn_treecode_cpu/wn_treecode_cpu_kernels.cpp:77-167 (CPU sign of A)
//
// Pinning (what checks this restatement; see tests/test_oracle.py):
//   * tree topology, node moments, A/AT/G: bit-for-bit against the reference's
//     own torch-free C++ core compiled into oracle/_ref by oracle/Makefile;
//   * every op: brute-force direct sums (opening factor -> infinity), the
//     An = 0 reduction (A2k -> A, A2Tk -> AT, A3 -> A), the sphere winding
//     number closed form, the dense identity A2Tk(An) = (-A2k(-An))^T,
//     adjoint identities, and the reference-held CAGR output bunny.xyz.
//   * PARITY UNPINNED: the An-dependent part of the continuous inner branch
//     of A2k / A3 (inside the width: e^{An w} [d.mu/w^3 + An/2 mu_k/w],
//     cuda_kernels.cu:140-147, 179-185, 216-222).  Only its An = 0 part
//     (d.mu/w^3, the CPU core's continuous A negated) is pinned; the factor
//     e^{An w} and the An/2 mu_k/w term are restated from the source and
//     checked by nothing reference-derived (the reference holds no output of
//     its GaussRecon query, and its CUDA kernels cannot be built here).
//
// All arithmetic is IEEE double with contraction off (-ffp-contract=off in
// oracle/Makefile), evaluated in the same operation order as the reference.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr int kKids = 8;

// ---------------------------------------------------------------- tree build
// The reference builds the tree recursively from the full ascending index list
// (treeutils.cpp:90-177): root centre (0,0,0), half-width 1
// (torch_interface.cpp:55-56); a node is a leaf iff depth >= max_depth or it
// holds <= leaf_size points (treeutils.cpp:122); octant = (x>=cx) + 2(y>=cy) +
// 4(z>=cz) (treeutils.cpp:136-139); child centre = c -/+ half_w/2 per axis
// (treeutils.cpp:143-168); node ids are assigned in DFS preorder, children in
// octant order (treeutils.cpp:109,119).  Each node keeps its own point list in
// ascending original order; serialisation concatenates them in preorder
// (treeutils.cpp:202-258).
struct Tree {
    int64_t n_points = 0;
    int64_t depth = 0;
    int64_t n_leaves = 0;
    std::vector<int64_t> parent, children, num_points, n2p_start, n2p_index;
    std::vector<uint8_t> is_leaf;
    std::vector<double> half_w;
};

struct Builder {
    const double* pts;
    int64_t max_depth, leaf_size;
    Tree* t;
    std::vector<int64_t> scratch;

    int64_t node(std::vector<int64_t>& ids, int64_t lo, int64_t hi, int64_t par,
                 double cx, double cy, double cz, double hw, int64_t depth) {
        const int64_t cnt = hi - lo;
        const int64_t me = (int64_t)t->parent.size();
        t->parent.push_back(par);
        t->children.insert(t->children.end(), kKids, -1);
        t->is_leaf.push_back(0);
        t->half_w.push_back(hw);
        t->num_points.push_back(cnt);
        t->n2p_start.push_back((int64_t)t->n2p_index.size());
        // the node's list, captured before its span is partitioned (ascending)
        t->n2p_index.insert(t->n2p_index.end(), ids.begin() + lo, ids.begin() + hi);
        if (depth > t->depth) t->depth = depth;

        if ((max_depth >= 0 && depth >= max_depth) || cnt <= leaf_size) {
            t->is_leaf[me] = 1;
            t->n_leaves++;
            return me;
        }
        // stable 8-way partition of ids[lo, hi) by octant
        int64_t cnt8[kKids] = {0};
        std::vector<uint8_t> oct((size_t)cnt);
        for (int64_t i = 0; i < cnt; ++i) {
            const double* p = pts + 3 * ids[lo + i];
            int o = 0;
            if (p[0] >= cx) o += 1;
            if (p[1] >= cy) o += 2;
            if (p[2] >= cz) o += 4;
            oct[i] = (uint8_t)o;
            cnt8[o]++;
        }
        int64_t off[kKids + 1];
        off[0] = 0;
        for (int k = 0; k < kKids; ++k) off[k + 1] = off[k] + cnt8[k];
        scratch.resize((size_t)cnt);
        int64_t cur[kKids];
        for (int k = 0; k < kKids; ++k) cur[k] = off[k];
        for (int64_t i = 0; i < cnt; ++i) scratch[cur[oct[i]]++] = ids[lo + i];
        std::copy(scratch.begin(), scratch.begin() + cnt, ids.begin() + lo);

        const double q = hw / 2.0;
        for (int k = 0; k < kKids; ++k) {
            if (cnt8[k] == 0) continue;
            const double ncx = (k & 1) ? cx + q : cx - q;
            const double ncy = (k & 2) ? cy + q : cy - q;
            const double ncz = (k & 4) ? cz + q : cz - q;
            int64_t c = node(ids, lo + off[k], lo + off[k + 1], me, ncx, ncy, ncz, q, depth + 1);
            t->children[me * kKids + k] = c;
        }
        return me;
    }
};

// ---------------------------------------------------------------- kernels
// Restated from wn_treecode_cuda_kernels.cu:79-414.  `an` is the convection
// strength (the reference hard-codes An = 1, wn_treecode_cuda.h:33).
inline double dist2_of(const double* a, const double* b) {
    // get_point2point_dist2: running sum of squared differences, axis order
    double s = 0.0;
    for (int d = 0; d < 3; ++d) { double e = a[d] - b[d]; s += e * e; }
    return s;
}

// eval_A_mu (cuda_kernels.cu:79-110); sign=-1 gives the CPU convention
// (cpu_kernels.cpp:77-103)
inline double k_A(const double* d, const double* mu, double w, bool cont, double sign) {
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = std::sqrt(r2);
    double res = 0.0;
    if (r >= w) {
        double den = r * r2;
        for (int c = 0; c < 3; ++c) res += (sign * d[c] * mu[c]) / den;
    } else if (cont) {
        double den = w * w * w;
        for (int c = 0; c < 3; ++c) res += (sign * d[c] * mu[c]) / den;
    }
    return res;
}

// eval_A2_mu{k+1} (cuda_kernels.cu:112-224)
inline double k_A2(int k, const double* d, const double* mu, double w, bool cont, double an) {
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = std::sqrt(r2);
    double e = std::exp(0.5 * an * d[k]);
    double res = 0.0;
    if (r >= w) {
        double den = r * r2;
        for (int c = 0; c < 3; ++c) {
            double cp = (c == k) ? an : 0.0;
            res += e * ((d[c] * mu[c]) / den + (cp * mu[c]) * 0.5 / r);
        }
    } else if (cont) {
        double den = w * w * w;
        e = std::exp(w * an);
        for (int c = 0; c < 3; ++c) {
            double cp = (c == k) ? an : 0.0;
            res += e * ((d[c] * mu[c]) / den + (cp * mu[c]) * 0.5 / w);
        }
    }
    return res;
}

// eval_AT_s_add_ (cuda_kernels.cu:228-258)
inline void k_AT(double* out, const double* d, double s, double w) {
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = std::sqrt(r2);
    if (r >= w) {
        double den = r * r2;
        for (int c = 0; c < 3; ++c) out[c] += (d[c] * s) / den;
    }
}

// eval_A2T{k+1}_s_add_ (cuda_kernels.cu:262-379)
inline void k_A2T(int k, double* out, const double* d, double s, double w, double an) {
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = std::sqrt(r2);
    double e = std::exp(0.5 * an * d[k]);
    if (r >= w) {
        double den = r * r2;
        for (int c = 0; c < 3; ++c) {
            double cp = (c == k) ? an : 0.0;
            out[c] += e * ((d[c] * s) / den + (cp * s) * 0.5 / r);
        }
    }
}

// eval_G_mu_add_ (cuda_kernels.cu:384-414; cpu_kernels.cpp:137-167)
inline void k_G(double* out, const double* d, const double* mu, double w) {
    double r2 = 0.0;
    for (int c = 0; c < 3; ++c) r2 += d[c] * d[c];
    double r = std::sqrt(r2);
    double r3 = r * r2, r5 = r2 * r3;
    double dm = 0.0;
    for (int c = 0; c < 3; ++c) dm += d[c] * mu[c];
    if (r >= w)
        for (int c = 0; c < 3; ++c) out[c] += (mu[c] / r3 - 3 * d[c] * dm / r5);
}

enum Op {
    OP_A = 0, OP_A_CPU = 1, OP_AT = 2, OP_G = 3,
    OP_A21 = 4, OP_A22 = 5, OP_A23 = 6,
    OP_A2T1 = 7, OP_A2T2 = 8, OP_A2T3 = 9,
    OP_A3 = 10,
};

inline int op_in_dim(int op) {
    return (op == OP_AT || op == OP_A2T1 || op == OP_A2T2 || op == OP_A2T3) ? 1 : 3;
}
inline int op_out_dim(int op) {
    return (op == OP_AT || op == OP_G || op == OP_A2T1 || op == OP_A2T2 || op == OP_A2T3) ? 3 : 1;
}

// one source interaction of `op`, accumulated into out[0..out_dim)
inline void interact(int op, double* out, const double* d, const double* a, double w,
                     bool cont, double an) {
    switch (op) {
        case OP_A: out[0] += k_A(d, a, w, cont, 1.0); break;
        case OP_A_CPU: out[0] += k_A(d, a, w, cont, -1.0); break;
        case OP_AT: k_AT(out, d, a[0], w); break;
        case OP_G: k_G(out, d, a, w); break;
        case OP_A21: case OP_A22: case OP_A23:
            out[0] += k_A2(op - OP_A21, d, a, w, cont, an); break;
        case OP_A2T1: case OP_A2T2: case OP_A2T3:
            k_A2T(op - OP_A2T1, out, d, a[0], w, an); break;
        case OP_A3:   // multiply_by_A3: three adds per interaction (cuda_kernels.cu:735-737)
            out[0] += k_A2(0, d, a, w, cont, an);
            out[0] += k_A2(1, d, a, w, cont, an);
            out[0] += k_A2(2, d, a, w, cont, an);
            break;
    }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- tree API
void* orc_tree_build(const double* pts, int64_t n, int64_t max_depth, int64_t leaf_size) {
    Tree* t = new Tree;
    t->n_points = n;
    if (n <= 0) return t;   // build_tree_cpu_recursive returns nullptr: zero nodes
    std::vector<int64_t> ids((size_t)n);
    for (int64_t i = 0; i < n; ++i) ids[i] = i;
    Builder b{pts, max_depth, leaf_size, t, {}};
    b.node(ids, 0, n, -1, 0.0, 0.0, 0.0, 1.0, 0);
    return t;
}

void orc_tree_info(const void* h, int64_t* num_nodes, int64_t* num_leaves, int64_t* depth,
                   int64_t* list_len) {
    const Tree* t = (const Tree*)h;
    *num_nodes = (int64_t)t->parent.size();
    *num_leaves = t->n_leaves;
    *depth = t->depth;
    *list_len = (int64_t)t->n2p_index.size();
}

void orc_tree_export(const void* h, int64_t* parent, int64_t* children, uint8_t* is_leaf,
                     double* half_w, int64_t* num_points, int64_t* n2p_index,
                     int64_t* n2p_start) {
    const Tree* t = (const Tree*)h;
    const size_t m = t->parent.size();
    std::memcpy(parent, t->parent.data(), m * 8);
    std::memcpy(children, t->children.data(), m * 8 * kKids);
    std::memcpy(is_leaf, t->is_leaf.data(), m);
    std::memcpy(half_w, t->half_w.data(), m * 8);
    std::memcpy(num_points, t->num_points.data(), m * 8);
    std::memcpy(n2p_start, t->n2p_start.data(), m * 8);
    std::memcpy(n2p_index, t->n2p_index.data(), t->n2p_index.size() * 8);
}

void orc_tree_free(void* h) { delete (Tree*)h; }

// ---------------------------------------------------------------- upward pass
// scatter_point_attrs_to_nodes (cpu_kernels.cpp:173-433): leaves sum their
// points in list order; internal nodes sum their children in octant order;
// centre = sum(W_c * centre_c) * (1/W), or the zero-weight fallback
// sum(centre_c) * (1/num_points) (cuda_kernels.cu:459-463, 513-517).
// Preorder ids put every child after its parent, so one reverse sweep over the
// ids finishes all children before their parent.
void orc_upward(int64_t m, const int64_t* children, const uint8_t* is_leaf,
                const int64_t* num_points, const int64_t* n2p_index, const int64_t* n2p_start,
                const double* pts, const double* weights, const double* attrs, int64_t c_dim,
                double* node_attrs, double* node_rep, double* node_w) {
    std::memset(node_attrs, 0, sizeof(double) * m * c_dim);
    for (int64_t n = m - 1; n >= 0; --n) {
        double rep[3] = {0, 0, 0}, rep0[3] = {0, 0, 0}, tw = 0;
        double* na = node_attrs + n * c_dim;
        if (is_leaf[n]) {
            for (int64_t j = 0; j < num_points[n]; ++j) {
                int64_t p = n2p_index[n2p_start[n] + j];
                tw += weights[p];
                for (int64_t c = 0; c < c_dim; ++c) na[c] += attrs[p * c_dim + c];
                for (int d = 0; d < 3; ++d) rep[d] += pts[3 * p + d] * weights[p];
                for (int d = 0; d < 3; ++d) rep0[d] += pts[3 * p + d];
            }
        } else {
            for (int k = 0; k < kKids; ++k) {
                int64_t ch = children[n * kKids + k];
                if (ch == -1) continue;
                tw += node_w[ch];
                for (int64_t c = 0; c < c_dim; ++c) na[c] += node_attrs[ch * c_dim + c];
                for (int d = 0; d < 3; ++d) rep[d] += node_rep[3 * ch + d] * node_w[ch];
                for (int d = 0; d < 3; ++d) rep0[d] += node_rep[3 * ch + d];
            }
        }
        if (tw > 0) {
            double inv = 1.0 / tw;
            for (int d = 0; d < 3; ++d) node_rep[3 * n + d] = rep[d] * inv;
        } else {
            double inv = 1.0 / (double)num_points[n];
            for (int d = 0; d < 3; ++d) node_rep[3 * n + d] = rep0[d] * inv;
        }
        node_w[n] = tw;
    }
}

// ---------------------------------------------------------------- traversal
// Per-query DFS with an explicit stack (cuda_kernels.cu:575-620): pop; far iff
// dist^2(query, centre) > (open_factor * half_w)^2 (open_factor = 2/theta; the
// reference's 2*TREECODE_THRESHOLD = 4, i.e. theta = 0.5); far -> one
// interaction with the node moment at the centre; near leaf -> every point of
// the leaf in list order; near internal -> push children 0..7 (popped 7..0).
// Counters (optional, may be null): far node interactions, leaf point
// interactions, nodes popped -- per query.
void orc_apply(int64_t op, int64_t nq, const double* q, const double* qw, const double* pts,
               const double* attrs, int64_t m, const int64_t* children, const uint8_t* is_leaf,
               const double* half_w, const int64_t* num_points, const int64_t* n2p_index,
               const int64_t* n2p_start, const double* node_attrs, const double* node_rep,
               double open_factor, int64_t continuous, double an, double* out,
               int64_t* n_far, int64_t* n_near, int64_t* n_visit) {
    const int in_dim = op_in_dim((int)op), out_dim = op_out_dim((int)op);
    const bool cont = continuous != 0;
#pragma omp parallel
    {
        std::vector<int64_t> stack;
        stack.reserve(256);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nq; ++i) {
            const double* x = q + 3 * i;
            double acc[3] = {0, 0, 0};
            int64_t cf = 0, cn = 0, cv = 0;
            stack.clear();
            if (m > 0) stack.push_back(0);
            while (!stack.empty()) {
                int64_t nd = stack.back();
                stack.pop_back();
                ++cv;
                const double* c = node_rep + 3 * nd;
                double thr = open_factor * half_w[nd];
                if (dist2_of(x, c) > thr * thr) {
                    double d[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
                    interact((int)op, acc, d, node_attrs + nd * in_dim, qw[i], cont, an);
                    ++cf;
                } else if (!is_leaf[nd]) {
                    for (int k = 0; k < kKids; ++k) {
                        int64_t ch = children[nd * kKids + k];
                        if (ch != -1) stack.push_back(ch);
                    }
                } else {
                    for (int64_t j = 0; j < num_points[nd]; ++j) {
                        int64_t p = n2p_index[n2p_start[nd] + j];
                        const double* y = pts + 3 * p;
                        double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
                        interact((int)op, acc, d, attrs + p * in_dim, qw[i], cont, an);
                    }
                    cn += num_points[nd];
                }
            }
            if (op == OP_A3) acc[0] = acc[0] / 3.0;
            for (int c = 0; c < out_dim; ++c) out[i * out_dim + c] = acc[c];
            if (n_far) n_far[i] = cf;
            if (n_near) n_near[i] = cn;
            if (n_visit) n_visit[i] = cv;
        }
    }
}

// Brute-force direct sum over every source point (the theta -> 0 limit).
void orc_direct(int64_t op, int64_t nq, const double* q, const double* qw, int64_t n,
                const double* pts, const double* attrs, int64_t continuous, double an,
                double* out) {
    const int in_dim = op_in_dim((int)op), out_dim = op_out_dim((int)op);
    const bool cont = continuous != 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nq; ++i) {
        const double* x = q + 3 * i;
        double acc[3] = {0, 0, 0};
        for (int64_t p = 0; p < n; ++p) {
            const double* y = pts + 3 * p;
            double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
            interact((int)op, acc, d, attrs + p * in_dim, qw[i], cont, an);
        }
        if (op == OP_A3) acc[0] = acc[0] / 3.0;
        for (int c = 0; c < out_dim; ++c) out[i * out_dim + c] = acc[c];
    }
}

// Exact adjoint of the linear treecode operator `op` (fixed node centres
// node_rep, node moments = plain sums of the input): the same traversal as
// orc_apply, but every interaction scatters d(out_i)/d(input) * g_i back to the
// sources -- far-node interactions into a per-node accumulator that a downward
// sweep then adds to every point of the node, near interactions straight to the
// point.  Serial, for small N only.  in_dim/out_dim are those of `op`; g is
// [nq, out_dim]; result is [n, in_dim].  Only ops whose kernel is linear in the
// source attribute are accepted (A, A_CPU, AT, G, A2k, A2Tk, A3).
void orc_adjoint(int64_t op, int64_t nq, const double* q, const double* qw, int64_t n,
                 const double* pts, int64_t m, const int64_t* children, const uint8_t* is_leaf,
                 const double* half_w, const int64_t* num_points, const int64_t* n2p_index,
                 const int64_t* n2p_start, const double* node_rep, double open_factor,
                 int64_t continuous, double an, const double* g, double* result) {
    const int in_dim = op_in_dim((int)op), out_dim = op_out_dim((int)op);
    const bool cont = continuous != 0;
    std::vector<double> node_acc((size_t)(m * in_dim), 0.0);
    std::memset(result, 0, sizeof(double) * n * in_dim);
    // the derivative of out w.r.t. a unit source attribute, per component
    auto jac = [&](const double* d, double w, double* jrow /*[out_dim*in_dim]*/) {
        for (int a = 0; a < in_dim; ++a) {
            double e[3] = {0, 0, 0};
            e[a] = 1.0;
            double o[3] = {0, 0, 0};
            interact((int)op, o, d, e, w, cont, an);
            for (int b = 0; b < out_dim; ++b) jrow[b * in_dim + a] = o[b];
        }
    };
    std::vector<int64_t> stack;
    for (int64_t i = 0; i < nq; ++i) {
        const double* x = q + 3 * i;
        double gi[3] = {0, 0, 0};
        for (int b = 0; b < out_dim; ++b) gi[b] = g[i * out_dim + b] / (op == OP_A3 ? 3.0 : 1.0);
        stack.clear();
        if (m > 0) stack.push_back(0);
        double J[9];
        while (!stack.empty()) {
            int64_t nd = stack.back();
            stack.pop_back();
            const double* c = node_rep + 3 * nd;
            double thr = open_factor * half_w[nd];
            if (dist2_of(x, c) > thr * thr) {
                double d[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
                jac(d, qw[i], J);
                for (int a = 0; a < in_dim; ++a)
                    for (int b = 0; b < out_dim; ++b) node_acc[nd * in_dim + a] += J[b * in_dim + a] * gi[b];
            } else if (!is_leaf[nd]) {
                for (int k = 0; k < kKids; ++k) {
                    int64_t ch = children[nd * kKids + k];
                    if (ch != -1) stack.push_back(ch);
                }
            } else {
                for (int64_t j = 0; j < num_points[nd]; ++j) {
                    int64_t p = n2p_index[n2p_start[nd] + j];
                    const double* y = pts + 3 * p;
                    double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
                    jac(d, qw[i], J);
                    for (int a = 0; a < in_dim; ++a)
                        for (int b = 0; b < out_dim; ++b) result[p * in_dim + a] += J[b * in_dim + a] * gi[b];
                }
            }
        }
    }
    // downward sweep: every point receives the accumulators of all its nodes
    for (int64_t nd = 0; nd < m; ++nd)
        for (int64_t j = 0; j < num_points[nd]; ++j) {
            int64_t p = n2p_index[n2p_start[nd] + j];
            for (int a = 0; a < in_dim; ++a) result[p * in_dim + a] += node_acc[nd * in_dim + a];
        }
}

int64_t orc_num_threads(void) { return omp_get_max_threads(); }

// OpenMP team size of the following parallel loops (the reference's CPU
// kernels hard-code 20, cpu_kernels.cpp:244; timing runs compare both)
void orc_set_num_threads(int64_t n) { omp_set_num_threads((int)n); }

}  // extern "C"
