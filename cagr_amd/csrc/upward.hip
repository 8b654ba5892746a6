// Upward pass ("dipole moments") and per-call source records.
//
// Restates scatter_point_attrs_to_nodes
// (CUDA-based/ext/wn_treecode/wn_treecode_cuda/wn_treecode_cuda_kernels.cu:419-553,
//  wn_treecode_cpu/wn_treecode_cpu_kernels.cpp:173-433):
//   leaf:     W = sum w_j, M = sum a_j, centre = (sum w_j y_j) * (1/W), points in
//             ascending original index (the sorted order inside a leaf);
//   internal: W = sum W_c, M = sum M_c, centre = (sum W_c centre_c) * (1/W),
//             children in octant order;
//   W == 0:   centre = (sum of the point / child centres) * (1/num_points).
// Point weights are ||a_j||_2 (wn_treecode_func.py:131) unless the caller
// passes fixed weights.  All FP64 moment arithmetic runs with contraction off
// in that order, so the FP64 moments equal the reference CPU core's bit for
// bit.  Instead of the reference's 1 + 2*tree_depth launches over all nodes
// (find_next + nonleaf per level, cuda_iface:77-131), this runs one launch for
// all leaves and one per internal depth, each over only that depth's nodes.
//
// Several "sets" (K) of scalar attributes can share one pass (the three
// inputs of the fused sum_k A2Tk): each set has its own weights and centres.
#include "common.h"

namespace cagr {
namespace {

constexpr int kThreads = 256;
inline int64_t blocks_for(int64_t n) { return (n + kThreads - 1) / kThreads; }

// attrs may be null (centres only, from fixed weights): read as zero
template <typename TI>
__device__ inline double ld64(const TI* p, int64_t i) { return p ? (double)p[i] : 0.0; }

// point weights in their own dtype (they may differ from the attributes')
__device__ inline double wload(const void* w, int wdtype, int64_t i) {
    return wdtype == CAGR_F32 ? (double)((const float*)w)[i] : ((const double*)w)[i];
}

// record writers --------------------------------------------------------------
template <int FAM, int K, int P>
struct Rec;

// vector family: one set of 3-vectors
template <int P>
struct Rec<0, 1, P> {
    static constexpr int C = 3;
    __device__ static void node(void* nodes, int64_t n, int k, const double* c, const double* mm,
                                double an, int /*dir0*/) {
        if constexpr (P == CAGR_PREC_FP64) {
            NodeV64 r;
            for (int d = 0; d < 3; ++d) { r.c[d] = c[d]; r.m[d] = mm[d]; }
            ((NodeV64*)nodes)[n] = r;
        } else {
            NodeV32 r;
            for (int d = 0; d < 3; ++d) {
                const double e = exp(-0.5 * an * c[d]);
                r.c[d] = c[d];
                r.m[d] = (float)mm[d];
                r.e[d] = (float)e;
                r.q[d] = (float)(0.5 * an * e * mm[d]);
            }
            r.pad = 0.f;
            ((NodeV32*)nodes)[n] = r;
        }
    }
    template <typename TI>
    __device__ static void src(void* src, int64_t s, int64_t j, const double* pts64,
                               const float* pts32, const TI* attrs, double an, int /*dir0*/) {
        if constexpr (P == CAGR_PREC_FP64) {
            SrcV64 r;
            for (int d = 0; d < 3; ++d) { r.y[d] = pts64[3 * s + d]; r.mu[d] = ld64(attrs, 3 * j + d); }
            ((SrcV64*)src)[s] = r;
        } else {
            SrcV32 r;
            for (int d = 0; d < 3; ++d) {
                const double mu = ld64(attrs, 3 * j + d);
                const double e = exp(-0.5 * an * pts64[3 * s + d]);
                r.y[d] = pts32[3 * s + d];
                r.mu[d] = (float)mu;
                r.e[d] = (float)e;
                r.q[d] = (float)(0.5 * an * e * mu);
            }
            ((SrcV32*)src)[s] = r;
        }
    }
};

// scalar family: K sets of scalars; set k convects along axis dir0 + k
template <int K, int P>
struct Rec<1, K, P> {
    static constexpr int C = 1;
    __device__ static void node(void* nodes, int64_t n, int k, const double* c, const double* mm,
                                double an, int dir0) {
        if constexpr (P == CAGR_PREC_FP64) {
            NodeS64 r;
            for (int d = 0; d < 3; ++d) r.c[d] = c[d];
            r.s = mm[0];
            ((NodeS64*)nodes)[n * K + k] = r;
        } else {
            NodeS32 r;
            for (int d = 0; d < 3; ++d) r.c[d] = c[d];
            r.s = (float)mm[0];
            r.e = (float)exp(-0.5 * an * c[dir0 + k]);
            ((NodeS32*)nodes)[n * K + k] = r;
        }
    }
    template <typename TI>
    __device__ static void src(void* src, int64_t s, int64_t j, const double* pts64,
                               const float* pts32, const TI* attrs, double an, int dir0) {
        if constexpr (P == CAGR_PREC_FP64) {
            SrcS64<K> r;
            for (int d = 0; d < 3; ++d) r.y[d] = pts64[3 * s + d];
            for (int k = 0; k < K; ++k) r.s[k] = ld64(attrs, j * K + k);
            ((SrcS64<K>*)src)[s] = r;
        } else {
            SrcS32<K> r;
            for (int d = 0; d < 3; ++d) r.y[d] = pts32[3 * s + d];
            for (int k = 0; k < K; ++k)
                r.a[k] = (float)(ld64(attrs, j * K + k) * exp(-0.5 * an * pts64[3 * s + dir0 + k]));
            ((SrcS32<K>*)src)[s] = r;
        }
    }
};

template <typename TI, int FAM, int K, int P>
__global__ __launch_bounds__(kThreads) void k_src(int64_t n, const int32_t* __restrict__ perm,
                                                  const double* __restrict__ pts64,
                                                  const float* __restrict__ pts32,
                                                  const TI* __restrict__ attrs,
                                                  void* __restrict__ src, double an, int dir0) {
    const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    Rec<FAM, K, P>::src(src, s, perm[s], pts64, pts32, attrs, an, dir0);
}

template <typename TI, int FAM, int K, int P>
__global__ __launch_bounds__(kThreads) void k_up_leaf(
    const int32_t* __restrict__ order, int64_t cnt, const TopoRec* __restrict__ topo,
    const int32_t* __restrict__ perm, const double* __restrict__ pts64,
    const TI* __restrict__ attrs, const void* __restrict__ weights, int wdtype, int64_t m,
    double* __restrict__ w64, double* __restrict__ r64, double* __restrict__ m64,
    void* __restrict__ nodes, double an, int dir0) {
#pragma clang fp contract(off)
    constexpr int C = Rec<FAM, K, P>::C;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const int32_t n = order[i];
    const TopoRec t = topo[n];
    double tw[K], rep[K][3], acc[K][C], rep0[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < K; ++k) {
        tw[k] = 0.0;
        for (int d = 0; d < 3; ++d) rep[k][d] = 0.0;
        for (int c = 0; c < C; ++c) acc[k][c] = 0.0;
    }
    for (int32_t s = t.begin; s < t.end; ++s) {
        const int64_t j = perm[s];
        const double y[3] = {pts64[3 * s], pts64[3 * s + 1], pts64[3 * s + 2]};
        for (int k = 0; k < K; ++k) {
            double a[C];
            for (int c = 0; c < C; ++c) a[c] = ld64(attrs, j * (K * C) + k * C + c);
            double w;
            if (weights) {
                w = wload(weights, wdtype, j);
            } else {
                double ss = a[0] * a[0];
                for (int c = 1; c < C; ++c) ss = ss + a[c] * a[c];
                w = sqrt(ss);
            }
            tw[k] += w;
            for (int c = 0; c < C; ++c) acc[k][c] += a[c];
            for (int d = 0; d < 3; ++d) rep[k][d] += y[d] * w;
        }
        for (int d = 0; d < 3; ++d) rep0[d] += y[d];
    }
    const double np = (double)(t.end - t.begin);
    for (int k = 0; k < K; ++k) {
        double c[3];
        if (tw[k] > 0) {
            const double inv = 1.0 / tw[k];
            for (int d = 0; d < 3; ++d) c[d] = rep[k][d] * inv;
        } else {
            const double inv = 1.0 / np;
            for (int d = 0; d < 3; ++d) c[d] = rep0[d] * inv;
        }
        const int64_t o = (int64_t)k * m + n;
        w64[o] = tw[k];
        for (int d = 0; d < 3; ++d) r64[3 * o + d] = c[d];
        for (int cc = 0; cc < C; ++cc) m64[C * o + cc] = acc[k][cc];
        Rec<FAM, K, P>::node(nodes, n, k, c, acc[k], an, dir0);
    }
}

template <int FAM, int K, int P>
__global__ __launch_bounds__(kThreads) void k_up_internal(
    const int32_t* __restrict__ order, int64_t cnt, const TopoRec* __restrict__ topo, int64_t m,
    double* __restrict__ w64, double* __restrict__ r64, double* __restrict__ m64,
    void* __restrict__ nodes, double an, int dir0) {
#pragma clang fp contract(off)
    constexpr int C = Rec<FAM, K, P>::C;
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt) return;
    const int32_t n = order[i];
    const TopoRec t = topo[n];
    double tw[K], rep[K][3], rep0[K][3], acc[K][C];
    for (int k = 0; k < K; ++k) {
        tw[k] = 0.0;
        for (int d = 0; d < 3; ++d) { rep[k][d] = 0.0; rep0[k][d] = 0.0; }
        for (int c = 0; c < C; ++c) acc[k][c] = 0.0;
    }
    // children of a preorder node: n+1, then each child's escape, up to n's escape
    for (int32_t ch = n + 1; ch < t.escape; ch = topo[ch].escape) {
        for (int k = 0; k < K; ++k) {
            const int64_t o = (int64_t)k * m + ch;
            const double wc = w64[o];
            tw[k] += wc;
            for (int c = 0; c < C; ++c) acc[k][c] += m64[C * o + c];
            for (int d = 0; d < 3; ++d) {
                const double cd = r64[3 * o + d];
                rep[k][d] += cd * wc;
                rep0[k][d] += cd;
            }
        }
    }
    const double np = (double)(t.end - t.begin);
    for (int k = 0; k < K; ++k) {
        double c[3];
        if (tw[k] > 0) {
            const double inv = 1.0 / tw[k];
            for (int d = 0; d < 3; ++d) c[d] = rep[k][d] * inv;
        } else {
            const double inv = 1.0 / np;
            for (int d = 0; d < 3; ++d) c[d] = rep0[k][d] * inv;
        }
        const int64_t o = (int64_t)k * m + n;
        w64[o] = tw[k];
        for (int d = 0; d < 3; ++d) r64[3 * o + d] = c[d];
        for (int cc = 0; cc < C; ++cc) m64[C * o + cc] = acc[k][cc];
        Rec<FAM, K, P>::node(nodes, n, k, c, acc[k], an, dir0);
    }
}

template <typename TI, int FAM, int K, int P>
int run_upward(cagr_tree* t, const TI* attrs, const void* weights, int wdtype, double an, int dir0,
               hipStream_t st) {
    const int64_t n = t->n, m = t->m;
    if (m == 0) return CAGR_OK;
    k_src<TI, FAM, K, P><<<blocks_for(n), kThreads, 0, st>>>(n, t->perm, t->pts64, t->pts32,
                                                              attrs, t->src, an, dir0);
    CAGR_LAUNCHED();
    const auto& g = t->group_off;
    const int64_t nleaf = g[1] - g[0];
    k_up_leaf<TI, FAM, K, P><<<blocks_for(nleaf), kThreads, 0, st>>>(
        t->order + g[0], nleaf, t->topo, t->perm, t->pts64, attrs, weights, wdtype, m, t->w64, t->r64,
        t->m64, t->nodes, an, dir0);
    CAGR_LAUNCHED();
    for (size_t gi = 1; gi + 1 < g.size(); ++gi) {
        const int64_t c = g[gi + 1] - g[gi];
        if (c == 0) continue;
        k_up_internal<FAM, K, P><<<blocks_for(c), kThreads, 0, st>>>(
            t->order + g[gi], c, t->topo, m, t->w64, t->r64, t->m64, t->nodes, an, dir0);
        CAGR_LAUNCHED();
    }
    return CAGR_OK;
}

template <typename TI>
int dispatch_upward(cagr_tree* t, int family, int precision, int K, int dir0, const TI* attrs,
                    const void* weights, int wdtype, double an, hipStream_t st) {
    const bool p64 = precision == CAGR_PREC_FP64;
    if (family == 0) {
        return p64 ? run_upward<TI, 0, 1, CAGR_PREC_FP64>(t, attrs, weights, wdtype, an, dir0, st)
                   : run_upward<TI, 0, 1, CAGR_PREC_FP32C>(t, attrs, weights, wdtype, an, dir0, st);
    }
    if (K == 1)
        return p64 ? run_upward<TI, 1, 1, CAGR_PREC_FP64>(t, attrs, weights, wdtype, an, dir0, st)
                   : run_upward<TI, 1, 1, CAGR_PREC_FP32C>(t, attrs, weights, wdtype, an, dir0, st);
    return p64 ? run_upward<TI, 1, 3, CAGR_PREC_FP64>(t, attrs, weights, wdtype, an, dir0, st)
               : run_upward<TI, 1, 3, CAGR_PREC_FP32C>(t, attrs, weights, wdtype, an, dir0, st);
}

}  // namespace

// family 0 = vector (attrs [n,3]), 1 = scalar sets (attrs [n,K]); weights in
// wdtype (-1: the attributes' dtype)
int upward(cagr_tree* t, int family, int precision, int K, int dir0, const void* attrs,
           int dtype, const void* weights, double an, hipStream_t st, int wdtype) {
    CAGR_REQUIRE(family == 0 ? K == 1 : (K == 1 || K == 3), CAGR_ERR_ARG, "bad upward set count");
    if (wdtype < 0) wdtype = dtype;
    if (dtype == CAGR_F32)
        return dispatch_upward<float>(t, family, precision, K, dir0, (const float*)attrs, weights,
                                      wdtype, an, st);
    return dispatch_upward<double>(t, family, precision, K, dir0, (const double*)attrs, weights,
                                   wdtype, an, st);
}

}  // namespace cagr
