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
// (find_next + nonleaf per level, cuda_iface:77-131), this runs
// 1 + ceil(depth/3) launches: one over the leaves (points staged through LDS,
// the points' per-call source records written on the way), then one per band
// of up to 3 depths, deepest first, a block per node at the band's top depth
// taking the band's internal nodes level by level.
//
// Several "sets" (K) of scalar attributes can share one pass (the three
// inputs of the fused sum_k A2Tk): each set has its own weights and centres,
// and its own wave in every block of both launches.
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

// y[i] for a run-time i in [0, 3) by selects, so y can stay in registers
__device__ inline double pick3(const double* y, int i) {
    return i == 0 ? y[0] : (i == 1 ? y[1] : y[2]);
}

// float bounds of a double, rounded outwards
__device__ inline float up_float(double x) {
    const float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}
__device__ inline float down_float(double x) {
    const float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}

// The FP32 walks decide far/near on the float centre cf = (float)c of each
// set (NodeV32, NodeS32): they form
// d = (x_f - cf) (+ x_lo, the low part of an FP64 query split into floats)
// and r2 = fma(d0, d0, fma(d1, d1, d2 * d2)).  A query point of the tree
// itself is taken as its float copy x_f, off by at most qr per component
// from the FP64 point (0 when the points are FP32 values).  Per component d
// is then off by at most E = |c - cf| + qr + 2^-23 (|d| + |c - cf| + qr) +
// 2^-46 (|c| + 2 |d|) (two FP32 roundings and the split of the query), so at
// |d| = thr
// |r2 - dist^2| <= B = 2 sqrt(3) thr E + 3 E^2 + 2^-22 thr^2 (r2's own
// roundings) + 2^-50 thr^2 (the FP64 reference's rounding of its dist^2).
// The error grows more slowly than |dist^2 - thr^2| away from the threshold,
// so r2 > hi2 = thr^2 + 2B proves the FP64 test far and r2 < lo2 = thr^2 - 2B
// proves it near; lanes between them retest in FP64.  ce = |c - cf| + qr +
// 2^-46 (|c| + 2) is the part of E that is not FP32 arithmetic, which the
// walks' cutoff test (r >= w) widens its own band by.
__device__ inline void open_band(const double* c, double thr2, double qr, float* cf, float& hi2,
                                 float& lo2, float& ce) {
    double delta = 0.0, cmax = 0.0;
    for (int d = 0; d < 3; ++d) {
        cf[d] = (float)c[d];
        delta = fmax(delta, fabs(c[d] - (double)cf[d]));
        cmax = fmax(cmax, fabs(c[d]));
    }
    const double thr = sqrt(thr2);
    const double E = delta + qr + 0x1p-23 * (thr + delta + qr) + 0x1p-46 * (cmax + 2.0 * thr);
    const double B = 3.4641016151377546 * thr * E + 3.0 * E * E + (0x1p-22 + 0x1p-50) * thr2;
    hi2 = up_float(thr2 + 2.0 * B);
    lo2 = down_float(thr2 - 2.0 * B);
    ce = up_float(delta + qr + 0x1p-46 * (cmax + 2.0));
}

// record writers --------------------------------------------------------------
template <int FAM, int K, int P>
struct Rec;

// vector family: one set of 3-vectors
template <int P>
struct Rec<0, 1, P> {
    static constexpr int C = 3;
    __device__ static void node(void* nodes, int64_t n, int k, const double* c, const double* mm,
                                double an, int /*dir0*/, double thr2, double qr) {
        if constexpr (P == CAGR_PREC_FP64) {
            NodeV64 r;
            for (int d = 0; d < 3; ++d) { r.c[d] = c[d]; r.m[d] = mm[d]; }
            ((NodeV64*)nodes)[n] = r;
        } else {
            NodeV32 r;
            open_band(c, thr2, qr, r.cf, r.hi2, r.lo2, r.ce);
            for (int d = 0; d < 3; ++d) {
                const double e = exp(-0.5 * an * c[d]);
                r.m[d] = (float)mm[d];
                r.e[d] = (float)e;
                r.q[d] = (float)(0.5 * an * e * mm[d]);
            }
            r.pad = 0.f;
            ((NodeV32*)nodes)[n] = r;
        }
    }
    // source record of sorted point s from its loaded point (FP64 and FP32)
    // and attributes
    __device__ static void src(void* src, int64_t s, const double* y64, const float* y32,
                               const double* a, double an, int /*dir0*/) {
        if constexpr (P == CAGR_PREC_FP64) {
            SrcV64 r;
            for (int d = 0; d < 3; ++d) { r.y[d] = y64[d]; r.mu[d] = a[d]; }
            ((SrcV64*)src)[s] = r;
        } else {
            SrcV32 r;
            for (int d = 0; d < 3; ++d) {
                const double mu = a[d];
                const double e = exp(-0.5 * an * y64[d]);
                r.y[d] = y32[d];
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
                                double an, int dir0, double thr2, double qr) {
        if constexpr (P == CAGR_PREC_FP64) {
            NodeS64 r;
            for (int d = 0; d < 3; ++d) r.c[d] = c[d];
            r.s = mm[0];
            ((NodeS64*)nodes)[n * K + k] = r;
        } else {
            NodeS32 r;
            open_band(c, thr2, qr, r.cf, r.hi2, r.lo2, r.ce);
            r.se = (float)(mm[0] * exp(-0.5 * an * pick3(c, dir0 + k)));
            r.pad = 0.f;
            ((NodeS32*)nodes)[n * K + k] = r;
        }
    }
    __device__ static void src(void* src, int64_t s, const double* y64, const float* y32,
                               const double* a, double an, int dir0) {
        if constexpr (P == CAGR_PREC_FP64) {
            SrcS64<K> r;
            for (int d = 0; d < 3; ++d) r.y[d] = y64[d];
            for (int k = 0; k < K; ++k) r.s[k] = a[k];
            ((SrcS64<K>*)src)[s] = r;
        } else {
            SrcS32<K> r;
            for (int d = 0; d < 3; ++d) r.y[d] = y32[d];
            for (int k = 0; k < K; ++k) r.a[k] = (float)(a[k] * exp(-0.5 * an * pick3(y64, dir0 + k)));
            ((SrcS32<K>*)src)[s] = r;
        }
    }
};

// one sorted point's loaded inputs
template <int KC, int P>
struct PtIn {
    double y[3], a[KC], w;
    float y32[P == CAGR_PREC_FP64 ? 1 : 3];
};

template <typename TI, int KC, int P>
__device__ inline void load_point(PtIn<KC, P>& q, int32_t s, int64_t j,
                                  const double* __restrict__ pts64,
                                  const float* __restrict__ pts32, const TI* __restrict__ attrs,
                                  const void* __restrict__ weights, int wdtype) {
    for (int d = 0; d < 3; ++d) q.y[d] = pts64[3 * (int64_t)s + d];
    if constexpr (P != CAGR_PREC_FP64)
        for (int d = 0; d < 3; ++d) q.y32[d] = pts32[3 * (int64_t)s + d];
    for (int c = 0; c < KC; ++c) q.a[c] = ld64(attrs, j * KC + c);
    q.w = weights ? wload(weights, wdtype, j) : 0.0;
}

// a leaf's moments of set k from its sums over its points: centre =
// weighted mean, or the plain mean when the total weight is 0
template <int FAM, int K, int P>
__device__ inline void leaf_finish(int32_t nd, const TopoRec& t, int k, double tw,
                                   const double (&rep)[3],
                                   const double (&acc)[Rec<FAM, K, P>::C],
                                   const double (&rep0)[3], int64_t m, double* __restrict__ w64,
                                   double* __restrict__ r64, double* __restrict__ m64,
                                   void* __restrict__ nodes, double an, int dir0,
                                   double qr) {
#pragma clang fp contract(off)
    constexpr int C = Rec<FAM, K, P>::C;
    const double np = (double)(t.end - t.begin);
    double c[3];
    if (tw > 0) {
        const double inv = 1.0 / tw;
        for (int d = 0; d < 3; ++d) c[d] = rep[d] * inv;
    } else {
        const double inv = 1.0 / np;
        for (int d = 0; d < 3; ++d) c[d] = rep0[d] * inv;
    }
    const int64_t o = (int64_t)k * m + nd;
    w64[o] = tw;
    for (int d = 0; d < 3; ++d) r64[3 * o + d] = c[d];
    for (int cc = 0; cc < C; ++cc) m64[C * o + cc] = acc[cc];
    Rec<FAM, K, P>::node(nodes, nd, k, c, acc, an, dir0, t.thr2, qr);
}

// child ids of internal node t in octant order (ascending id): positions
// kids.. of the depth order; entries past the child count repeat the last
// child (in bounds, never summed)
__device__ inline void load_kids(const TopoRec& t, const int32_t* __restrict__ order, int32_t ch[8]) {
    const int nk = topo_nkids(t);
#pragma unroll
    for (int j = 0; j < 8; ++j) ch[j] = order[t.kids + (j < nk ? j : nk - 1)];
}

// set k of internal node nd from its children ch[0 .. nk): the child moments
// are loaded and summed in order before any store
template <int FAM, int K, int P>
__device__ inline void up_internal_node(int32_t nd, const TopoRec& t, const int32_t ch[8], int k,
                                        int64_t m, double* __restrict__ w64,
                                        double* __restrict__ r64, double* __restrict__ m64,
                                        void* __restrict__ nodes, double an, int dir0,
                                        double qr) {
#pragma clang fp contract(off)
    constexpr int C = Rec<FAM, K, P>::C;
    const int nk = topo_nkids(t);
    double tw = 0.0, rep[3] = {0.0, 0.0, 0.0}, rep0[3] = {0.0, 0.0, 0.0}, acc[C];
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t o = (int64_t)k * m + ch[j];
        const double wc = w64[o];
        double mc[C], cd[3];
        for (int c = 0; c < C; ++c) mc[c] = m64[C * o + c];
        for (int d = 0; d < 3; ++d) cd[d] = r64[3 * o + d];
        if (j < nk) {
            tw += wc;
            for (int c = 0; c < C; ++c) acc[c] += mc[c];
            for (int d = 0; d < 3; ++d) {
                rep[d] += cd[d] * wc;
                rep0[d] += cd[d];
            }
        }
    }
    const double np = (double)(t.end - t.begin);
    double c[3];
    if (tw > 0) {
        const double inv = 1.0 / tw;
        for (int d = 0; d < 3; ++d) c[d] = rep[d] * inv;
    } else {
        const double inv = 1.0 / np;
        for (int d = 0; d < 3; ++d) c[d] = rep0[d] * inv;
    }
    const int64_t o = (int64_t)k * m + nd;
    w64[o] = tw;
    for (int d = 0; d < 3; ++d) r64[3 * o + d] = c[d];
    for (int cc = 0; cc < C; ++cc) m64[C * o + cc] = acc[cc];
    Rec<FAM, K, P>::node(nodes, nd, k, c, acc, an, dir0, t.thr2, qr);
}

// Leaf launch: a block of K waves per kLeafIds consecutive node ids
// (preorder), wave k summing set k.  The leaves among the ids cover one run
// of sorted points with nothing between them (leaves in preorder hold
// consecutive point ranges), so the run is staged through LDS T points at a
// time with coalesced loads by the whole block (the attributes are gathered
// by the permutation), the points' source records written on the way; then
// each leaf's lane in wave k sums set k over its points from LDS in order.
// Every point is staged once.  The sums are serial per leaf and set, so a
// wave per set shortens each block's sum phase K-fold.  (16 ids per block
// instead of 64 measured the same: the launch is held by its HBM traffic,
// about twice the compulsory bytes, the attributes' gather by the
// permutation fetching a line per point.)
constexpr int kLeafIds = 64;

template <typename TI, int FAM, int K, int P>
__global__ __launch_bounds__(64 * K) void k_up_leaves(
    int64_t m, const TopoRec* __restrict__ topo, const int32_t* __restrict__ perm,
    const double* __restrict__ pts64, const float* __restrict__ pts32,
    const TI* __restrict__ attrs, const void* __restrict__ weights, int wdtype,
    double* __restrict__ w64, double* __restrict__ r64, double* __restrict__ m64,
    void* __restrict__ nodes, void* __restrict__ src, double an, int dir0, double qr) {
#pragma clang fp contract(off)
    constexpr int C = Rec<FAM, K, P>::C;
    constexpr int KC = K * C;
    constexpr int F = 3 + KC + K;  // staged per point: y, attributes, a weight per set
    constexpr int T = 256;         // points per staged tile
    constexpr int NT = 64 * K;     // threads of the block
    constexpr int U = (T + NT - 1) / NT;
    __shared__ double s_pt[F * T];
    const int32_t tid = (int32_t)threadIdx.x;
    const int32_t lane = tid & 63;
    const int k = __builtin_amdgcn_readfirstlane(tid >> 6);   // this wave's set
    const int64_t nd = (int64_t)blockIdx.x * kLeafIds + lane;
    TopoRec t{};
    bool leaf = false;
    if (lane < kLeafIds && nd < m) {
        t = topo[nd];
        leaf = topo_leaf(t);
    }
    const uint64_t lm = __ballot(leaf);   // the same in every wave of the block
    if (lm == 0) return;
    const int first = __ffsll((unsigned long long)lm) - 1;
    const int last = 63 - __clzll((unsigned long long)lm);
    const int32_t lo = __shfl(t.begin, first), hi = __shfl(t.end, last);
    double tw = 0.0, rep[3] = {0.0, 0.0, 0.0}, acc[C], rep0[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    for (int32_t tl = lo; tl < hi; tl += T) {
        const int32_t th = min(tl + T, hi);
        PtIn<KC, P> pt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t i = tid + NT * u;
            if (i < T) {
                const int32_t s = min(tl + i, th - 1);
                load_point<TI, KC, P>(pt[u], s, (int64_t)perm[s], pts64, pts32, attrs, weights,
                                      wdtype);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t i = tid + NT * u;
            const int32_t s = tl + i;
            if (i < T && s < th) {
                Rec<FAM, K, P>::src(src, s, pt[u].y, pt[u].y32, pt[u].a, an, dir0);
                for (int d = 0; d < 3; ++d) s_pt[d * T + i] = pt[u].y[d];
                for (int c = 0; c < KC; ++c) s_pt[(3 + c) * T + i] = pt[u].a[c];
                for (int kk = 0; kk < K; ++kk) {
                    double w;
                    if (weights) {
                        w = pt[u].w;
                    } else {
                        const double* a = pt[u].a + kk * C;
                        double ss = a[0] * a[0];
                        for (int c = 1; c < C; ++c) ss = ss + a[c] * a[c];
                        w = sqrt(ss);
                    }
                    s_pt[(3 + KC + kk) * T + i] = w;
                }
            }
        }
        __syncthreads();
        if (leaf) {
            const int32_t a1 = min(t.end, th);
            for (int32_t s = max(t.begin, tl); s < a1; ++s) {
                const int i = s - tl;
                const double y[3] = {s_pt[i], s_pt[T + i], s_pt[2 * T + i]};
                const double w = s_pt[(3 + KC + k) * T + i];
                tw += w;
                for (int c = 0; c < C; ++c) acc[c] += s_pt[(3 + k * C + c) * T + i];
                for (int d = 0; d < 3; ++d) rep[d] += y[d] * w;
                for (int d = 0; d < 3; ++d) rep0[d] += y[d];
            }
        }
        __syncthreads();
    }
    if (leaf)
        leaf_finish<FAM, K, P>((int32_t)nd, t, k, tw, rep, acc, rep0, m, w64, r64, m64, nodes, an,
                               dir0, qr);
}

// Band launch, a block of K waves per node at the band's top depth (wave k
// for set k): the internal nodes of the band's levels nlev-1..0 (nlev <= 3)
// under it, with a barrier between levels; the leaves were done by the leaf
// launch.  Level l holds at most 8^l <= 64 of them, so each lane has at most
// one node per level, and the node ids and child ids of the next level (which
// no launch here writes) are loaded while the current level runs.
template <int FAM, int K, int P>
__global__ __launch_bounds__(64 * K) void k_up_band(const int2* __restrict__ seg, int nlev,
                                                    const int32_t* __restrict__ order,
                                                    const TopoRec* __restrict__ topo, int64_t m,
                                                    double* __restrict__ w64,
                                                    double* __restrict__ r64,
                                                    double* __restrict__ m64,
                                                    void* __restrict__ nodes, double an, int dir0,
                                                    double qr) {
    const int32_t lane = (int32_t)(threadIdx.x & 63);
    const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));   // this wave's set
    auto fetch = [&](int l, int32_t& nd, TopoRec& t, int32_t ch[8]) {
        const int2 sg = seg[(int64_t)blockIdx.x * nlev + l];
        nd = -1;
        if (sg.x + lane < sg.y) {
            const int32_t v = order[sg.x + lane];
            t = topo[v];
            if (!topo_leaf(t)) {
                nd = v;
                load_kids(t, order, ch);
            }
        }
    };
    int32_t nd, ch[8];
    TopoRec t;
    fetch(nlev - 1, nd, t, ch);
    for (int l = nlev - 1; l >= 0; --l) {
        int32_t nd2 = -1, ch2[8];
        TopoRec t2;
        if (l > 0) fetch(l - 1, nd2, t2, ch2);
        if (nd >= 0)
            up_internal_node<FAM, K, P>(nd, t, ch, k, m, w64, r64, m64, nodes, an, dir0, qr);
        __syncthreads();
        nd = nd2;
        t = t2;
        for (int j = 0; j < 8; ++j) ch[j] = ch2[j];
    }
}

template <typename TI, int FAM, int K, int P>
int run_upward(cagr_tree* t, const TI* attrs, const void* weights, int wdtype, double an, int dir0,
               hipStream_t st) {
    const int64_t m = t->m;
    if (m == 0) return CAGR_OK;
    bool ok = true;
    for (const int4& b : t->up_bands) ok = ok && b.y >= 1 && b.y <= 3;  // <= 64 nodes a level
    CAGR_REQUIRE(ok, CAGR_ERR_STATE, "upward: bad band");
    k_up_leaves<TI, FAM, K, P><<<(unsigned)((m + kLeafIds - 1) / kLeafIds), 64 * K, 0, st>>>(
        m, t->topo, t->perm, t->pts64, t->pts32, attrs, weights, wdtype, t->w64, t->r64, t->m64,
        t->nodes, t->src, an, dir0, t->qround);
    CAGR_LAUNCHED();
    for (const int4& b : t->up_bands) {
        k_up_band<FAM, K, P><<<b.z, 64 * K, 0, st>>>(t->up_seg + b.w, b.y, t->order, t->topo, m, t->w64,
                                                 t->r64, t->m64, t->nodes, an, dir0, t->qround);
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

// node records of one set from caller-given moments (preorder node ids):
// node_attrs [m, C] and node_reppoints [m, 3] in TI
// (r64, set 0, gets the same centres: the FP32 walks' FP64 retests read it)
template <typename TI, int FAM, int P>
__global__ __launch_bounds__(kThreads) void k_nodes_from_moments(int64_t m,
                                                                const TI* __restrict__ attrs,
                                                                const TI* __restrict__ rep,
                                                                const TopoRec* __restrict__ topo,
                                                                void* nodes, double* r64,
                                                                double an, int dir0, double qr) {
    using R = Rec<FAM, 1, P>;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m) return;
    double c[3], mm[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
        c[d] = (double)rep[n * 3 + d];
        r64[3 * n + d] = c[d];
    }
    for (int d = 0; d < R::C; ++d) mm[d] = (double)attrs[n * R::C + d];
    R::node(nodes, n, 0, c, mm, an, dir0, topo[n].thr2, qr);
}

template <typename TI>
int nodes_from_moments_t(cagr_tree* t, int family, int precision, int dir0, const TI* attrs,
                         const TI* rep, double an, hipStream_t st) {
    const int64_t m = t->m;
    if (m == 0) return CAGR_OK;
    const unsigned nb = (unsigned)blocks_for(m);
    const bool p64 = precision == CAGR_PREC_FP64;
    if (family == 0) {
        if (p64)
            k_nodes_from_moments<TI, 0, CAGR_PREC_FP64><<<nb, kThreads, 0, st>>>(
                m, attrs, rep, t->topo, t->nodes, t->r64, an, dir0, t->qround);
        else
            k_nodes_from_moments<TI, 0, CAGR_PREC_FP32C><<<nb, kThreads, 0, st>>>(
                m, attrs, rep, t->topo, t->nodes, t->r64, an, dir0, t->qround);
    } else {
        if (p64)
            k_nodes_from_moments<TI, 1, CAGR_PREC_FP64><<<nb, kThreads, 0, st>>>(
                m, attrs, rep, t->topo, t->nodes, t->r64, an, dir0, t->qround);
        else
            k_nodes_from_moments<TI, 1, CAGR_PREC_FP32C><<<nb, kThreads, 0, st>>>(
                m, attrs, rep, t->topo, t->nodes, t->r64, an, dir0, t->qround);
    }
    CAGR_LAUNCHED();
    return CAGR_OK;
}

}  // namespace

// node records of a one-set operator from caller-given moments in `dtype`
// (replaces the records of the last upward pass; the source records stay)
int nodes_from_moments(cagr_tree* t, int family, int precision, int dir0, const void* node_attrs,
                       const void* node_reppoints, int dtype, double an, hipStream_t st) {
    if (dtype == CAGR_F32)
        return nodes_from_moments_t<float>(t, family, precision, dir0, (const float*)node_attrs,
                                           (const float*)node_reppoints, an, st);
    return nodes_from_moments_t<double>(t, family, precision, dir0, (const double*)node_attrs,
                                        (const double*)node_reppoints, an, st);
}

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
