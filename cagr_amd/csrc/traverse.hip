// Treecode traversal kernels (gfx950).
//
// Reference semantics (CUDA-based/ext/wn_treecode/wn_treecode_cuda/wn_treecode_cuda_kernels.cu:556-1257):
// each query walks the tree depth-first from the root; a node whose centre is
// far (dist^2 > (open_factor*half_w)^2) contributes once through its moment,
// a near leaf contributes every one of its points, a near internal node is
// opened.  Kernel math: cuda_kernels.cu:79-414 (A, AT, A2_k, A2T_k, G) and
// :694-763 (A3).
//
// Execution model here.  One wavefront serves 64 queries that are
// consecutive in Morton order (the tree's own sorted points, a 4x4x4 block of
// a grid, or Morton-sorted external queries).  The wave walks the preorder
// node array with a single wave-uniform cursor n: every node record it reads
// is a wave-uniform (scalar) load.  Each lane keeps `resume`, the first node
// id it still has to consider: a lane is active at n iff resume <= n.  An
// active lane that finds n far takes the far interaction and skips n's
// subtree; at a near leaf it takes the leaf's points and also skips past it;
// at a near internal node it stays active for the children.  The wave moves
// to n+1 if any active lane opened n, else to escape(n).  Only the move to
// n+1 needs the far lanes' resume = escape(n): a move to escape(n) leaves
// every lane active at n active there anyway, so leaves and subtrees no lane
// enters update nothing.  Every lane therefore visits exactly the nodes of
// its own reference DFS and takes exactly the same far/near decisions; only
// the order of the summands differs.
//
// Far/near tests run in FP64 in both precisions (the decision must not
// depend on the precision of the kernel sums); in the FP32 form a cutoff
// test r >= w that FP32 cannot resolve (|r^2 - w^2| <= 1e-5 w^2) is redone in
// FP64.  FP32 sums are two-level: plain FP32 partial sums per leaf and per
// run of far interactions, added into an FP64 running total at every leaf
// (one conversion and one FP64 add per partial, where an FP32 Neumaier pair
// costs seven instructions per fold and keeps its old sum live past the new
// one, which costs register copies on the node steps).
//
// The convection factor exp(An*(x-y)_k/2) factors into exp(An*x_k/2) (per
// query) times exp(-An*y_k/2) (per source, precomputed in the source/node
// records), so the FP32 form needs no exp per interaction.  The FP64 form
// evaluates exp per interaction like the reference.
#include "common.h"
#include "trav_common.h"
#include <algorithm>
#include <climits>
#include <cmath>

namespace cagr {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;

enum { QM_SELF = 0, QM_EXPLICIT = 1, QM_GRID = 2 };
enum { VOP_A = 0, VOP_G = 1, VOP_A2 = 2, VOP_A3 = 3 };

struct TP {
    int64_t nq;               // lanes (grid: blocks * 64)
    int32_t root, stop;
    const int32_t* perm;
    const void* q;
    int qdtype;
    const int32_t* qorder;
    double glo[3], gstep[3];
    int64_t gdim[3], gblk[3];
    const void* widths;
    int wdtype;
    double width;
    void* out;
    int odtype;
    double an;
    double sign;              // A: +1 CUDA convention, -1 CPU convention
    int kdir;                 // V: output column for A2k (-1 = all three); S: direction of set 0
    int64_t* cnt_far;
    int64_t* cnt_near;
    int64_t* cnt_visit;
    double open_factor;       // 2 / theta
    const float4* lbound;     // leaf bounding spheres
    dev::CutBand cb;          // band of the FP32 near-pair cutoff test
    const int32_t* tile_order;// regular grid: dispatch slot -> wave (null: identity)
    const double* r64;        // FP64 centres of the upward pass, [set][node][3] (retests)
    int64_t nm;               // nodes (set stride of r64)
    // walk pieces (self queries of one tree; null: one wave per 64 queries
    // on the whole tree): {wave, first node, end node, partial slot or -1}
    const int4* pieces;
    double* planes;           // partial slots: [slot][64 lanes][3] FP64
};

// Where a walk's results go: the output array (slot < 0), or (a piece of a
// split walk) its partial slot in p.planes, 3 FP64 values per lane, which
// k_piece_combine adds up
struct Sink {
    int32_t slot;
    int lane;
};
__device__ inline void emit(const TP& p, const Sink& sk, int64_t oi, int no, int k, double v);

using namespace dev;

__device__ inline void store_out(const TP& p, int64_t i, double v) {
    if (p.odtype == CAGR_F32) ((float*)p.out)[i] = (float)v;
    else ((double*)p.out)[i] = v;
}

// value k of the no values of query oi
__device__ inline void emit(const TP& p, const Sink& sk, int64_t oi, int no, int k, double v) {
    if (sk.slot >= 0) p.planes[((int64_t)sk.slot * 64 + sk.lane) * 3 + k] = v;
    else store_out(p, no * oi + k, v);
}

template <int QM>
__device__ inline bool load_query(const TP& p, const double* __restrict__ pts64, int64_t idx,
                                  int64_t qend, double x[3], double& w, int64_t& oi) {
    bool valid;
    if constexpr (QM == QM_SELF) {
        valid = idx < qend;
        const int64_t s = valid ? idx : 0;
        x[0] = pts64[3 * s]; x[1] = pts64[3 * s + 1]; x[2] = pts64[3 * s + 2];
        oi = p.perm[s];
        w = p.widths ? load_as_double(p.widths, p.wdtype, oi) : p.width;
    } else if constexpr (QM == QM_EXPLICIT) {
        valid = idx < qend;
        const int64_t s = valid ? idx : 0;
        const int64_t i = p.qorder ? (int64_t)p.qorder[s] : s;
        for (int d = 0; d < 3; ++d) x[d] = load_as_double(p.q, p.qdtype, 3 * i + d);
        oi = i;
        w = p.widths ? load_as_double(p.widths, p.wdtype, i) : p.width;
    } else {
        const int64_t wave = idx >> 6;
        const int lane = (int)(idx & 63);
        const int64_t bz = wave % p.gblk[2];
        const int64_t by = (wave / p.gblk[2]) % p.gblk[1];
        const int64_t bx = wave / (p.gblk[2] * p.gblk[1]);
        const int64_t ix = bx * 4 + (lane >> 4), iy = by * 4 + ((lane >> 2) & 3),
                      iz = bz * 4 + (lane & 3);
        valid = bx < p.gblk[0] && ix < p.gdim[0] && iy < p.gdim[1] && iz < p.gdim[2];
        x[0] = p.glo[0] + (double)ix * p.gstep[0];
        x[1] = p.glo[1] + (double)iy * p.gstep[1];
        x[2] = p.glo[2] + (double)iz * p.gstep[2];
        oi = valid ? (ix * p.gdim[1] + iy) * p.gdim[2] + iz : 0;
        w = p.width;
    }
    return valid;
}

// ============================================================================
// vector family: A / A_cpu, G, A2 (all three directions), A3
// ============================================================================
template <int VOP, bool CONT>
struct VAcc32 {
    static constexpr int NO = (VOP == VOP_A) ? 1 : 3;
    double s[NO];         // running sums: FP64 totals of the FP32 partials
    float lp[NO], fp[NO];
    float P, Q[3];   // continuous-kernel sums inside the width: sum d.mu, sum mu
    __device__ void init() {
        for (int i = 0; i < NO; ++i) { s[i] = 0.0; lp[i] = 0.f; fp[i] = 0.f; }
        P = 0.f; Q[0] = Q[1] = Q[2] = 0.f;
    }
    // one interaction: d = x - y (FP32), mu / e / q of the source
    __device__ void add(float* dst, const float* d, float r2, bool in, const float* mu,
                        const float* e, const float* q) {
        const float ir = in ? 0.f : __builtin_amdgcn_rsqf(r2);
        const float dmu = add_ir(dst, d, ir, mu, e, q);
        if constexpr (CONT) {
            if (in) {
                P += dmu;
                for (int k = 0; k < 3; ++k) Q[k] += mu[k];
            }
        }
    }
    // the same with 1/r given (0 inside the width) and no continuous-kernel
    // sums; returns d.mu
    __device__ float add_ir(float* dst, const float* d, float ir, const float* mu, const float* e,
                            const float* q) {
        const float ir2 = ir * ir;
        const float ir3 = ir2 * ir;
        const float dmu = fmaf(d[0], mu[0], fmaf(d[1], mu[1], d[2] * mu[2]));
        if constexpr (VOP == VOP_A) {
            dst[0] = fmaf(dmu, ir3, dst[0]);
        } else if constexpr (VOP == VOP_G) {
            const float u = 3.f * dmu * (ir3 * ir2);
            for (int k = 0; k < 3; ++k) dst[k] = fmaf(-u, d[k], fmaf(mu[k], ir3, dst[k]));
        } else {
            const float t = dmu * ir3;
            for (int k = 0; k < 3; ++k) dst[k] = fmaf(e[k], t, fmaf(q[k], ir, dst[k]));
        }
        return dmu;
    }
    // the leaf partial and the far partial since the last leaf join the FP64
    // total (the pair sum is exact in FP64, one rounding per fold)
    __device__ void flush_leaf() {
        for (int i = 0; i < NO; ++i) {
            s[i] += (double)lp[i] + (double)fp[i];
            lp[i] = 0.f;
            fp[i] = 0.f;
        }
    }
};

template <int VOP, bool CONT>
struct VAcc64 {
    static constexpr int NO = (VOP == VOP_A) ? 1 : 3;
    double s[NO];
    double P, Q[3];
    __device__ void init() {
        for (int i = 0; i < NO; ++i) s[i] = 0.0;
        P = 0.0; Q[0] = Q[1] = Q[2] = 0.0;
    }
    __device__ void add(const double* d, double w, const double* mu, double an) {
        const double r2 = dist2_rn(d[0], d[1], d[2]);
        const double r = sqrt(r2);
        const double dmu = d[0] * mu[0] + d[1] * mu[1] + d[2] * mu[2];
        if (r >= w) {
            const double ir = 1.0 / r;
            const double ir3 = ir * ir * ir;
            if constexpr (VOP == VOP_A) {
                s[0] += dmu * ir3;
            } else if constexpr (VOP == VOP_G) {
                const double ir5 = ir3 * ir * ir;
                for (int k = 0; k < 3; ++k) s[k] += mu[k] * ir3 - 3.0 * d[k] * dmu * ir5;
            } else {
                const double t = dmu * ir3;
                const double h = 0.5 * an * ir;
                for (int k = 0; k < 3; ++k) s[k] += exp(0.5 * an * d[k]) * (t + h * mu[k]);
            }
        } else if constexpr (CONT) {
            P += dmu;
            for (int k = 0; k < 3; ++k) Q[k] += mu[k];
        }
    }
};

// outputs of one query of the vector family from its FP32c sums
template <int VOP, bool CONT>
__device__ inline void finish_v32(const TP& p, const Sink& sk, const VAcc32<VOP, CONT>& acc,
                                  int64_t oi, const double* x, double w) {
    double res[3];
    for (int k = 0; k < acc.NO; ++k) res[k] = acc.s[k];
    if constexpr (VOP == VOP_A) {
        double v = res[0];
        if constexpr (CONT) v += (double)acc.P / (w * w * w);
        emit(p, sk, oi, 1, 0, p.sign * v);
    } else if constexpr (VOP == VOP_G) {
        for (int k = 0; k < 3; ++k) emit(p, sk, oi, 3, k, res[k]);
    } else {
        double o[3];
        const double ew = CONT ? exp(p.an * w) : 0.0;
        for (int k = 0; k < 3; ++k) {
            o[k] = exp(0.5 * p.an * x[k]) * res[k];
            if constexpr (CONT)
                o[k] += ew * ((double)acc.P / (w * w * w) + 0.5 * p.an * (double)acc.Q[k] / w);
        }
        if constexpr (VOP == VOP_A3) {
            emit(p, sk, oi, 1, 0, ((o[0] + o[1]) + o[2]) / 3.0);
        } else if (p.kdir >= 0) {
            emit(p, sk, oi, 1, 0, sel3(o, p.kdir));
        } else {
            for (int k = 0; k < 3; ++k) emit(p, sk, oi, 3, k, o[k]);
        }
    }
}

// one wave: queries idx0 + lane (< qend), tree nodes [root, stop).  A piece
// (chain = true) starts at a node root inside the tree: each lane first
// replays the opening tests of root's ancestors, so that it is active at
// root only if its own walk reaches root, and otherwise resumes after the
// shallowest ancestor it takes far -- the piece then takes exactly the steps
// of the whole walk that fall in [root, stop).  Results go to sk.
// Y64: the near pairs' FP64 cutoff retests read the tree's FP64 points
// (false: the source records' float coordinates, exact only for a tree of
// FP32 points, but it keeps the points' address out of the walk's registers)
template <int P, int VOP, int QM, bool CONT, bool CNT, bool Y64 = true>
__device__ __forceinline__ void trav_v(
    const TP& p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, int64_t idx0, int64_t qend, int32_t root, int32_t stop,
    int32_t part = -1, bool chain = false) {
    const int lane = threadIdx.x & 63;
    const int64_t idx = idx0 + lane;
    double x[3], w;
    int64_t oi;
    const bool valid = load_query<QM>(p, pts64, idx, qend, x, w, oi);
    if (__ballot(valid) == 0) return;
    const Sink sk{part, lane};

    int32_t resume = valid ? root : INT_MAX;
    int64_t nfar = 0, nnear = 0, nvisit = 0;

    if constexpr (P == CAGR_PREC_FP32C) {
        const NodeV32* __restrict__ nodes = (const NodeV32*)nodes_;
        const SrcV32* __restrict__ src = (const SrcV32*)src_;
        // query as an unevaluated float sum hi + lo: d = (hi - y) + lo keeps the
        // FP64 query's position (tree points are exact floats, so lo = 0 there)
        float xf[3], xl[3] = {0.f, 0.f, 0.f};
        if constexpr (QM == QM_SELF) {
            const int64_t s = valid ? idx : 0;
            xf[0] = pts32[3 * s]; xf[1] = pts32[3 * s + 1]; xf[2] = pts32[3 * s + 2];
        } else {
            for (int k = 0; k < 3; ++k) {
                xf[k] = (float)x[k];
                xl[k] = (float)(x[k] - (double)xf[k]);
            }
        }
        const float wf = (float)w;
        const float w2 = wf * wf;
        const float cband = fmaf(p.cb.nb1, w2, p.cb.nb0);   // near pairs' cutoff band
        const int dsafe = __builtin_amdgcn_readfirstlane(
            far_safe_depth(p.open_factor, wave_max(valid ? w : 0.0)));
        if (chain) {   // the ancestors' opening tests, exactly as the steps below take them
            for (int32_t a = rec32(topo, root).parent; a >= 0;) {
                a = __builtin_amdgcn_readfirstlane(a);
                const TopoRec ta = rec32(topo, a);
                const NodeV32 na = rec32(nodes, a);
                float d[3] = {xf[0] - na.cf[0], xf[1] - na.cf[1], xf[2] - na.cf[2]};
                if constexpr (QM != QM_SELF)
                    for (int k = 0; k < 3; ++k) d[k] += xl[k];
                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                bool far = r2 > na.hi2;
                if (!far && !(r2 < na.lo2)) {
                    const double* c64 = p.r64 + 3 * (int64_t)a;
                    far = dist2_rn(x[0] - c64[0], x[1] - c64[1], x[2] - c64[2]) > ta.thr2;
                }
                if (far && valid) resume = ta.escape;
                a = ta.parent;
            }
        }
        VAcc32<VOP, CONT> acc;
        acc.init();
        for (int32_t n = root; n < stop;) {
            n = __builtin_amdgcn_readfirstlane(n);
            const TopoRec t = rec32(topo, n);
            const NodeV32 nr = rec32(nodes, n);
            // queries other than the tree's points: the FP64 centre (below)
            double c[3] = {0.0, 0.0, 0.0};
            if constexpr (QM != QM_SELF)
                for (int k = 0; k < 3; ++k) c[k] = p.r64[3 * (int64_t)n + k];
            // the step's records loaded together and waited for once (the far
            // interaction's fields would otherwise be a second round trip)
            CAGR_LOADS_HERE("s"(t.escape), "s"(t.info), "s"(nr.cf[0]), "s"(nr.cf[1]), "s"(nr.cf[2]),
                            "s"(nr.hi2), "s"(nr.lo2), "s"(nr.m[0]), "s"(nr.m[1]), "s"(nr.m[2]),
                            "s"(nr.e[0]), "s"(nr.e[1]), "s"(nr.e[2]), "s"(nr.q[0]), "s"(nr.q[1]),
                            "s"(nr.q[2]));
            if constexpr (QM != QM_SELF) CAGR_LOADS_HERE("s"(c[0]), "s"(c[1]), "s"(c[2]));
            __builtin_amdgcn_sched_barrier(0);
            // FP32 distance to the float centre: far above the record's hi2,
            // near below lo2; lanes between them take the FP64 test with the
            // upward pass's FP64 centre (open_band), so every decision is the
            // FP64 test's.  Each comparison is its own wave mask, combined on
            // the scalar unit (a ballot of a combined lane predicate, or a
            // predicate carried past the far branch, would go through a vector
            // register).
            // d: the tree's own points are floats, their difference to the
            // float centre is the FP32 one; other queries take the difference
            // to the FP64 centre in FP64 (as many instructions as the float
            // centre and the query's low part, without the centre's rounding,
            // which weighs in the far terms of deep nodes)
            float d[3];
            if constexpr (QM == QM_SELF) {
                for (int k = 0; k < 3; ++k) d[k] = xf[k] - nr.cf[k];
            } else {
                for (int k = 0; k < 3; ++k) d[k] = (float)(x[k] - c[k]);
            }
            const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
            const uint64_t am = __builtin_amdgcn_ballot_w64(resume <= n);
            const uint64_t fm = __builtin_amdgcn_ballot_w64(r2 > nr.hi2);
            const uint64_t lm = __builtin_amdgcn_ballot_w64(r2 < nr.lo2);
            uint64_t far_m = am & fm, near_m = am & lm;
            if (const uint64_t amb = am & ~(fm | lm); __builtin_expect(amb != 0, 0)) {
                const bool ak = __builtin_amdgcn_inverse_ballot_w64(amb);
                bool far64 = false;
                if (ak) {
                    const double* c64 = p.r64 + 3 * (int64_t)n;
                    far64 = dist2_rn(x[0] - c64[0], x[1] - c64[1], x[2] - c64[2]) > t.thr2;
                }
                far_m |= __builtin_amdgcn_ballot_w64(ak & far64);
                near_m |= __builtin_amdgcn_ballot_w64(ak & !far64);
            }
            if constexpr (CNT) nvisit += (resume <= n) ? 1 : 0;
            if (__builtin_amdgcn_inverse_ballot_w64(far_m)) {
                if constexpr (CNT) {
                    ++nfar;
                } else {
                    // far at depth <= dsafe proves r > w: the cutoff test (and
                    // its select) only below that depth
                    const double* c64 = p.r64 + 3 * (int64_t)n;
                    if constexpr (CONT) {
                        bool in = false;
                        if (topo_depth(t) > dsafe) in = inside_centre(r2, w2, wf, nr.ce, x, c64, w);
                        acc.add(acc.fp, d, r2, in, nr.m, nr.e, nr.q);
                    } else {
                        float ir = __builtin_amdgcn_rsqf(r2);
                        if (topo_depth(t) > dsafe) {
                            if (inside_centre(r2, w2, wf, nr.ce, x, c64, w)) ir = 0.f;
                        }
                        acc.add_ir(acc.fp, d, ir, nr.m, nr.e, nr.q);
                    }
                }
            }
            const bool leaf = topo_leaf(t);
            if (leaf) {
                const bool nearl = __builtin_amdgcn_inverse_ballot_w64(near_m);
                if constexpr (CNT) {
                    if (nearl) nnear += t.end - t.begin;
                } else if (near_m != 0) {
                    const bool sure = leaf_surely_outside(xf, p.lbound[n], wf);
                    const bool fast = __ballot(nearl && !sure) == 0;
                    const uint64_t nm = near_m;
                    const int lr = fast ? spread_log2(__popcll(nm), t.end - t.begin) : 0;
                    if (lr > 0) {   // few near lanes: the idle lanes share the points
                        const Spread sp = spread_setup(nm, nearl, lr);
                        float qx[3], ql[3] = {0.f, 0.f, 0.f}, part[3] = {0.f, 0.f, 0.f};
                        for (int k = 0; k < 3; ++k) qx[k] = lane_pull(xf[k], sp.owner);
                        if constexpr (QM != QM_SELF)
                            for (int k = 0; k < 3; ++k) ql[k] = lane_pull(xl[k], sp.owner);
                        if (sp.part) {
                            for (int32_t j = t.begin + sp.g; j < t.end; j += 1 << lr) {
                                const SrcV32 sr = src[j];
                                float d[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1], qx[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += ql[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                acc.add(part, d, r2, false, sr.mu, sr.e, sr.q);
                            }
                        }
                        spread_sums<acc.NO>(part, lr);
                        for (int k = 0; k < acc.NO; ++k) {
                            const float v = lane_pull(part[k], sp.rank << lr);
                            if (nearl) acc.lp[k] += v;
                        }
                        if (nearl) acc.flush_leaf();
                    } else if (nearl) {
                        if (fast) {   // no pair of this leaf is within any lane's width
                            for_each_rec(src, t.begin, t.end, [&](const SrcV32& sr, int32_t) {
                                float d[3] = {xf[0] - sr.y[0], xf[1] - sr.y[1], xf[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += xl[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                acc.add(acc.lp, d, r2, false, sr.mu, sr.e, sr.q);
                            });
                        } else {
                            for_each_rec(src, t.begin, t.end, [&](const SrcV32& sr, int32_t j) {
                                float d[3] = {xf[0] - sr.y[0], xf[1] - sr.y[1], xf[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += xl[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                // the FP64 source: the tree's FP64 point, or (FP32
                                // points only) the record's own coordinates
                                double yr[3];
                                for (int k = 0; k < 3; ++k) yr[k] = (double)sr.y[k];
                                const bool in = inside_fp32(r2, w2, cband, x,
                                                            Y64 ? pts64 + 3 * (int64_t)j : yr, w);
                                acc.add(acc.lp, d, r2, in, sr.mu, sr.e, sr.q);
                            });
                        }
                        acc.flush_leaf();
                    }
                }
            }
            // one latch: past a leaf or a subtree no lane enters, the escape;
            // else the first child, with the far lanes parked past the subtree
            if (leaf || near_m == 0) {
                n = t.escape;
            } else {
                if (__builtin_amdgcn_inverse_ballot_w64(far_m)) resume = t.escape;
                n = n + 1;
            }
        }
        if constexpr (!CNT) {
            acc.flush_leaf();
            if (valid) finish_v32<VOP, CONT>(p, sk, acc, oi, x, w);
        }
    } else {
        const NodeV64* __restrict__ nodes = (const NodeV64*)nodes_;
        const SrcV64* __restrict__ src = (const SrcV64*)src_;
        if (chain) {   // the ancestors' opening tests (see the FP32c walk)
            for (int32_t a = rec32(topo, root).parent; a >= 0;) {
                a = __builtin_amdgcn_readfirstlane(a);
                const TopoRec ta = rec32(topo, a);
                const NodeV64 na = rec32(nodes, a);
                const bool far =
                    dist2_rn(x[0] - na.c[0], x[1] - na.c[1], x[2] - na.c[2]) > ta.thr2;
                if (far && valid) resume = ta.escape;
                a = ta.parent;
            }
        }
        VAcc64<VOP, CONT> acc;
        acc.init();
        for (int32_t n = root; n < stop;) {
            n = __builtin_amdgcn_readfirstlane(n);
            const TopoRec t = rec32(topo, n);
            const NodeV64 nr = rec32(nodes, n);
            // the step's records loaded together and waited for once
            CAGR_LOADS_HERE("s"(t.thr2), "s"(t.escape), "s"(t.info), "s"(nr.c[0]), "s"(nr.c[1]),
                            "s"(nr.c[2]), "s"(nr.m[0]), "s"(nr.m[1]), "s"(nr.m[2]));
            __builtin_amdgcn_sched_barrier(0);
            const double d[3] = {x[0] - nr.c[0], x[1] - nr.c[1], x[2] - nr.c[2]};
            // comparisons as wave masks, as in the FP32c walk
            const uint64_t am = __builtin_amdgcn_ballot_w64(resume <= n);
            const uint64_t fm = __builtin_amdgcn_ballot_w64(dist2_rn(d[0], d[1], d[2]) > t.thr2);
            const uint64_t near_m = am & ~fm, far_m = am & fm;
            if constexpr (CNT) nvisit += (resume <= n) ? 1 : 0;
            if (__builtin_amdgcn_inverse_ballot_w64(far_m)) {
                if constexpr (CNT) ++nfar;
                else acc.add(d, w, nr.m, p.an);
            }
            const bool leaf = topo_leaf(t);
            if (leaf) {
                const bool nearl = __builtin_amdgcn_inverse_ballot_w64(near_m);
                if constexpr (CNT) {
                    if (nearl) nnear += t.end - t.begin;
                } else if (const uint64_t nm = near_m; nm != 0) {
                    const int lr = spread_log2(__popcll(nm), t.end - t.begin);
                    if (lr > 0) {   // few near lanes: the idle lanes share the points
                        const Spread sp = spread_setup(nm, nearl, lr);
                        double qx[3];
                        for (int k = 0; k < 3; ++k) qx[k] = lane_pull(x[k], sp.owner);
                        const double qw = lane_pull(w, sp.owner);
                        VAcc64<VOP, CONT> part;
                        part.init();
                        if (sp.part) {
                            for (int32_t j = t.begin + sp.g; j < t.end; j += 1 << lr) {
                                const SrcV64 sr = src[j];
                                const double dd[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1],
                                                      qx[2] - sr.y[2]};
                                part.add(dd, qw, sr.mu, p.an);
                            }
                        }
                        const int from = sp.rank << lr;
                        for (int k = 0; k < acc.NO; ++k) {
                            const double v = lane_pull(spread_sum(part.s[k], lr), from);
                            if (nearl) acc.s[k] += v;
                        }
                        if constexpr (CONT) {
                            const double v = lane_pull(spread_sum(part.P, lr), from);
                            if (nearl) acc.P += v;
                            for (int k = 0; k < 3; ++k) {
                                const double u = lane_pull(spread_sum(part.Q[k], lr), from);
                                if (nearl) acc.Q[k] += u;
                            }
                        }
                    } else if (nearl) {
                        for_each_rec(src, t.begin, t.end, [&](const SrcV64& sr, int32_t) {
                            const double dd[3] = {x[0] - sr.y[0], x[1] - sr.y[1], x[2] - sr.y[2]};
                            acc.add(dd, w, sr.mu, p.an);
                        });
                    }
                }
            }
            // one latch: past a leaf or a subtree no lane enters, the escape;
            // else the first child, with the far lanes parked past the subtree
            if (leaf || near_m == 0) {
                n = t.escape;
            } else {
                if (__builtin_amdgcn_inverse_ballot_w64(far_m)) resume = t.escape;
                n = n + 1;
            }
        }
        if constexpr (!CNT) {
            if (valid) {
                if constexpr (VOP == VOP_A) {
                    double v = acc.s[0];
                    if constexpr (CONT) v += acc.P / (w * w * w);
                    emit(p, sk, oi, 1, 0, p.sign * v);
                } else if constexpr (VOP == VOP_G) {
                    for (int k = 0; k < 3; ++k) emit(p, sk, oi, 3, k, acc.s[k]);
                } else {
                    double o[3];
                    for (int k = 0; k < 3; ++k) {
                        o[k] = acc.s[k];
                        if constexpr (CONT)
                            o[k] += exp(p.an * w) * (acc.P / (w * w * w) + 0.5 * p.an * acc.Q[k] / w);
                    }
                    if constexpr (VOP == VOP_A3) {
                        emit(p, sk, oi, 1, 0, ((o[0] + o[1]) + o[2]) / 3.0);
                    } else if (p.kdir >= 0) {
                        emit(p, sk, oi, 1, 0, sel3(o, p.kdir));
                    } else {
                        for (int k = 0; k < 3; ++k) emit(p, sk, oi, 3, k, o[k]);
                    }
                }
            }
        }
    }
    if constexpr (CNT) {
        if (valid) {
            if (p.cnt_far) p.cnt_far[oi] = nfar;
            if (p.cnt_near) p.cnt_near[oi] = nnear;
            if (p.cnt_visit) p.cnt_visit[oi] = nvisit;
        }
    }
}

// ============================================================================
// scalar family: AT (an = 0), A2Tk (K = 1, direction kdir), sum_k A2Tk (K = 3)
// ============================================================================
// (pieces: as trav_v, one resume per set)
template <int P, int K, int QM, bool CNT, bool Y64 = true>
__device__ __forceinline__ void trav_s(
    const TP& p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, int64_t idx0, int64_t qend, int32_t root, int32_t stop,
    int32_t part = -1, bool chain = false) {
    const int lane = threadIdx.x & 63;
    const int64_t idx = idx0 + lane;
    double x[3], w;
    int64_t oi;
    const bool valid = load_query<QM>(p, pts64, idx, qend, x, w, oi);
    if (__ballot(valid) == 0) return;
    const Sink sk{part, lane};

    int32_t resume[K];
    for (int k = 0; k < K; ++k) resume[k] = valid ? root : INT_MAX;
    int64_t nfar = 0, nnear = 0, nvisit = 0;
    const double han = 0.5 * p.an;

    if constexpr (P == CAGR_PREC_FP32C) {
        const NodeS32* __restrict__ nodes = (const NodeS32*)nodes_;
        const SrcS32<K>* __restrict__ src = (const SrcS32<K>*)src_;
        float xf[3], xl[3] = {0.f, 0.f, 0.f};
        if constexpr (QM == QM_SELF) {
            const int64_t s = valid ? idx : 0;
            xf[0] = pts32[3 * s]; xf[1] = pts32[3 * s + 1]; xf[2] = pts32[3 * s + 2];
        } else {
            for (int k = 0; k < 3; ++k) {
                xf[k] = (float)x[k];
                xl[k] = (float)(x[k] - (double)xf[k]);
            }
        }
        float eq[K];   // exp(An * x_dir / 2) of each set's direction
        for (int k = 0; k < K; ++k) eq[k] = (float)exp(han * sel3(x, p.kdir + k));
        const float hanf = (float)han;
        const float wf = (float)w;
        const float w2 = wf * wf;
        const float cband = fmaf(p.cb.nb1, w2, p.cb.nb0);   // near pairs' cutoff band
        const int dsafe = __builtin_amdgcn_readfirstlane(
            far_safe_depth(p.open_factor, wave_max(valid ? w : 0.0)));
        if (chain) {   // the ancestors' opening tests of every set (see trav_v)
            for (int32_t a = rec32(topo, root).parent; a >= 0;) {
                a = __builtin_amdgcn_readfirstlane(a);
                const TopoRec ta = rec32(topo, a);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const NodeS32 na = rec32(nodes, a * K + k);
                    float d[3] = {xf[0] - na.cf[0], xf[1] - na.cf[1], xf[2] - na.cf[2]};
                    if constexpr (QM != QM_SELF)
                        for (int cc = 0; cc < 3; ++cc) d[cc] += xl[cc];
                    const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                    bool far = r2 > na.hi2;
                    if (!far && !(r2 < na.lo2)) {
                        const double* c64 = p.r64 + 3 * ((int64_t)k * p.nm + a);
                        far = dist2_rn(x[0] - c64[0], x[1] - c64[1], x[2] - c64[2]) > ta.thr2;
                    }
                    if (far && valid) resume[k] = ta.escape;
                }
                a = ta.parent;
            }
        }
        double s[3] = {0.0, 0.0, 0.0};   // FP64 totals of the FP32 partials
        float lp[3] = {0.f, 0.f, 0.f}, fp[3] = {0.f, 0.f, 0.f};
        // one near-pair interaction of all K sets (weights mk), into dst
        auto near_pair = [&](float* dst, const SrcS32<K>& sr, const float* d, float r2, bool in,
                             const float* mk) {
            const float ir = in ? 0.f : __builtin_amdgcn_rsqf(r2);
            const float ir3 = ir * ir * ir;
            float g[K];
#pragma unroll
            for (int k = 0; k < K; ++k) g[k] = mk[k] * sr.a[k];
            float gsum = g[0];
#pragma unroll
            for (int k = 1; k < K; ++k) gsum += g[k];
            const float gi3 = gsum * ir3;
            for (int cc = 0; cc < 3; ++cc) dst[cc] = fmaf(gi3, d[cc], dst[cc]);
            const float hir = hanf * ir;
            if constexpr (K == 3) {
#pragma unroll
                for (int k = 0; k < 3; ++k) dst[k] = fmaf(g[k], hir, dst[k]);
            } else {
                const float v = g[0] * hir;
                for (int cc = 0; cc < 3; ++cc) dst[cc] += (cc == p.kdir) ? v : 0.f;
            }
        };
        for (int32_t n = root; n < stop;) {
            // internal nodes until a leaf or the end of the tree (this inner
            // loop leaves the leaf-only state alone, so the register allocator
            // keeps its loop-carried values in place), then the leaf.  The
            // near lanes of each set travel as wave masks (scalar registers):
            // per-lane booleans carried out of the loop would be kept as
            // 0/1 values in vector registers and converted back on every step
            TopoRec t;
            uint64_t nm[K];
            for (;;) {
            n = __builtin_amdgcn_readfirstlane(n);
            t = rec32(topo, n);
            NodeS32 nrs[K];
#pragma unroll
            for (int k = 0; k < K; ++k) nrs[k] = rec32(nodes, n * K + k);
            // one wait for the whole step's records
            if constexpr (K == 3) {
                CAGR_LOADS_HERE("s"(t.escape), "s"(t.info),
                                "s"(nrs[0].cf[0]), "s"(nrs[0].cf[1]), "s"(nrs[0].cf[2]),
                                "s"(nrs[0].se), "s"(nrs[0].hi2), "s"(nrs[0].lo2),
                                "s"(nrs[1].cf[0]), "s"(nrs[1].cf[1]), "s"(nrs[1].cf[2]),
                                "s"(nrs[1].se), "s"(nrs[1].hi2), "s"(nrs[1].lo2),
                                "s"(nrs[2].cf[0]), "s"(nrs[2].cf[1]), "s"(nrs[2].cf[2]),
                                "s"(nrs[2].se), "s"(nrs[2].hi2), "s"(nrs[2].lo2));
            } else {
                CAGR_LOADS_HERE("s"(t.escape), "s"(t.info),
                                "s"(nrs[0].cf[0]), "s"(nrs[0].cf[1]), "s"(nrs[0].cf[2]),
                                "s"(nrs[0].se), "s"(nrs[0].hi2), "s"(nrs[0].lo2));
            }
            __builtin_amdgcn_sched_barrier(0);
            const bool safe = topo_depth(t) <= dsafe;
            // far interaction of set k (FP32 d and r2 to its float centre)
            auto far_set = [&](int k, const NodeS32& nr, const float* d, float r2) {
                // far at depth <= dsafe proves r > w: the cutoff test (and its
                // select) only below that depth
                float ir = __builtin_amdgcn_rsqf(r2);
                if (!safe) {
                    if (inside_centre(r2, w2, wf, nr.ce, x, p.r64 + 3 * ((int64_t)k * p.nm + n), w))
                        ir = 0.f;
                }
                const float gi = eq[k] * nr.se * ir;   // g / r
                const float hg = hanf * gi;
                const float gi3 = gi * (ir * ir);      // g / r^3
                for (int cc = 0; cc < 3; ++cc) fp[cc] = fmaf(gi3, d[cc], fp[cc]);
                if constexpr (K == 3) {
                    fp[k] += hg;
                } else {
                    for (int cc = 0; cc < 3; ++cc) fp[cc] += (cc == p.kdir) ? hg : 0.f;
                }
            };
            // d to set k's centre: FP32 to the float centre for the tree's
            // own points, FP64 to the FP64 centre for other queries (see
            // trav_v)
            auto dist_set = [&](int k, const NodeS32& nr, float* d) {
                if constexpr (QM == QM_SELF) {
                    for (int cc = 0; cc < 3; ++cc) d[cc] = xf[cc] - nr.cf[cc];
                } else {
                    const double* c64 = p.r64 + 3 * ((int64_t)k * p.nm + n);
                    for (int cc = 0; cc < 3; ++cc) d[cc] = (float)(x[cc] - c64[cc]);
                }
                return fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
            };
            uint64_t amb[K], fmk[K];
            uint64_t any_amb = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                // FP32 distance to the float centre: far above the record's
                // hi2, near below lo2 (lo2 < hi2); lanes between them wait for
                // the FP64 test after the loop (open_band)
                const NodeS32& nr = nrs[k];
                float d[3];
                const float r2 = dist_set(k, nr, d);
                // each comparison as its own wave mask, combined on the
                // scalar unit (a ballot of a combined lane predicate would be
                // taken through a vector register)
                const uint64_t am = __builtin_amdgcn_ballot_w64(resume[k] <= n);
                const uint64_t fm = __builtin_amdgcn_ballot_w64(r2 > nr.hi2);
                const uint64_t lm = __builtin_amdgcn_ballot_w64(r2 < nr.lo2);
                nm[k] = am & lm;
                fmk[k] = am & fm;
                amb[k] = am & ~(fm | lm);
                any_amb |= amb[k];
                if constexpr (CNT) nvisit += (resume[k] <= n) ? 1 : 0;
                if (__builtin_amdgcn_inverse_ballot_w64(fmk[k])) {
                    if constexpr (CNT) ++nfar;
                    else far_set(k, nr, d, r2);
                }
            }
            if (__builtin_expect(any_amb != 0, 0)) {
                // the FP64 test with the upward pass's FP64 centre, then the
                // same far / near work as above
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const bool ak = __builtin_amdgcn_inverse_ballot_w64(amb[k]);
                    bool far64 = false;
                    if (ak) {
                        const double* c64 = p.r64 + 3 * ((int64_t)k * p.nm + n);
                        far64 = dist2_rn(x[0] - c64[0], x[1] - c64[1], x[2] - c64[2]) > t.thr2;
                    }
                    if (ak & far64) {
                        if constexpr (CNT) {
                            ++nfar;
                        } else {
                            float d[3];
                            const float r2 = dist_set(k, nrs[k], d);
                            far_set(k, nrs[k], d, r2);
                        }
                    }
                    fmk[k] |= __builtin_amdgcn_ballot_w64(ak & far64);
                    nm[k] |= __builtin_amdgcn_ballot_w64(ak & !far64);
                }
            }
            if (topo_leaf(t)) break;
            uint64_t nm_any = nm[0];
#pragma unroll
            for (int k = 1; k < K; ++k) nm_any |= nm[k];
            // into the first child with each set's far lanes parked past the
            // subtree, or past the subtree (nothing to update)
            if (nm_any == 0) {
                n = t.escape;
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (__builtin_amdgcn_inverse_ballot_w64(fmk[k])) resume[k] = t.escape;
                n = n + 1;
            }
            if (n >= stop) break;
            }
            bool nearl[K];
            uint64_t nm_any = nm[0];
#pragma unroll
            for (int k = 1; k < K; ++k) nm_any |= nm[k];
#pragma unroll
            for (int k = 0; k < K; ++k) nearl[k] = __builtin_amdgcn_inverse_ballot_w64(nm[k]);
            const bool any_near = __builtin_amdgcn_inverse_ballot_w64(nm_any);
            const bool leaf = topo_leaf(t);
            if (leaf) {
                if constexpr (CNT) {
                    for (int k = 0; k < K; ++k) nnear += nearl[k] ? t.end - t.begin : 0;
                } else if (nm_any != 0) {
                    const bool sure = leaf_surely_outside(xf, p.lbound[n], wf);
                    const bool fast = __ballot(any_near && !sure) == 0;
                    const int lr = fast ? spread_log2(__popcll(nm_any), t.end - t.begin) : 0;
                    if (lr > 0) {   // few near lanes: the idle lanes share the points
                        const Spread sp = spread_setup(nm_any, any_near, lr);
                        float mk[K], qx[3], ql[3] = {0.f, 0.f, 0.f}, part[3] = {0.f, 0.f, 0.f};
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            mk[k] = lane_pull(nearl[k] ? eq[k] : 0.f, sp.owner);
                        for (int k = 0; k < 3; ++k) qx[k] = lane_pull(xf[k], sp.owner);
                        if constexpr (QM != QM_SELF)
                            for (int k = 0; k < 3; ++k) ql[k] = lane_pull(xl[k], sp.owner);
                        if (sp.part) {
                            for (int32_t j = t.begin + sp.g; j < t.end; j += 1 << lr) {
                                const SrcS32<K> sr = src[j];
                                float d[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1], qx[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += ql[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                near_pair(part, sr, d, r2, false, mk);
                            }
                        }
                        spread_sums<3>(part, lr);
                        for (int cc = 0; cc < 3; ++cc) {
                            const float v = lane_pull(part[cc], sp.rank << lr);
                            if (any_near) lp[cc] += v;
                        }
                        if (any_near)
                            for (int cc = 0; cc < 3; ++cc) {
                                s[cc] += (double)lp[cc] + (double)fp[cc];
                                lp[cc] = 0.f;
                                fp[cc] = 0.f;
                            }
                    } else if (any_near) {
                        float mk[K];
#pragma unroll
                        for (int k = 0; k < K; ++k) mk[k] = nearl[k] ? eq[k] : 0.f;
                        if (fast) {   // no pair of this leaf is within any lane's width
                            for_each_rec(src, t.begin, t.end, [&](const SrcS32<K>& sr, int32_t) {
                                float d[3] = {xf[0] - sr.y[0], xf[1] - sr.y[1], xf[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += xl[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                near_pair(lp, sr, d, r2, false, mk);
                            });
                        } else {
                            for_each_rec(src, t.begin, t.end, [&](const SrcS32<K>& sr, int32_t j) {
                                float d[3] = {xf[0] - sr.y[0], xf[1] - sr.y[1], xf[2] - sr.y[2]};
                                if constexpr (QM != QM_SELF)
                                    for (int k = 0; k < 3; ++k) d[k] += xl[k];
                                const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
                                // the FP64 source: the tree's FP64 point, or (FP32
                                // points only) the record's own coordinates
                                double yr[3];
                                for (int k = 0; k < 3; ++k) yr[k] = (double)sr.y[k];
                                const bool in = inside_fp32(r2, w2, cband, x,
                                                            Y64 ? pts64 + 3 * (int64_t)j : yr, w);
                                near_pair(lp, sr, d, r2, in, mk);
                            });
                        }
                        for (int cc = 0; cc < 3; ++cc) {
                            s[cc] += (double)lp[cc] + (double)fp[cc];
                            lp[cc] = 0.f;
                            fp[cc] = 0.f;
                        }
                    }
                }
                n = t.escape;
            }
        }
        if constexpr (!CNT) {
            if (valid)
                for (int cc = 0; cc < 3; ++cc) emit(p, sk, oi, 3, cc, s[cc] + (double)fp[cc]);
        }
    } else {
        const NodeS64* __restrict__ nodes = (const NodeS64*)nodes_;
        const SrcS64<K>* __restrict__ src = (const SrcS64<K>*)src_;
        if (chain) {   // the ancestors' opening tests of every set (see trav_v)
            for (int32_t a = rec32(topo, root).parent; a >= 0;) {
                a = __builtin_amdgcn_readfirstlane(a);
                const TopoRec ta = rec32(topo, a);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const NodeS64 na = rec32(nodes, a * K + k);
                    const bool far =
                        dist2_rn(x[0] - na.c[0], x[1] - na.c[1], x[2] - na.c[2]) > ta.thr2;
                    if (far && valid) resume[k] = ta.escape;
                }
                a = ta.parent;
            }
        }
        double acc[3] = {0.0, 0.0, 0.0};
        // one A2T_dir interaction (AT when an = 0) of scalar sv at d, width wq,
        // into dst
        auto add_to = [&](double* dst, const double* d, double sv, int dir, double wq) {
            const double r2 = dist2_rn(d[0], d[1], d[2]);
            const double r = sqrt(r2);
            if (r >= wq) {
                const double ir = 1.0 / r;
                const double g = exp(han * sel3(d, dir)) * sv;
                const double gi3 = g * (ir * ir * ir);
                const double hg = han * g * ir;
                for (int cc = 0; cc < 3; ++cc) dst[cc] += gi3 * d[cc] + (cc == dir ? hg : 0.0);
            }
        };
        auto add = [&](const double* d, double sv, int dir) { add_to(acc, d, sv, dir, w); };
        for (int32_t n = root; n < stop;) {
            // internal nodes until a leaf or the end of the tree, then the leaf
            // (see the FP32c walk)
            TopoRec t;
            uint64_t nm[K], fmk[K];   // near / far lanes of each set (wave masks, as in the FP32c walk)
            for (;;) {
            n = __builtin_amdgcn_readfirstlane(n);
            t = rec32(topo, n);
            NodeS64 nrs[K];
#pragma unroll
            for (int k = 0; k < K; ++k) nrs[k] = rec32(nodes, n * K + k);
            // the step's records loaded together and waited for once
            if constexpr (K == 3) {
                CAGR_LOADS_HERE("s"(t.thr2), "s"(t.escape), "s"(t.info),
                                "s"(nrs[0].c[0]), "s"(nrs[0].c[1]), "s"(nrs[0].c[2]), "s"(nrs[0].s),
                                "s"(nrs[1].c[0]), "s"(nrs[1].c[1]), "s"(nrs[1].c[2]), "s"(nrs[1].s),
                                "s"(nrs[2].c[0]), "s"(nrs[2].c[1]), "s"(nrs[2].c[2]), "s"(nrs[2].s));
            } else {
                CAGR_LOADS_HERE("s"(t.thr2), "s"(t.escape), "s"(t.info),
                                "s"(nrs[0].c[0]), "s"(nrs[0].c[1]), "s"(nrs[0].c[2]), "s"(nrs[0].s));
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const NodeS64& nr = nrs[k];
                const double d[3] = {x[0] - nr.c[0], x[1] - nr.c[1], x[2] - nr.c[2]};
                const uint64_t am = __builtin_amdgcn_ballot_w64(resume[k] <= n);
                const uint64_t fm =
                    __builtin_amdgcn_ballot_w64(dist2_rn(d[0], d[1], d[2]) > t.thr2);
                nm[k] = am & ~fm;
                fmk[k] = am & fm;
                if constexpr (CNT) nvisit += (resume[k] <= n) ? 1 : 0;
                if (__builtin_amdgcn_inverse_ballot_w64(fmk[k])) {
                    if constexpr (CNT) ++nfar;
                    else add(d, nr.s, p.kdir + k);
                }
            }
            if (topo_leaf(t)) break;
            uint64_t nm_any = nm[0];
#pragma unroll
            for (int k = 1; k < K; ++k) nm_any |= nm[k];
            if (nm_any == 0) {
                n = t.escape;
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (__builtin_amdgcn_inverse_ballot_w64(fmk[k])) resume[k] = t.escape;
                n = n + 1;
            }
            if (n >= stop) break;
            }
            bool nearl[K];
            uint64_t nm_any = nm[0];
#pragma unroll
            for (int k = 1; k < K; ++k) nm_any |= nm[k];
#pragma unroll
            for (int k = 0; k < K; ++k) nearl[k] = __builtin_amdgcn_inverse_ballot_w64(nm[k]);
            const bool any_near = __builtin_amdgcn_inverse_ballot_w64(nm_any);
            const bool leaf = topo_leaf(t);
            if (leaf) {
                if constexpr (CNT) {
                    for (int k = 0; k < K; ++k) nnear += nearl[k] ? t.end - t.begin : 0;
                } else if (nm_any != 0) {
                    const int lr = spread_log2(__popcll(nm_any), t.end - t.begin);
                    if (lr > 0) {   // few near lanes: the idle lanes share the points
                        const Spread sp = spread_setup(nm_any, any_near, lr);
                        int bits = 0;
                        for (int k = 0; k < K; ++k) bits |= nearl[k] ? 1 << k : 0;
                        const int qbits = lane_pull(bits, sp.owner);
                        double qx[3], part[3] = {0.0, 0.0, 0.0};
                        for (int k = 0; k < 3; ++k) qx[k] = lane_pull(x[k], sp.owner);
                        const double qw = lane_pull(w, sp.owner);
                        if (sp.part) {
                            for (int32_t j = t.begin + sp.g; j < t.end; j += 1 << lr) {
                                const SrcS64<K> sr = src[j];
                                const double d[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1],
                                                     qx[2] - sr.y[2]};
                                for (int k = 0; k < K; ++k)
                                    if (qbits & (1 << k)) add_to(part, d, sr.s[k], p.kdir + k, qw);
                            }
                        }
                        spread_sums<3>(part, lr);
                        for (int cc = 0; cc < 3; ++cc) {
                            const double v = lane_pull(part[cc], sp.rank << lr);
                            if (any_near) acc[cc] += v;
                        }
                    } else if (any_near) {
                        for_each_rec(src, t.begin, t.end, [&](const SrcS64<K>& sr, int32_t) {
                            const double d[3] = {x[0] - sr.y[0], x[1] - sr.y[1], x[2] - sr.y[2]};
                            for (int k = 0; k < K; ++k)
                                if (nearl[k]) add(d, sr.s[k], p.kdir + k);
                        });
                    }
                }
                n = t.escape;
            }
        }
        if constexpr (!CNT) {
            if (valid)
                for (int cc = 0; cc < 3; ++cc) emit(p, sk, oi, 3, cc, acc[cc]);
        }
    }
    if constexpr (CNT) {
        if (valid) {
            if (p.cnt_far) p.cnt_far[oi] = nfar;
            if (p.cnt_near) p.cnt_near[oi] = nnear;
            if (p.cnt_visit) p.cnt_visit[oi] = nvisit;
        }
    }
}

// ---------------------------------------------------------------- kernels
// what the wave in dispatch slot `slot` walks: a piece, or (no pieces) the
// slot's wave in dispatch order on the whole tree
struct WaveWork {
    int64_t q0;
    int32_t lo, hi;
    int32_t part;
    bool chain;
};
__device__ inline WaveWork wave_work(const TP& p, int64_t slot_) {
    // the slot is the same for the whole wave: made a scalar, the piece or
    // dispatch-order entry is a scalar load and its fields stay in scalar
    // registers (a vector copy would turn the walk's loop bounds into
    // per-lane values)
    const int32_t slot = __builtin_amdgcn_readfirstlane((int32_t)slot_);
    WaveWork ww{0, p.root, p.stop, -1, false};
    if (p.pieces) {
        const int4 pc = p.pieces[slot];
        ww.q0 = (int64_t)pc.x * 64;
        ww.lo = pc.y;
        ww.hi = pc.z;
        ww.part = pc.w;
        ww.chain = pc.y != p.root;
    } else {
        ww.q0 = (p.tile_order ? (int64_t)p.tile_order[slot] : slot) * 64;
    }
    return ww;
}

// one wave per 64 consecutive queries, every wave on the whole tree; or, for
// the self queries of a tree with walk pieces, one wave per piece
// {wave, first node, end node, partial slot}: the walk of that wave over the
// node range, into the output (slot -1) or its partial slot
template <int P, int VOP, int QM, bool CONT, bool CNT>
__global__ __launch_bounds__(kThreads) void k_trav_v(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32) {
    const int64_t slot =
        xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const WaveWork ww = wave_work(p, slot);
    trav_v<P, VOP, QM, CONT, CNT>(p, topo, nodes_, src_, pts64, pts32, ww.q0, p.nq, ww.lo, ww.hi,
                                  ww.part, ww.chain);
}

template <int P, int K, int QM, bool CNT>
__global__ __launch_bounds__(kThreads) void k_trav_s(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32) {
    const int64_t slot =
        xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const WaveWork ww = wave_work(p, slot);
    trav_s<P, K, QM, CNT>(p, topo, nodes_, src_, pts64, pts32, ww.q0, p.nq, ww.lo, ww.hi, ww.part,
                          ww.chain);
}

// The FP32c walks over the tree's own points, held to 64 VGPRs so that 8
// waves fit per SIMD: a wave walks for about a third of the launch, and more
// resident waves shorten the launch's tail (A2 -5 % at 1M points).  The FP64
// walks, the continuous kernel and the other query modes would spill.
template <int VOP, int QM, bool CONT>
__global__ __launch_bounds__(kThreads, 8) void k_trav_v32(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32) {
    const int64_t slot =
        xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const WaveWork ww = wave_work(p, slot);
    trav_v<CAGR_PREC_FP32C, VOP, QM, CONT, false, false>(p, topo, nodes_, src_, pts64, pts32, ww.q0,
                                                         p.nq, ww.lo, ww.hi, ww.part, ww.chain);
}

template <int K, int QM>
__global__ __launch_bounds__(kThreads, 8) void k_trav_s32(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32) {
    const int64_t slot =
        xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const WaveWork ww = wave_work(p, slot);
    trav_s<CAGR_PREC_FP32C, K, QM, false, false>(p, topo, nodes_, src_, pts64, pts32, ww.q0, p.nq,
                                                 ww.lo, ww.hi, ww.part, ww.chain);
}

// Persistent form: a grid sized to the device pulls wave items {root, stop,
// q0, qn} (a cloud's tree and up to 64 of its sorted points) from atomic
// queues until all are empty.  The pull order is cut into one chunk per XCD
// (block b runs on XCD b mod 8): a chunk is a contiguous Morton range, so an
// XCD's L2 holds the nodes and sources of its own part of space, and, once
// the items are costed, the chunks carry equal cost.  A wave drains its own
// chunk, then steals from the others.  Any block shape up to 512 threads.
struct Items {
    const int4* items;
    const int32_t* order;        // pull position -> item (null: identity)
    int32_t bound[kQueues + 1];  // chunk c = pull positions [bound[c], bound[c+1])
    int32_t* counter;            // chunk c's counter at counter[c * kCounterStride]
};

__device__ __forceinline__ bool next_item(const Items& q, int4& item) {
    const int home = (int)(blockIdx.x % kQueues);
    for (int s = 0; s < kQueues; ++s) {
        const int c = (home + s) % kQueues;
        const int32_t lo = q.bound[c], hi = q.bound[c + 1];
        if (lo >= hi) continue;
        int it = 0;
        if ((threadIdx.x & 63) == 0) it = atomicAdd(q.counter + c * kCounterStride, 1);
        it = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(it, 0));
        if (it < hi - lo) {
            const int32_t pos = lo + it;
            item = q.items[q.order ? q.order[pos] : pos];
            return true;
        }
    }
    return false;
}

template <int P, int VOP, bool CONT, bool CNT>
__global__ __launch_bounds__(512) void k_trav_v_items(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, Items q) {
    int4 it;
    while (next_item(q, it))
        trav_v<P, VOP, QM_SELF, CONT, CNT>(p, topo, nodes_, src_, pts64, pts32, it.z, it.z + it.w,
                                           it.x, it.y);
}

template <int P, int K, bool CNT>
__global__ __launch_bounds__(512) void k_trav_s_items(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, Items q) {
    int4 it;
    while (next_item(q, it))
        trav_s<P, K, QM_SELF, CNT>(p, topo, nodes_, src_, pts64, pts32, it.z, it.z + it.w, it.x,
                                   it.y);
}

// the FP32c persistent walks held to 64 VGPRs: the default grid is sized
// for 8 waves per SIMD (as the regular-grid FP32c walks)
template <int VOP>
__global__ __launch_bounds__(512, 8) void k_trav_v_items32(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, Items q) {
    int4 it;
    while (next_item(q, it))
        trav_v<CAGR_PREC_FP32C, VOP, QM_SELF, false, false, false>(p, topo, nodes_, src_, pts64,
                                                                  pts32, it.z, it.z + it.w, it.x,
                                                                  it.y);
}

template <int K>
__global__ __launch_bounds__(512, 8) void k_trav_s_items32(
    TP p, const TopoRec* __restrict__ topo, const void* __restrict__ nodes_,
    const void* __restrict__ src_, const double* __restrict__ pts64,
    const float* __restrict__ pts32, Items q) {
    int4 it;
    while (next_item(q, it))
        trav_s<CAGR_PREC_FP32C, K, QM_SELF, false, false>(p, topo, nodes_, src_, pts64, pts32,
                                                          it.z, it.z + it.w, it.x, it.y);
}

__global__ void k_reset_counter(int32_t* c) {
    if (threadIdx.x < kQueues) c[threadIdx.x * kCounterStride] = 0;
}

// ---------------------------------------------------------------- walk pieces
// The output of every split wave {wave, first slot, pieces}: the sum of its
// pieces' partial slots in piece order (a fixed order, so the result does not
// depend on when the pieces ran), no values per query.  One wave per split
// wave, four to a block; a wave's partials are loaded 8 slots at a time before
// any is added (one memory round trip per 8 pieces, not one per piece).
// NO is a template argument so that a query's values are written by one
// store (the rows are scattered by the permutation: one store per row, not
// one per value).
constexpr int kCombineWaves = 4;
template <int NO>
__global__ __launch_bounds__(64 * kCombineWaves) void k_piece_combine(
    TP p, const int4* __restrict__ split, int64_t nsplit) {
    constexpr int no = NO;
    const int64_t ws = (int64_t)blockIdx.x * kCombineWaves + (threadIdx.x >> 6);
    if (ws >= nsplit) return;   // whole waves
    const int4 sw = split[ws];
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)sw.x * 64 + lane;
    if (s >= p.nq) return;
    const int64_t oi = p.perm[s];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < sw.z; j0 += 8) {
        double v[8][3];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double* pl = p.planes + ((int64_t)(sw.y + j0 + j) * 64 + lane) * 3;
#pragma unroll
            for (int k = 0; k < 3; ++k) v[j][k] = (j0 + j < sw.z && k < no) ? pl[k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (j0 + j < sw.z) acc[k] += v[j][k];
    }
    if (p.odtype == CAGR_F32) {
        struct Row { float v[NO]; };
        Row r;
        for (int k = 0; k < NO; ++k) r.v[k] = (float)acc[k];
        ((Row*)p.out)[oi] = r;
    } else {
        struct Row { double v[NO]; };
        Row r;
        for (int k = 0; k < NO; ++k) r.v[k] = acc[k];
        ((Row*)p.out)[oi] = r;
    }
}

// Cost of the steps of each wave's self-query walk (the decisions of the FP64
// test on the FP64 centres c64 [m, 3]): 1 per node step, plus kNearLeafCost
// per leaf step with near lanes (a near leaf's points, measured at about two
// to three node steps a leaf).  np == null: the total of each wave into
// total.  Else, for a wave cut into np[w] pieces, the node at which the cost
// before the step first reaches c * total[w] / np[w] (c = 1 .. np[w]-1) into
// cuts[w * pmax + c - 1] (m where it never does: an empty last piece).
constexpr int kNearLeafCost = 3;
__global__ __launch_bounds__(64) void k_walk_cost(const TopoRec* __restrict__ topo,
                                                  const double* __restrict__ c64,
                                                  const double* __restrict__ pts64, int64_t n,
                                                  int32_t m, const int32_t* __restrict__ np,
                                                  int32_t* __restrict__ total,
                                                  int32_t* __restrict__ cuts, int pmax) {
    const int64_t wave = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t s = wave * 64 + lane;
    const bool valid = s < n;
    const int64_t sc = valid ? s : 0;
    const double x[3] = {pts64[3 * sc], pts64[3 * sc + 1], pts64[3 * sc + 2]};
    const int npw = np ? np[wave] : 1;
    const int64_t tot = np ? total[wave] : 0;
    if (np && lane < npw - 1) cuts[wave * pmax + lane] = m;
    int32_t resume = valid ? 0 : INT_MAX;
    int64_t cost = 0;
    int c = 1;
    for (int32_t nd = 0; nd < m;) {
        nd = __builtin_amdgcn_readfirstlane(nd);
        while (c < npw && cost * npw >= c * tot) {
            if (lane == 0) cuts[wave * pmax + c - 1] = nd;
            ++c;
        }
        const TopoRec t = rec32(topo, nd);
        const double* cn = c64 + 3 * (int64_t)nd;
        const uint64_t am = __builtin_amdgcn_ballot_w64(resume <= nd);
        const uint64_t fm = __builtin_amdgcn_ballot_w64(
            dist2_rn(x[0] - cn[0], x[1] - cn[1], x[2] - cn[2]) > t.thr2);
        const uint64_t far_m = am & fm, near_m = am & ~fm;
        cost += 1;
        if (topo_leaf(t)) {
            if (near_m != 0) cost += kNearLeafCost;
            nd = t.escape;
        } else if (near_m != 0) {
            if (__builtin_amdgcn_inverse_ballot_w64(far_m)) resume = t.escape;
            nd = nd + 1;
        } else {
            nd = t.escape;
        }
    }
    if (!np && lane == 0) total[wave] = (int32_t)min(cost, (int64_t)INT_MAX);
}

// ---------------------------------------------------------------- dispatch
template <int P, int VOP, int QM, bool CNT>
int launch_v(const TP& tp, cagr_tree* t, int continuous, int64_t blocks, hipStream_t st) {
    if constexpr (P == CAGR_PREC_FP32C && !CNT && QM == QM_SELF) {
        if (!continuous && t->qround == 0.0) {
            k_trav_v32<VOP, QM, false><<<blocks, 64 * t->grid_waves, 0, st>>>(
                tp, t->topo, t->nodes, t->src, t->pts64, t->pts32);
            CAGR_LAUNCHED();
            return CAGR_OK;
        }
    }
    if (continuous)
        k_trav_v<P, VOP, QM, true, CNT><<<blocks, 64 * t->grid_waves, 0, st>>>(
            tp, t->topo, t->nodes, t->src, t->pts64, t->pts32);
    else
        k_trav_v<P, VOP, QM, false, CNT><<<blocks, 64 * t->grid_waves, 0, st>>>(
            tp, t->topo, t->nodes, t->src, t->pts64, t->pts32);
    CAGR_LAUNCHED();
    return CAGR_OK;
}

template <int P, int QM, bool CNT>
int launch_vop(const TP& tp, cagr_tree* t, int vop, int continuous, int64_t blocks,
               hipStream_t st) {
    switch (vop) {
        case VOP_A: return launch_v<P, VOP_A, QM, CNT>(tp, t, continuous, blocks, st);
        case VOP_G: return launch_v<P, VOP_G, QM, CNT>(tp, t, 0, blocks, st);
        case VOP_A2: return launch_v<P, VOP_A2, QM, CNT>(tp, t, continuous, blocks, st);
        default: return launch_v<P, VOP_A3, QM, CNT>(tp, t, continuous, blocks, st);
    }
}

template <int P, int K, int QM, bool CNT>
int launch_s(const TP& tp, cagr_tree* t, int64_t blocks, hipStream_t st) {
    if constexpr (P == CAGR_PREC_FP32C && !CNT && QM == QM_SELF) {
        if (t->qround == 0.0) {   // the 64-VGPR walk: a tree of FP32 points only
            k_trav_s32<K, QM><<<blocks, 64 * t->grid_waves, 0, st>>>(tp, t->topo, t->nodes, t->src,
                                                                     t->pts64, t->pts32);
            CAGR_LAUNCHED();
            return CAGR_OK;
        }
    }
    k_trav_s<P, K, QM, CNT><<<blocks, 64 * t->grid_waves, 0, st>>>(tp, t->topo, t->nodes, t->src,
                                                                   t->pts64, t->pts32);
    CAGR_LAUNCHED();
    return CAGR_OK;
}

template <int P, bool CNT>
int launch_items(const TP& tp, cagr_tree* t, int op, int continuous, hipStream_t st) {
    Items q{t->items, t->item_order, {}, t->item_counter};
    for (int c = 0; c <= kQueues; ++c) q.bound[c] = t->item_bound[c];
    k_reset_counter<<<1, 64, 0, st>>>(t->item_counter);
    CAGR_LAUNCHED();
    const int threads = 64 * t->waves_per_block;
    // default: enough blocks for 8 waves per SIMD on every CU
    const int64_t blocks = t->grid_blocks > 0 ? t->grid_blocks
                                              : (int64_t)t->num_cus * 32 / t->waves_per_block;
    const auto* topo = t->topo;
    const void* nodes = t->nodes;
    const void* src = t->src;
    // the 64-VGPR walks only for trees of FP32 points (their cutoff retests
    // take the source records' coordinates)
    const bool fp32_tree = t->qround == 0.0;
    switch (op) {
        case CAGR_OP_A:
        case CAGR_OP_A_CPU:
            if (continuous)
                k_trav_v_items<P, VOP_A, true, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            else
                k_trav_v_items<P, VOP_A, false, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        case CAGR_OP_G:
            k_trav_v_items<P, VOP_G, false, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        case CAGR_OP_A21:
        case CAGR_OP_A22:
        case CAGR_OP_A23:
        case CAGR_OP_A2:
            if (continuous)
                k_trav_v_items<P, VOP_A2, true, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            else if (fp32_tree && P == CAGR_PREC_FP32C && !CNT)
                k_trav_v_items32<VOP_A2><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            else
                k_trav_v_items<P, VOP_A2, false, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        case CAGR_OP_A3:
            if (continuous)
                k_trav_v_items<P, VOP_A3, true, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            else
                k_trav_v_items<P, VOP_A3, false, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        case CAGR_OP_AT:
        case CAGR_OP_A2T1:
        case CAGR_OP_A2T2:
        case CAGR_OP_A2T3:
            k_trav_s_items<P, 1, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        case CAGR_OP_A2T_SUM:
            if (fp32_tree && P == CAGR_PREC_FP32C && !CNT)
                k_trav_s_items32<3><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            else
                k_trav_s_items<P, 3, CNT><<<blocks, threads, 0, st>>>(tp, topo, nodes, src, t->pts64, t->pts32, q);
            break;
        default:
            set_error("unknown op");
            return CAGR_ERR_ARG;
    }
    CAGR_LAUNCHED();
    return CAGR_OK;
}

template <int P, int QM, bool CNT>
int launch_all(const TP& tp, cagr_tree* t, int op, int continuous, int64_t blocks,
               hipStream_t st) {
    switch (op) {
        case CAGR_OP_A:
        case CAGR_OP_A_CPU: return launch_vop<P, QM, CNT>(tp, t, VOP_A, continuous, blocks, st);
        case CAGR_OP_G: return launch_vop<P, QM, CNT>(tp, t, VOP_G, 0, blocks, st);
        case CAGR_OP_A21:
        case CAGR_OP_A22:
        case CAGR_OP_A23:
        case CAGR_OP_A2: return launch_vop<P, QM, CNT>(tp, t, VOP_A2, continuous, blocks, st);
        case CAGR_OP_A3: return launch_vop<P, QM, CNT>(tp, t, VOP_A3, continuous, blocks, st);
        case CAGR_OP_AT:
        case CAGR_OP_A2T1:
        case CAGR_OP_A2T2:
        case CAGR_OP_A2T3:
            if constexpr (QM == QM_GRID) return CAGR_ERR_ARG;
            else return launch_s<P, 1, QM, CNT>(tp, t, blocks, st);
        case CAGR_OP_A2T_SUM:
            if constexpr (QM == QM_GRID) return CAGR_ERR_ARG;
            else return launch_s<P, 3, QM, CNT>(tp, t, blocks, st);
    }
    set_error("unknown op");
    return CAGR_ERR_ARG;
}

}  // namespace

int traverse(cagr_tree* t, int op, int precision, const Queries& q, int continuous, double an,
             void* out, int odtype, int64_t* cnt_far, int64_t* cnt_near, int64_t* cnt_visit,
             hipStream_t st) {
    TP tp{};
    tp.root = 0;
    tp.stop = (int32_t)t->m;
    tp.perm = t->perm;
    tp.q = q.q;
    tp.qdtype = q.qdtype;
    tp.qorder = q.qorder;
    tp.widths = q.widths;
    tp.wdtype = q.wdtype;
    tp.width = q.width;
    tp.out = out;
    tp.odtype = odtype;
    tp.an = an;
    tp.sign = (op == CAGR_OP_A_CPU) ? -1.0 : 1.0;
    tp.kdir = -1;
    if (op >= CAGR_OP_A21 && op <= CAGR_OP_A23) tp.kdir = op - CAGR_OP_A21;
    if (op >= CAGR_OP_A2T1 && op <= CAGR_OP_A2T3) tp.kdir = op - CAGR_OP_A2T1;
    if (op == CAGR_OP_AT || op == CAGR_OP_A2T_SUM) tp.kdir = 0;
    if (op == CAGR_OP_AT) tp.an = 0.0;
    tp.cnt_far = cnt_far;
    tp.cnt_near = cnt_near;
    tp.cnt_visit = cnt_visit;
    tp.open_factor = t->open_factor;
    tp.lbound = t->lbound;
    tp.cb = dev::cut_band(t->qround);
    tp.r64 = t->r64;
    tp.nm = t->m;
    int64_t lanes;
    if (q.mode == QM_GRID) {
        for (int d = 0; d < 3; ++d) {
            tp.glo[d] = q.glo[d];
            tp.gstep[d] = q.gstep[d];
            tp.gdim[d] = q.gdim[d];
            tp.gblk[d] = (q.gdim[d] + 3) / 4;
        }
        lanes = tp.gblk[0] * tp.gblk[1] * tp.gblk[2] * 64;
    } else {
        lanes = q.nq;
    }
    tp.nq = lanes;
    if (lanes == 0) return CAGR_OK;
    const int64_t waves = (lanes + 63) / 64;
    int64_t blocks = (waves + t->grid_waves - 1) / t->grid_waves;
    const bool cnt = cnt_far || cnt_near || cnt_visit;
    const bool p64 = precision == CAGR_PREC_FP64;
    if (q.mode == QM_SELF && (t->persistent || t->num_clouds > 1)) {
        if (cnt) return launch_items<CAGR_PREC_FP64, true>(tp, t, op, continuous, st);
        if (p64) return launch_items<CAGR_PREC_FP64, false>(tp, t, op, continuous, st);
        return launch_items<CAGR_PREC_FP32C, false>(tp, t, op, continuous, st);
    }
    CAGR_REQUIRE(t->num_clouds == 1, CAGR_ERR_ARG,
                 "external and grid queries need a single-cloud tree, not a forest");
    int rc;
#define CAGR_LAUNCH_QM(QMV)                                                                   \
    if (cnt) rc = launch_all<CAGR_PREC_FP64, QMV, true>(tp, t, op, continuous, blocks, st);    \
    else if (p64) rc = launch_all<CAGR_PREC_FP64, QMV, false>(tp, t, op, continuous, blocks, st); \
    else rc = launch_all<CAGR_PREC_FP32C, QMV, false>(tp, t, op, continuous, blocks, st);
    if (q.mode == QM_SELF) {
        const bool pieces = t->grid_waves == 1 && !cnt && t->num_pieces > 0;
        if (pieces) {
            tp.pieces = t->pieces;
            tp.planes = t->planes;
            blocks = t->num_pieces;
        } else if (t->grid_waves == 1 && !cnt) {
            tp.tile_order = t->tile_order;
        }
        CAGR_LAUNCH_QM(QM_SELF)
        if (rc == CAGR_OK && pieces && t->num_split > 0) {
            // values per query of the op
            const int no = (op == CAGR_OP_G || op == CAGR_OP_A2 || op == CAGR_OP_AT ||
                            op == CAGR_OP_A2T_SUM || (op >= CAGR_OP_A2T1 && op <= CAGR_OP_A2T3))
                               ? 3 : 1;
            const unsigned cb = (unsigned)((t->num_split + kCombineWaves - 1) / kCombineWaves);
            if (no == 3)
                k_piece_combine<3><<<cb, 64 * kCombineWaves, 0, st>>>(tp, t->split, t->num_split);
            else
                k_piece_combine<1><<<cb, 64 * kCombineWaves, 0, st>>>(tp, t->split, t->num_split);
            CAGR_LAUNCHED();
        }
    } else if (q.mode == QM_EXPLICIT) {
        CAGR_LAUNCH_QM(QM_EXPLICIT)
    } else {
        CAGR_LAUNCH_QM(QM_GRID)
    }
#undef CAGR_LAUNCH_QM
    return rc;
}

void split_walks_free(cagr_tree* t) {
    (void)hipFree(t->pieces);
    (void)hipFree(t->split);
    (void)hipFree(t->planes);
    t->pieces = nullptr;
    t->split = nullptr;
    t->planes = nullptr;
    t->num_pieces = t->num_split = 0;
}

int split_walks(cagr_tree* t, double piece_cost, int max_pieces, hipStream_t st) {
    split_walks_free(t);
    if (!(piece_cost > 0.0) || t->n == 0 || t->num_clouds != 1) return CAGR_OK;
    CAGR_REQUIRE(max_pieces >= 1 && max_pieces <= 64, CAGR_ERR_ARG, "max_pieces must be in [1, 64]");
    const int64_t nw = (t->n + 63) / 64;
    const int32_t m = (int32_t)t->m;
    int32_t *dtot = nullptr, *dnp = nullptr, *dcuts = nullptr;
    struct Free {
        int32_t **a, **b, **c;
        ~Free() { (void)hipFree(*a); (void)hipFree(*b); (void)hipFree(*c); }
    } free_{&dtot, &dnp, &dcuts};
    CAGR_HIP(hipMalloc(&dtot, nw * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&dnp, nw * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&dcuts, nw * max_pieces * sizeof(int32_t)));
    k_walk_cost<<<(unsigned)nw, 64, 0, st>>>(t->topo, t->r64, t->pts64, t->n, m, nullptr, dtot,
                                              nullptr, max_pieces);
    CAGR_LAUNCHED();
    std::vector<int32_t> tot(nw), np(nw, 1), cuts(nw * max_pieces, m);
    CAGR_HIP(hipMemcpyAsync(tot.data(), dtot, nw * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    double mean = 0.0;
    for (int32_t c : tot) mean += c;
    mean /= (double)nw;
    const double target = std::max(1.0, piece_cost * mean);
    int64_t nsplit = 0, slots = 0;
    for (int64_t w = 0; w < nw; ++w) {
        np[w] = (int32_t)std::min<double>(max_pieces, std::max(1.0, std::ceil(tot[w] / target)));
        if (np[w] > 1) {
            ++nsplit;
            slots += np[w];
        }
    }
    if (nsplit == 0) return CAGR_OK;
    CAGR_HIP(hipMemcpyAsync(dnp, np.data(), nw * sizeof(int32_t), hipMemcpyHostToDevice, st));
    k_walk_cost<<<(unsigned)nw, 64, 0, st>>>(t->topo, t->r64, t->pts64, t->n, m, dnp, dtot, dcuts,
                                              max_pieces);
    CAGR_LAUNCHED();
    CAGR_HIP(hipMemcpyAsync(cuts.data(), dcuts, nw * max_pieces * sizeof(int32_t),
                            hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    // pieces in wave (Morton) order, a wave's pieces one after another: the
    // pieces are about equally long, so no longest-first order is needed,
    // and the waves running side by side (each XCD takes a contiguous range
    // of the slots) walk neighbouring queries' node sequences, which the
    // caches share (longest first within each XCD's range measured 8 %
    // slower on the A2 walk at 1M points)
    std::vector<int4> pcs;
    std::vector<int4> split;
    pcs.reserve(nw + slots);
    int32_t slot = 0;
    for (int64_t w = 0; w < nw; ++w) {
        const int P = np[w];
        if (P == 1) {
            pcs.push_back(make_int4((int)w, 0, m, -1));
            continue;
        }
        split.push_back(make_int4((int)w, slot, P, 0));
        for (int j = 0; j < P; ++j) {
            const int32_t lo = j ? cuts[w * max_pieces + j - 1] : 0;
            const int32_t hi = j + 1 < P ? cuts[w * max_pieces + j] : m;
            pcs.push_back(make_int4((int)w, lo, std::max(lo, hi), slot + j));
        }
        slot += P;
    }
    const int64_t G = (int64_t)pcs.size();
    // the tree holds them (and split_walks_free releases them) from here on,
    // counts set only once everything is in place
    CAGR_HIP(hipMalloc(&t->pieces, G * sizeof(int4)));
    CAGR_HIP(hipMalloc(&t->split, nsplit * sizeof(int4)));
    CAGR_HIP(hipMalloc(&t->planes, (size_t)slots * 192 * sizeof(double)));
    CAGR_HIP(hipMemcpyAsync(t->pieces, pcs.data(), G * sizeof(int4), hipMemcpyHostToDevice, st));
    CAGR_HIP(hipMemcpyAsync(t->split, split.data(), nsplit * sizeof(int4), hipMemcpyHostToDevice, st));
    CAGR_HIP(hipStreamSynchronize(st));
    t->num_pieces = G;
    t->num_split = nsplit;
    return CAGR_OK;
}

}  // namespace cagr
