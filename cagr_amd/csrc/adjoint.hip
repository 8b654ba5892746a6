// Exact adjoint of the linear treecode operator (gfx950).
//
// With node centres fixed by caller-given weights (cagr_apply_weighted), the
// treecode y = A x of the vector-family operators (A / A_cpu, A2k, A3, fused
// A2) is linear in x:
//   y_i = sum_{S far for i} K(x_i - c_S) . X_S + sum_{j in near leaves of i} K(x_i - y_j) . x_j,
//   X_S = sum_{j in S} x_j.
// Its exact adjoint is
//   (A^T g)_j = N_j + sum_{S containing j} L_S,
//   L_S = sum_{i: S far for i} g_i K(x_i - c_S),   N_j = sum_{i: leaf(j) near for i} g_i K(x_i - y_j),
// computed here with the forward traversal itself: the same wave-coherent
// walk (same decisions, node for node), but every interaction is scattered
// back to its source -- far nodes through a wave reduction and one FP64 atomic
// per node visit into L, near leaves transposed (lanes hold the leaf's
// sources, the wave's near queries are broadcast one by one) with one FP64
// atomic per source and visit into N -- then a downward sweep adds the L of
// every node on a point's root path.  Nothing here needs the reference's
// A2Tk: that is the adjoint of the direction-reversed operator, not of A2k
// (A2Tk(An) = (-A2k(-An))^T).
#include "common.h"
#include "trav_common.h"
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

namespace cagr {
namespace {

using namespace dev;
constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;

enum { GM_A = 0, GM_A2K = 1, GM_A3 = 2, GM_A2 = 3 };
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt at their maxima, gfx9 encoding):
// a list chunk's loads are complete before its steps, so the steps' loop
// carries no wait for them (nor for the steps' own stores)
constexpr int kWaitVm0 = 0x0F70;
// modes of the FP32 walk: the plain adjoint (far sums as FP64 atomics into L),
// and the two walks that build the adjoint plan (count, then record)
enum { FM_ATOMIC = 0, FM_COUNT = 2, FM_FILL = 3 };

struct AP {
    int64_t n;
    int32_t root, stop;
    const int32_t* perm;
    const int32_t* tile_order;// dispatch slot -> wave (null: identity); one wave per block
    const void* g;            // cotangent [n, gcols], original order
    int gdtype;
    int gmode;                // GM_*
    int kdir;                 // GM_A2K: direction
    const void* widths;
    int wdtype;
    double an, sign;
    double open_factor;       // 2 / theta
    dev::CutBand cb;          // band of the FP32 near-pair cutoff test
    const double* c64;        // FP64 centres of the node records, [m, 3] (retests of the
                              // FP32 opening and cutoff tests)
    double* adj_node;         // [m,3] accumulators L
    double* adj_pt;           // [3][n] accumulators N (sorted order, one plane per component,
                              // so a wave's atomics on one component touch consecutive words)
    // plan (fixed far-step sequence per wave)
    const int32_t* wave_off;  // [waves + 1] first slot of each wave
    const int32_t* slot_pos;  // [slots] node-major position of each slot
    int32_t* wave_cnt;        // [waves] far steps per wave (counting walk)
    int32_t* node_cnt;        // [m] far steps per node (counting walk)
    uint64_t* slot_node;      // [slots] node of each slot (filling walk)
    int32_t* slot_id;         // [slots] the slot's own index (filling walk)
    float4* far_val;          // [slots] wave sums of the far steps
    const int32_t* pieces;    // replay pieces, 8 ints each (null: whole waves)
    // the walk itself, recorded (filling walk) and replayed (list kernels)
    int32_t* far_node;        // [slots] node of each far step, wave-major
    uint64_t* far_mask;       // [slots] lanes that take it
    const int32_t* near_off;  // [waves + 1] first near-leaf entry of each wave
    int32_t* near_cnt;        // [waves] near-leaf steps per wave (counting walk)
    int32_t* near_leaf;       // [entries] leaf of each near step
    uint64_t* near_mask;      // [entries] lanes near to it
};

// the wave (64 consecutive sorted points) this block serves: blocks of one
// wave, dispatched XCD-aware, in the tree's longest-first order when it has one
__device__ inline int64_t wave_of(const AP& p) {
    // a scalar: the order entry is a scalar load and the wave index stays in
    // a scalar register
    const int32_t slot = __builtin_amdgcn_readfirstlane(
        (int32_t)(xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6)));
    return p.tile_order ? (int64_t)p.tile_order[slot] : (int64_t)slot;
}

// cotangent of query oi as the per-direction weights of the k-sum form
// (GM_A: scalar in gk[0])
__device__ inline void load_cot(const AP& p, int64_t oi, double gk[3]) {
    if (p.gmode == GM_A2) {
        for (int k = 0; k < 3; ++k) gk[k] = load_as_double(p.g, p.gdtype, 3 * oi + k);
    } else {
        const double v = load_as_double(p.g, p.gdtype, oi);
        if (p.gmode == GM_A) {
            gk[0] = v; gk[1] = 0.0; gk[2] = 0.0;
        } else if (p.gmode == GM_A3) {
            gk[0] = gk[1] = gk[2] = v / 3.0;
        } else {
            for (int k = 0; k < 3; ++k) gk[k] = (k == p.kdir) ? v : 0.0;
        }
    }
}

// ---------------------------------------------------------------- FP32c
// interaction of (query with factors h, width state) with a source at d:
// adds the contribution to the source's adjoint accumulator z.  Lanes with
// use = false add exact zeros, so a wave can run it unmasked.
template <bool GA, bool CONT>
__device__ inline void adj_term32(float* z, const float* d, float r2, bool use, bool in,
                                  const float* h, const float* e, float hanf, const float* gin,
                                  float winv, float ew) {
    const float ir = (use && !in) ? __builtin_amdgcn_rsqf(r2) : 0.f;
    const float ir3 = ir * ir * ir;
    const bool ci = use && in;
    if constexpr (GA) {
        const float u = h[0] * ir3;
        for (int c = 0; c < 3; ++c) z[c] = fmaf(u, d[c], z[c]);
        if constexpr (CONT) {
            const float v = ci ? gin[0] * winv * winv * winv : 0.f;
            for (int c = 0; c < 3; ++c) z[c] = fmaf(v, d[c], z[c]);
        }
    } else {
        const float a0 = h[0] * e[0], a1 = h[1] * e[1], a2 = h[2] * e[2];
        const float gi3 = ((a0 + a1) + a2) * ir3;
        const float hir = hanf * ir;
        z[0] = fmaf(gi3, d[0], fmaf(a0, hir, z[0]));
        z[1] = fmaf(gi3, d[1], fmaf(a1, hir, z[1]));
        z[2] = fmaf(gi3, d[2], fmaf(a2, hir, z[2]));
        if constexpr (CONT) {
            const float gs = (gin[0] + gin[1]) + gin[2];
            const float v = ci ? ew * gs * winv * winv * winv : 0.f;
            const float hw = ci ? ew * hanf * winv : 0.f;
            for (int c = 0; c < 3; ++c) z[c] = fmaf(v, d[c], fmaf(hw, gin[c], z[c]));
        }
    }
}

// Group layout of a transposed near block (lanes hold a leaf's sources):
// groups of 2^lg lanes, the smallest power of two holding the leaf's points
// (64 lanes and chunks of 64 points for leaves of more than 32), and the
// 64 >> lg groups take consecutive near queries of the wave, so a leaf of 14
// points serves 4 queries per step.  inv: lane r < nq holds the lane of the
// r-th near query (a permutation sent through ds_permute).  All 64 lanes must
// be active.
struct NearGroups {
    int lg, grp, nq, inv;
};

__device__ inline NearGroups near_groups(uint64_t qmask, int32_t npts, int lane) {
    NearGroups g;
    g.lg = 0;
    while ((1 << g.lg) < npts && g.lg < 6) ++g.lg;
    g.grp = lane >> g.lg;
    g.nq = __popcll(qmask);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi(
        (uint32_t)(qmask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)qmask, 0u));
    const bool isq = (qmask >> lane) & 1;
    g.inv = __builtin_amdgcn_ds_permute((isq ? rank : g.nq + (lane - rank)) << 2, lane);
    return g;
}

// lane of the near query this lane's group takes in the step starting at
// rank b0 (ok = false past the last query)
__device__ inline int near_query(const NearGroups& g, int b0, bool& ok) {
    const int r = b0 + g.grp;
    ok = r < g.nq;
    return __builtin_amdgcn_ds_bpermute((ok ? r : 0) << 2, g.inv);
}

// The transposed near block of one leaf for the wave's near lanes qmask
// (group layout above); the groups' sums meet in every group (group_sum).
// Y64: the FP64 cutoff retests read the tree's FP64 points (false: the
// records' float coordinates, exact only for a tree of FP32 points, and
// fewer registers at the retest).
// One FP64 atomic per source and component into N.  All 64 lanes must be
// active.
template <bool GA, bool CONT, bool Y64>
__device__ inline void adj_near_leaf(const AP& p, const TopoRec& t, uint64_t qmask, int lane,
                                     const SrcV32* __restrict__ src,
                                     const double* __restrict__ pts64, const double* x,
                                     const float* xf, const float* h, const float* gin, float w2,
                                     double w, float winv, float ew, float hanf) {
    const NearGroups ng = near_groups(qmask, t.end - t.begin, lane);
    const int S = 1 << ng.lg, G = 64 >> ng.lg;
    for (int32_t base = t.begin; base < t.end; base += S) {
        const int32_t j = base + (lane & (S - 1));
        const bool has = j < t.end;
        const SrcV32 sr = src[has ? j : t.begin];   // lanes past the leaf read a real record
        float zz[3] = {0.f, 0.f, 0.f};
        for (int b0 = 0; b0 < ng.nq; b0 += G) {
            bool qok;
            const int qsrc = near_query(ng, b0, qok);
            // the group's query values, fetched through the LDS crossbar
            auto pick = [&](auto v) { return __shfl(v, qsrc, 64); };
            const float qx[3] = {pick(xf[0]), pick(xf[1]), pick(xf[2])};
            const float qh[3] = {pick(h[0]), pick(h[1]), pick(h[2])};
            const float qw2 = pick(w2);
            float qg[3] = {0.f, 0.f, 0.f}, qwinv = 0.f, qew = 0.f;
            if constexpr (CONT) {
                qg[0] = pick(gin[0]);
                qg[1] = pick(gin[1]);
                qg[2] = pick(gin[2]);
                qwinv = pick(winv);
                qew = pick(ew);
            }
            const bool use = has && qok;
            const float d[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1], qx[2] - sr.y[2]};
            const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
            const bool amb = use && fabsf(r2 - qw2) <= fmaf(p.cb.nb1, qw2, p.cb.nb0);
            bool in = r2 < qw2;
            if (__builtin_expect(__ballot(amb) != 0, 0)) {
                const double qx64[3] = {pick(x[0]), pick(x[1]), pick(x[2])};
                const double qw64 = pick(w);
                double y[3];
                if constexpr (Y64) {
                    const double* yp = pts64 + 3 * (int64_t)(has ? j : t.begin);
                    for (int k = 0; k < 3; ++k) y[k] = yp[k];
                } else {
                    for (int k = 0; k < 3; ++k) y[k] = (double)sr.y[k];
                }
                if (amb)
                    in = !(sqrt(dist2_rn(qx64[0] - y[0], qx64[1] - y[1], qx64[2] - y[2])) >= qw64);
            }
            adj_term32<GA, CONT>(zz, d, r2, use, in, qh, sr.e, hanf, qg, qwinv, qew);
        }
        for (int k = 0; k < 3; ++k) zz[k] = group_sum(zz[k], S);
        if (has && ng.grp == 0) {
            atomicAdd(p.adj_pt + j, (double)zz[0]);
            atomicAdd(p.adj_pt + p.n + j, (double)zz[1]);
            atomicAdd(p.adj_pt + 2 * p.n + j, (double)zz[2]);
        }
    }
}

template <int GMODE_A, bool CONT, int FM, bool Y64 = false>
__global__ __launch_bounds__(kThreads) void k_adj32(AP p, const TopoRec* __restrict__ topo,
                                                    const NodeV32* __restrict__ nodes,
                                                    const SrcV32* __restrict__ src,
                                                    const double* __restrict__ pts64,
                                                    const float* __restrict__ pts32) {
    constexpr bool GA = GMODE_A != 0;
    const int lane = threadIdx.x & 63;
    const int64_t wave = wave_of(p);
    const int64_t s = wave * 64 + lane;
    const bool valid = s < p.n;
    if (__ballot(valid) == 0) return;
    const int64_t sc = valid ? s : 0;
    const double x[3] = {pts64[3 * sc], pts64[3 * sc + 1], pts64[3 * sc + 2]};
    const float xf[3] = {pts32[3 * sc], pts32[3 * sc + 1], pts32[3 * sc + 2]};
    constexpr bool MATH = FM == FM_ATOMIC;   // the walks of the plan do no math
    const int64_t oi = p.perm[sc];
    const double w = MATH ? load_as_double(p.widths, p.wdtype, oi) : 1.0;
    double gk[3] = {0.0, 0.0, 0.0};
    if constexpr (MATH) load_cot(p, oi, gk);
    if (!valid) gk[0] = gk[1] = gk[2] = 0.0;
    float h[3], gin[3];
    if constexpr (GA) {
        h[0] = (float)(p.sign * gk[0]);
        h[1] = h[2] = 0.f;
        gin[0] = h[0];
        gin[1] = gin[2] = 0.f;
    } else {
        for (int k = 0; k < 3; ++k) {
            h[k] = (float)(gk[k] * exp(0.5 * p.an * x[k]));
            gin[k] = (float)gk[k];
        }
    }
    const float wf = (float)w, w2 = wf * wf, winv = 1.f / wf;
    const float ew = CONT ? (float)exp(p.an * w) : 0.f;
    const float hanf = (float)(0.5 * p.an);
    int32_t resume = valid ? p.root : INT_MAX;
    const int dsafe = __builtin_amdgcn_readfirstlane(
        far_safe_depth(p.open_factor, wave_max(valid ? w : 0.0)));
    // the wave's slots (plan): [slot0, slot0 + kmax)
    int64_t slot0 = 0;
    int32_t kmax = 0;
    if constexpr (FM == FM_FILL) {
        slot0 = p.wave_off[wave];
        kmax = p.wave_off[wave + 1] - (int32_t)slot0;
    }
    int32_t k = 0;   // far steps so far
    int64_t near0 = 0;
    int32_t knmax = 0, kn = 0;   // near-leaf steps (plan walks)
    if constexpr (FM == FM_FILL) {
        near0 = p.near_off[wave];
        knmax = p.near_off[wave + 1] - (int32_t)near0;
    }

    // every branch is wave-uniform; lanes that do not take a step add zeros
    for (int32_t n = p.root; n < p.stop;) {
        n = __builtin_amdgcn_readfirstlane(n);
        const TopoRec t = rec32(topo, n);
        const NodeV32 nr = rec32(nodes, n);
        // the step's records loaded together and waited for once
        CAGR_LOADS_HERE("s"(t.escape), "s"(t.info), "s"(nr.cf[0]), "s"(nr.cf[1]), "s"(nr.cf[2]),
                        "s"(nr.hi2), "s"(nr.lo2), "s"(nr.e[0]), "s"(nr.e[1]), "s"(nr.e[2]));
        __builtin_amdgcn_sched_barrier(0);
        // the forward walk's decisions: FP32 distance to the float centre
        // against the record's bounds, the FP64 test in between
        const float d[3] = {xf[0] - nr.cf[0], xf[1] - nr.cf[1], xf[2] - nr.cf[2]};
        const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
        const uint64_t am = __builtin_amdgcn_ballot_w64(resume <= n);
        const uint64_t fm = __builtin_amdgcn_ballot_w64(r2 > nr.hi2);
        const uint64_t lm = __builtin_amdgcn_ballot_w64(r2 < nr.lo2);
        uint64_t far_m = am & fm, near_m = am & lm;
        if (const uint64_t amb = am & ~(fm | lm); __builtin_expect(amb != 0, 0)) {
            const bool ak = __builtin_amdgcn_inverse_ballot_w64(amb);
            bool far64 = false;
            if (ak) {
                const double* c64 = p.c64 + 3 * (int64_t)n;
                far64 = dist2_rn(x[0] - c64[0], x[1] - c64[1], x[2] - c64[2]) > t.thr2;
            }
            far_m |= __builtin_amdgcn_ballot_w64(ak & far64);
            near_m |= __builtin_amdgcn_ballot_w64(ak & !far64);
        }
        const bool takefar = __builtin_amdgcn_inverse_ballot_w64(far_m);
        if (far_m != 0) {
            if constexpr (FM == FM_COUNT) {
                if (lane == 0) atomicAdd(p.node_cnt + n, 1);
            } else if constexpr (FM == FM_FILL) {
                if (lane == 0 && k < kmax) {
                    p.slot_node[slot0 + k] = (uint64_t)n;
                    p.slot_id[slot0 + k] = (int32_t)(slot0 + k);
                    p.far_node[slot0 + k] = n;
                    p.far_mask[slot0 + k] = far_m;
                }
            } else {
                bool in = false;   // far at depth <= dsafe proves r > w
                if (topo_depth(t) > dsafe)
                    in = inside_centre(r2, w2, wf, nr.ce, x, p.c64 + 3 * (int64_t)n, w, takefar);
                float z[3] = {0.f, 0.f, 0.f};
                adj_term32<GA, CONT>(z, d, r2, takefar, in, h, nr.e, hanf, gin, winv, ew);
                float s0, s1, s2;
                wave_sum3(z[0], z[1], z[2], s0, s1, s2);
                if (lane == 0) {
                    atomicAdd(p.adj_node + 3 * (int64_t)n, (double)s0);
                    atomicAdd(p.adj_node + 3 * (int64_t)n + 1, (double)s1);
                    atomicAdd(p.adj_node + 3 * (int64_t)n + 2, (double)s2);
                }
            }
            ++k;
        }
        const uint64_t qmask = near_m;
        // past a leaf or a subtree no lane enters, the escape (every lane
        // active at n stays active there); into the first child with the far
        // lanes parked past the subtree
        if (topo_leaf(t)) {
            if (MATH && qmask != 0)
                adj_near_leaf<GA, CONT, Y64>(p, t, qmask, lane, src, pts64, x, xf, h, gin, w2, w, winv, ew, hanf);
            if constexpr (FM == FM_COUNT || FM == FM_FILL) {
                if (qmask != 0) {
                    if constexpr (FM == FM_FILL) {
                        if (lane == 0 && kn < knmax) {
                            p.near_leaf[near0 + kn] = n;
                            p.near_mask[near0 + kn] = qmask;
                        }
                    }
                    ++kn;
                }
            }
            n = t.escape;
        } else if (qmask != 0) {
            if (takefar) resume = t.escape;
            n = n + 1;
        } else {
            n = t.escape;
        }
    }
    if constexpr (FM == FM_COUNT)
        if (lane == 0) {
            p.wave_cnt[wave] = k;
            p.near_cnt[wave] = kn;
        }
}

// The planned adjoint: the recorded walk replayed.  Far steps come from the
// wave's list (node, lanes) and need no opening test; their wave sums go to
// the slots; near leaves come from the wave's near list.  With pieces, a
// wave's lists are cut into ranges replayed by separate waves (the slots
// and the near atomics do not depend on which wave takes a step).
template <int GMODE_A, bool CONT, bool Y64>
__global__ __launch_bounds__(kThreads) void k_adj32_list(AP p, const TopoRec* __restrict__ topo,
                                                         const NodeV32* __restrict__ nodes,
                                                         const SrcV32* __restrict__ src,
                                                         const double* __restrict__ pts64,
                                                         const float* __restrict__ pts32) {
    constexpr bool GA = GMODE_A != 0;
    const int lane = threadIdx.x & 63;
    int64_t wave;
    int32_t far_lo, far_hi, near_lo, near_hi;
    if (p.pieces) {
        const int32_t slot = __builtin_amdgcn_readfirstlane(
            (int32_t)(xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6)));
        const int32_t* pc = p.pieces + 8 * (int64_t)slot;
        wave = pc[0];
        far_lo = pc[1];
        far_hi = pc[2];
        near_lo = pc[3];
        near_hi = pc[4];
    } else {
        wave = __builtin_amdgcn_readfirstlane((int)wave_of(p));
        far_lo = __builtin_amdgcn_readfirstlane(p.wave_off[wave]);
        far_hi = __builtin_amdgcn_readfirstlane(p.wave_off[wave + 1]);
        near_lo = __builtin_amdgcn_readfirstlane(p.near_off[wave]);
        near_hi = __builtin_amdgcn_readfirstlane(p.near_off[wave + 1]);
    }
    const int64_t s = wave * 64 + lane;
    const bool valid = s < p.n;
    if (__ballot(valid) == 0) return;
    const int64_t sc = valid ? s : 0;
    const double x[3] = {pts64[3 * sc], pts64[3 * sc + 1], pts64[3 * sc + 2]};
    const float xf[3] = {pts32[3 * sc], pts32[3 * sc + 1], pts32[3 * sc + 2]};
    const int64_t oi = p.perm[sc];
    const double w = load_as_double(p.widths, p.wdtype, oi);
    double gk[3];
    load_cot(p, oi, gk);
    if (!valid) gk[0] = gk[1] = gk[2] = 0.0;
    float h[3], gin[3];
    if constexpr (GA) {
        h[0] = (float)(p.sign * gk[0]);
        h[1] = h[2] = 0.f;
        gin[0] = h[0];
        gin[1] = gin[2] = 0.f;
    } else {
        for (int c = 0; c < 3; ++c) {
            h[c] = (float)(gk[c] * exp(0.5 * p.an * x[c]));
            gin[c] = (float)gk[c];
        }
    }
    const float wf = (float)w, w2 = wf * wf, winv = 1.f / wf;
    const float ew = CONT ? (float)exp(p.an * w) : 0.f;
    const float hanf = (float)(0.5 * p.an);
    const int dsafe = __builtin_amdgcn_readfirstlane(
        far_safe_depth(p.open_factor, wave_max(valid ? w : 0.0)));

    // the lists are read 64 entries at a time, one per lane (coalesced), and
    // entry k of a chunk is taken from lane k: no dependent load in front of
    // each step.  The chunk's loads are waited for once, in front of its
    // steps; a step's slot store is not waited for (a load issued inside the
    // steps' loop would make every step wait for all outstanding memory
    // operations, the previous step's store among them)
    const int32_t slot0 = far_lo;
    const int32_t nf = far_hi - far_lo;
    for (int32_t k0 = 0; k0 < nf; k0 += 64) {
        const int32_t kl = k0 + lane;
        const int32_t node_l = kl < nf ? p.far_node[slot0 + kl] : 0;
        const uint64_t mask_l = kl < nf ? p.far_mask[slot0 + kl] : 0;
        const int32_t pos_l = kl < nf ? p.slot_pos[slot0 + kl] : 0;
        const int32_t kn = min(64, nf - k0);
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        for (int32_t k = 0; k < kn; ++k) {
            const int32_t n = __builtin_amdgcn_readlane(node_l, k);
            const uint64_t m = lane_bcast_u64(mask_l, k);
            const int32_t pos = __builtin_amdgcn_readlane(pos_l, k);
            const bool take = __builtin_amdgcn_inverse_ballot_w64(m);
            const TopoRec t = rec32(topo, n);
            const NodeV32 nr = rec32(nodes, n);
            CAGR_LOADS_HERE("s"(t.info), "s"(nr.cf[0]), "s"(nr.cf[1]), "s"(nr.cf[2]), "s"(nr.e[0]),
                            "s"(nr.e[1]), "s"(nr.e[2]));
            __builtin_amdgcn_sched_barrier(0);
            const float d[3] = {xf[0] - nr.cf[0], xf[1] - nr.cf[1], xf[2] - nr.cf[2]};
            const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
            bool in = false;   // far at depth <= dsafe proves r > w
            if (topo_depth(t) > dsafe)
                in = inside_centre(r2, w2, wf, nr.ce, x, p.c64 + 3 * (int64_t)n, w, take);
            float z[3] = {0.f, 0.f, 0.f};
            adj_term32<GA, CONT>(z, d, r2, take, in, h, nr.e, hanf, gin, winv, ew);
            // the three totals stay in their rows (row 0: z0, 1: z2, 2: z1);
            // the first lane of each row stores its component of the slot
            const float tot = wave_sum3_rows(z[0], z[1], z[2]);
            if ((lane & 15) == 0 && lane < 48)
                ((float*)p.far_val)[4 * (int64_t)pos + (lane == 0 ? 0 : (lane == 16 ? 2 : 1))] = tot;
        }
    }
    const int32_t near0 = near_lo;
    const int32_t nn = near_hi - near_lo;
    for (int32_t j0 = 0; j0 < nn; j0 += 64) {
        const int32_t jl = j0 + lane;
        const int32_t leaf_l = jl < nn ? p.near_leaf[near0 + jl] : 0;
        const uint64_t nmask_l = jl < nn ? p.near_mask[near0 + jl] : 0;
        const int32_t jn = min(64, nn - j0);
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        for (int32_t j = 0; j < jn; ++j) {
            const int32_t n = __builtin_amdgcn_readlane(leaf_l, j);
            const uint64_t qmask = lane_bcast_u64(nmask_l, j);
            const TopoRec t = rec32(topo, n);
            adj_near_leaf<GA, CONT, Y64>(p, t, qmask, lane, src, pts64, x, xf, h, gin, w2, w, winv, ew, hanf);
        }
    }
}

// L of every node from its run of slots, one wave per node: lanes stride the
// run (coalesced), then a fixed-order wave reduction in FP64, so the planned
// adjoint is deterministic.  Waves take the nodes in depth order (order[],
// shallowest first): the top nodes, whose runs hold a slot for nearly every
// wave of the walk, start first instead of wherever their preorder ids fall.
__global__ __launch_bounds__(kThreads) void k_far_gather(int64_t m, const int32_t* __restrict__ order,
                                                         const int32_t* __restrict__ node_off,
                                                         const float4* __restrict__ far_val,
                                                         double* __restrict__ adj_node) {
    const int64_t k = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (k >= m) return;   // whole waves
    const int64_t nd = order[k];
    const int lane = threadIdx.x & 63;
    double z0 = 0.0, z1 = 0.0, z2 = 0.0;
    const int32_t lo = node_off[nd], hi = node_off[nd + 1];
    // B of a lane's slots loaded before any is added (one memory round trip
    // per 64 B slots of the run), added in the lane's slot order whatever B:
    // 16 for long runs (the top nodes have a slot for nearly every wave), 4
    // for the rest
    auto batch = [&](int32_t i0, auto bc) {
        constexpr int B = decltype(bc)::value;
        float4 v[B];
#pragma unroll
        for (int u = 0; u < B; ++u)
            v[u] = i0 + 64 * u < hi ? far_val[i0 + 64 * u] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (i0 + 64 * u < hi) {
                z0 += (double)v[u].x;
                z1 += (double)v[u].y;
                z2 += (double)v[u].z;
            }
        }
    };
    int32_t i0 = lo + lane;
    for (; i0 + 64 * 15 < hi; i0 += 64 * 16) batch(i0, std::integral_constant<int, 16>{});
    for (; i0 < hi; i0 += 256) batch(i0, std::integral_constant<int, 4>{});
    wave_sum3(z0, z1, z2, z0, z1, z2);
    if (lane == 0) {
        adj_node[3 * nd] = z0;
        adj_node[3 * nd + 1] = z1;
        adj_node[3 * nd + 2] = z2;
    }
}

__global__ __launch_bounds__(kThreads) void k_invert(const int32_t* __restrict__ perm, int64_t len,
                                                     int32_t* __restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < len) pos[perm[i]] = (int32_t)i;
}

__global__ __launch_bounds__(kThreads) void k_to_f64(const void* __restrict__ a, int dtype,
                                                     int64_t len, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < len) out[i] = load_as_double(a, dtype, i);
}

__global__ __launch_bounds__(kThreads) void k_fill0_i32(int32_t* __restrict__ a, int64_t len) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < len) a[i] = 0;
}

// ---------------------------------------------------------------- FP64
template <bool GA, bool CONT>
__device__ inline void adj_term64(double* z, const double* d, double w, const double* gk,
                                  double an, double sign) {
    const double r2 = dist2_rn(d[0], d[1], d[2]);
    const double r = sqrt(r2);
    if (r >= w) {
        const double ir = 1.0 / r;
        const double ir3 = ir * ir * ir;
        if constexpr (GA) {
            const double u = sign * gk[0] * ir3;
            for (int c = 0; c < 3; ++c) z[c] += u * d[c];
        } else {
            const double a0 = gk[0] * exp(0.5 * an * d[0]);
            const double a1 = gk[1] * exp(0.5 * an * d[1]);
            const double a2 = gk[2] * exp(0.5 * an * d[2]);
            const double gi3 = ((a0 + a1) + a2) * ir3;
            const double hir = 0.5 * an * ir;
            z[0] += gi3 * d[0] + a0 * hir;
            z[1] += gi3 * d[1] + a1 * hir;
            z[2] += gi3 * d[2] + a2 * hir;
        }
    } else if constexpr (CONT) {
        const double w3 = w * w * w;
        if constexpr (GA) {
            const double u = sign * gk[0] / w3;
            for (int c = 0; c < 3; ++c) z[c] += u * d[c];
        } else {
            const double ew = exp(an * w);
            const double gs = (gk[0] + gk[1]) + gk[2];
            for (int c = 0; c < 3; ++c) z[c] += ew * (gs * d[c] / w3 + 0.5 * an * gk[c] / w);
        }
    }
}

template <int GMODE_A, bool CONT>
__global__ __launch_bounds__(kThreads) void k_adj64(AP p, const TopoRec* __restrict__ topo,
                                                    const NodeV64* __restrict__ nodes,
                                                    const SrcV64* __restrict__ src,
                                                    const double* __restrict__ pts64) {
    constexpr bool GA = GMODE_A != 0;
    const int lane = threadIdx.x & 63;
    const int64_t wave = wave_of(p);
    const int64_t s = wave * 64 + lane;
    const bool valid = s < p.n;
    if (__ballot(valid) == 0) return;
    const int64_t sc = valid ? s : 0;
    const double x[3] = {pts64[3 * sc], pts64[3 * sc + 1], pts64[3 * sc + 2]};
    const int64_t oi = p.perm[sc];
    const double w = load_as_double(p.widths, p.wdtype, oi);
    double gk[3];
    load_cot(p, oi, gk);
    if (!valid) gk[0] = gk[1] = gk[2] = 0.0;
    int32_t resume = valid ? p.root : INT_MAX;

    for (int32_t n = p.root; n < p.stop;) {
        n = __builtin_amdgcn_readfirstlane(n);
        const TopoRec t = rec32(topo, n);
        const NodeV64 nr = rec32(nodes, n);
        CAGR_LOADS_HERE("s"(t.thr2), "s"(t.escape), "s"(t.info), "s"(nr.c[0]), "s"(nr.c[1]),
                        "s"(nr.c[2]));
        __builtin_amdgcn_sched_barrier(0);
        const bool act = resume <= n;
        const double d[3] = {x[0] - nr.c[0], x[1] - nr.c[1], x[2] - nr.c[2]};
        const bool far = dist2_rn(d[0], d[1], d[2]) > t.thr2;
        const bool takefar = act && far;
        double z[3] = {0.0, 0.0, 0.0};
        if (takefar) adj_term64<GA, CONT>(z, d, w, gk, p.an, p.sign);
        if (__ballot(takefar) != 0) {
            double s0, s1, s2;
            wave_sum3(z[0], z[1], z[2], s0, s1, s2);
            if (lane == 0) {
                atomicAdd(p.adj_node + 3 * (int64_t)n, s0);
                atomicAdd(p.adj_node + 3 * (int64_t)n + 1, s1);
                atomicAdd(p.adj_node + 3 * (int64_t)n + 2, s2);
            }
        }
        const bool nearl = act && !far;
        // resume moves only into a subtree (as in the FP32c walk)
        if (topo_leaf(t)) {
            const uint64_t qmask = __ballot(nearl);
            if (qmask != 0) {   // transposed, in the group layout of near_groups
                const NearGroups ng = near_groups(qmask, t.end - t.begin, lane);
                const int S = 1 << ng.lg, G = 64 >> ng.lg;
                for (int32_t base = t.begin; base < t.end; base += S) {
                    const int32_t j = base + (lane & (S - 1));
                    const bool has = j < t.end;
                    const SrcV64 sr = src[has ? j : t.begin];
                    double zz[3] = {0.0, 0.0, 0.0};
                    for (int b0 = 0; b0 < ng.nq; b0 += G) {
                        bool qok;
                        const int qsrc = near_query(ng, b0, qok);
                        const double qx[3] = {lane_pull(x[0], qsrc), lane_pull(x[1], qsrc),
                                              lane_pull(x[2], qsrc)};
                        const double qg[3] = {lane_pull(gk[0], qsrc), lane_pull(gk[1], qsrc),
                                              lane_pull(gk[2], qsrc)};
                        const double qw = lane_pull(w, qsrc);
                        if (has && qok) {
                            const double dd[3] = {qx[0] - sr.y[0], qx[1] - sr.y[1], qx[2] - sr.y[2]};
                            adj_term64<GA, CONT>(zz, dd, qw, qg, p.an, p.sign);
                        }
                    }
                    for (int sh = S; sh < 64; sh <<= 1)
                        for (int k = 0; k < 3; ++k) zz[k] += __shfl_xor(zz[k], sh, 64);
                    if (has && ng.grp == 0) {
                        atomicAdd(p.adj_pt + j, zz[0]);
                        atomicAdd(p.adj_pt + p.n + j, zz[1]);
                        atomicAdd(p.adj_pt + 2 * p.n + j, zz[2]);
                    }
                }
            }
            n = t.escape;
        } else if (__ballot(nearl) != 0) {
            if (takefar) resume = t.escape;
            n = n + 1;
        } else {
            n = t.escape;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_fill0(double* __restrict__ a, int64_t len) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < len) a[i] = 0.0;
}

// out[perm[s]] = N_s + sum of L over the nodes containing s
template <typename TO>
__global__ __launch_bounds__(kThreads) void k_adj_down(int64_t n, const int32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ leaf_of,
                                                       const TopoRec* __restrict__ topo,
                                                       const double* __restrict__ adj_node,
                                                       const double* __restrict__ adj_pt,
                                                       TO* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    double z[3] = {adj_pt[s], adj_pt[n + s], adj_pt[2 * n + s]};
    for (int32_t nd = leaf_of[s]; nd >= 0; nd = topo[nd].parent)
        for (int c = 0; c < 3; ++c) z[c] += adj_node[3 * (int64_t)nd + c];
    const int64_t o = perm[s];
    for (int c = 0; c < 3; ++c) out[3 * o + c] = (TO)z[c];
}

}  // namespace

int adjoint(cagr_tree* t, int op, int precision, const void* g, int dtype, const void* widths,
            int continuous, double an, void* out, hipStream_t st, bool planned,
            const void* nodes_in, const void* src_in, const double* c64_in) {
    if (!nodes_in) nodes_in = t->nodes;
    if (!src_in) src_in = t->src;
    AP p{};
    p.c64 = c64_in ? c64_in : t->r64;
    p.n = t->n;
    p.root = 0;
    p.stop = (int32_t)t->m;
    p.perm = t->perm;
    p.g = g;
    p.gdtype = dtype;
    p.widths = widths;
    p.wdtype = dtype;
    p.an = an;
    p.sign = op == CAGR_OP_A_CPU ? -1.0 : 1.0;
    p.open_factor = t->open_factor;
    p.cb = dev::cut_band(t->qround);
    p.adj_node = t->adj_node;
    p.adj_pt = t->adj_pt;
    p.kdir = 0;
    if (op == CAGR_OP_A || op == CAGR_OP_A_CPU) p.gmode = GM_A;
    else if (op == CAGR_OP_A3) p.gmode = GM_A3;
    else if (op == CAGR_OP_A2) p.gmode = GM_A2;
    else { p.gmode = GM_A2K; p.kdir = op - CAGR_OP_A21; }
    const bool ga = p.gmode == GM_A;
    const int64_t n = t->n, m = t->m;
    const bool slots = planned && precision == CAGR_PREC_FP32C;
    CAGR_REQUIRE(!planned || t->plan_slots >= 0, CAGR_ERR_STATE, "no adjoint plan on this tree");
    if (slots) {
        p.pieces = t->plan_pieces;
        p.wave_off = t->plan_wave_off;
        p.slot_pos = t->plan_slot_pos;
        p.far_val = t->plan_far_val;
        p.far_node = t->plan_far_node;
        p.far_mask = t->plan_far_mask;
        p.near_off = t->plan_near_off;
        p.near_leaf = t->plan_near_leaf;
        p.near_mask = t->plan_near_mask;
    } else {
        k_fill0<<<(3 * m + kThreads - 1) / kThreads, kThreads, 0, st>>>(t->adj_node, 3 * m);
        CAGR_LAUNCHED();
    }
    k_fill0<<<(3 * n + kThreads - 1) / kThreads, kThreads, 0, st>>>(t->adj_pt, 3 * n);
    CAGR_LAUNCHED();
    const int64_t blocks = (n + 63) / 64;   // one wave per block
    p.tile_order = t->tile_order;
    if (precision == CAGR_PREC_FP32C) {
        auto nodes = (const NodeV32*)nodes_in;
        auto src = (const SrcV32*)src_in;
        // trees of FP64 points retest the near pairs' cutoffs with them
        const bool y64 = t->qround != 0.0;
#define CAGR_ADJ32_Y(GMA, CT, Y)                                                                 \
    do {                                                                                         \
        if (slots)                                                                               \
            k_adj32_list<GMA, CT, Y><<<p.pieces ? t->plan_num_pieces : blocks, 64, 0, st>>>(      \
                p, t->topo, nodes, src, t->pts64, t->pts32);                                     \
        else                                                                                     \
            k_adj32<GMA, CT, FM_ATOMIC, Y><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src,         \
                                                                    t->pts64, t->pts32);         \
    } while (0)
#define CAGR_ADJ32(GMA, CT)                                                                      \
    do {                                                                                         \
        if (y64) CAGR_ADJ32_Y(GMA, CT, true);                                                    \
        else CAGR_ADJ32_Y(GMA, CT, false);                                                       \
    } while (0)
        if (ga) {
            if (continuous) CAGR_ADJ32(1, true);
            else CAGR_ADJ32(1, false);
        } else {
            if (continuous) CAGR_ADJ32(0, true);
            else CAGR_ADJ32(0, false);
        }
#undef CAGR_ADJ32
#undef CAGR_ADJ32_Y
    } else {
        auto nodes = (const NodeV64*)nodes_in;
        auto src = (const SrcV64*)src_in;
        if (ga) {
            if (continuous) k_adj64<1, true><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64);
            else k_adj64<1, false><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64);
        } else {
            if (continuous) k_adj64<0, true><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64);
            else k_adj64<0, false><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64);
        }
    }
    CAGR_LAUNCHED();
    if (slots) {
        k_far_gather<<<(m + kWavesPerBlock - 1) / kWavesPerBlock, kThreads, 0, st>>>(
            m, t->order, t->plan_node_off, t->plan_far_val, t->adj_node);
        CAGR_LAUNCHED();
    }
    const int64_t nb = (n + kThreads - 1) / kThreads;
    if (dtype == CAGR_F32)
        k_adj_down<float><<<nb, kThreads, 0, st>>>(n, t->perm, t->leaf_of, t->topo, t->adj_node, t->adj_pt, (float*)out);
    else
        k_adj_down<double><<<nb, kThreads, 0, st>>>(n, t->perm, t->leaf_of, t->topo, t->adj_node, t->adj_pt, (double*)out);
    CAGR_LAUNCHED();
    return CAGR_OK;
}

void adjoint_plan_free(cagr_tree* t) {
    (void)hipFree(t->plan_w);
    (void)hipFree(t->plan_wave_off);
    (void)hipFree(t->plan_node_off);
    (void)hipFree(t->plan_slot_pos);
    (void)hipFree(t->plan_far_val);
    (void)hipFree(t->plan_far_node);
    (void)hipFree(t->plan_far_mask);
    (void)hipFree(t->plan_near_off);
    (void)hipFree(t->plan_near_leaf);
    (void)hipFree(t->plan_near_mask);
    (void)hipFree(t->plan_pieces);
    (void)hipFree(t->plan_nodes);
    (void)hipFree(t->plan_src);
    (void)hipFree(t->plan_c64);
    t->plan_nodes = t->plan_src = nullptr;
    t->plan_c64 = nullptr;
    t->plan_rec_an = __builtin_nan("");
    t->plan_w = nullptr;
    t->plan_wave_off = t->plan_node_off = t->plan_slot_pos = nullptr;
    t->plan_far_node = t->plan_near_off = t->plan_near_leaf = nullptr;
    t->plan_far_mask = t->plan_near_mask = nullptr;
    t->plan_pieces = nullptr;
    t->plan_num_pieces = 0;
    t->plan_far_val = nullptr;
    t->plan_slots = -1;
}

// The replay's pieces: a wave's cost is its far steps plus kNearCost per
// near entry (a near entry's transposed block costs about as much as that
// many far steps); a wave costing more than a quarter of the mean is cut into
// ranges of about that cost (at most 16), far and near lists alike; the
// pieces run in wave (Morton) order, so the waves running side by side
// replay neighbouring queries' lists (half the mean and longest first within
// each XCD's share measured 4 % slower at 1M points).
constexpr int kNearCost = 10;
constexpr double kReplayPieceCost = 0.25;
constexpr int kReplayMaxPieces = 16;
int plan_pieces(cagr_tree* t, int64_t nw, hipStream_t st) {
    std::vector<int32_t> woff(nw + 1), noff(nw + 1);
    CAGR_HIP(hipMemcpyAsync(woff.data(), t->plan_wave_off, (nw + 1) * sizeof(int32_t),
                            hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipMemcpyAsync(noff.data(), t->plan_near_off, (nw + 1) * sizeof(int32_t),
                            hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    std::vector<double> cost(nw);
    double mean = 0.0;
    for (int64_t w = 0; w < nw; ++w) {
        cost[w] = (double)(woff[w + 1] - woff[w]) + kNearCost * (double)(noff[w + 1] - noff[w]);
        mean += cost[w];
    }
    mean = std::max(1.0, mean / (double)std::max<int64_t>(nw, 1));
    std::vector<int32_t> pcs;
    for (int64_t w = 0; w < nw; ++w) {
        const int P = (int)std::min<double>(
            kReplayMaxPieces, std::max(1.0, std::ceil(cost[w] / (kReplayPieceCost * mean))));
        const int32_t nf = woff[w + 1] - woff[w], nn = noff[w + 1] - noff[w];
        for (int j = 0; j < P; ++j) {
            const int32_t pc[8] = {(int32_t)w,
                                   woff[w] + (int32_t)((int64_t)nf * j / P),
                                   woff[w] + (int32_t)((int64_t)nf * (j + 1) / P),
                                   noff[w] + (int32_t)((int64_t)nn * j / P),
                                   noff[w] + (int32_t)((int64_t)nn * (j + 1) / P), 0, 0, 0};
            pcs.insert(pcs.end(), pc, pc + 8);
        }
    }
    const int64_t G = (int64_t)pcs.size() / 8;
    CAGR_HIP(hipMalloc(&t->plan_pieces, std::max<int64_t>(G, 1) * 8 * sizeof(int32_t)));
    CAGR_HIP(hipMemcpyAsync(t->plan_pieces, pcs.data(), G * 8 * sizeof(int32_t),
                            hipMemcpyHostToDevice, st));
    CAGR_HIP(hipStreamSynchronize(st));
    t->plan_num_pieces = G;
    return CAGR_OK;
}

int adjoint_plan(cagr_tree* t, const void* weights, int dtype, hipStream_t st) {
    adjoint_plan_free(t);
    const int64_t n = t->n, m = t->m;
    CAGR_HIP(hipMalloc(&t->plan_w, n * sizeof(double)));
    k_to_f64<<<(n + kThreads - 1) / kThreads, kThreads, 0, st>>>(weights, dtype, n, t->plan_w);
    CAGR_LAUNCHED();
    // centres of these weights (the walk needs no moments)
    int rc = upward(t, 0, CAGR_PREC_FP32C, 1, 0, nullptr, CAGR_F64, t->plan_w, 0.0, st);
    if (rc) return rc;
    const int64_t blocks = (n + 63) / 64;   // one wave per block, as the adjoint launches
    const int64_t nw = blocks;
    CAGR_HIP(hipMalloc(&t->plan_wave_off, (nw + 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->plan_near_off, (nw + 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->plan_node_off, (m + 1) * sizeof(int32_t)));
    k_fill0_i32<<<(nw + 1 + kThreads - 1) / kThreads, kThreads, 0, st>>>(t->plan_wave_off, nw + 1);
    CAGR_LAUNCHED();
    k_fill0_i32<<<(nw + 1 + kThreads - 1) / kThreads, kThreads, 0, st>>>(t->plan_near_off, nw + 1);
    CAGR_LAUNCHED();
    k_fill0_i32<<<(m + 1 + kThreads - 1) / kThreads, kThreads, 0, st>>>(t->plan_node_off, m + 1);
    CAGR_LAUNCHED();
    AP p{};
    p.n = n;
    p.root = 0;
    p.stop = (int32_t)m;
    p.perm = t->perm;
    p.open_factor = t->open_factor;
    p.cb = dev::cut_band(t->qround);
    p.wave_cnt = t->plan_wave_off;
    p.near_cnt = t->plan_near_off;
    p.node_cnt = t->plan_node_off;
    p.c64 = t->r64;   // the centres of the upward pass above
    auto nodes = (const NodeV32*)t->nodes;
    auto src = (const SrcV32*)t->src;
    p.tile_order = t->tile_order;
    k_adj32<0, false, FM_COUNT><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64, t->pts32);
    CAGR_LAUNCHED();
    // counts -> offsets
    int32_t* tmp = nullptr;
    CAGR_HIP(hipMalloc(&tmp, scan_tmp_len(std::max(nw, m) + 1) * sizeof(int32_t)));
    struct FreeTmp { int32_t** q; ~FreeTmp() { (void)hipFree(*q); } } free_tmp{&tmp};
    if ((rc = scan_exclusive<int32_t>(t->plan_wave_off, nw + 1, tmp, st))) return rc;
    if ((rc = scan_exclusive<int32_t>(t->plan_node_off, m + 1, tmp, st))) return rc;
    if ((rc = scan_exclusive<int32_t>(t->plan_near_off, nw + 1, tmp, st))) return rc;
    int32_t tot[2] = {0, 0};
    CAGR_HIP(hipMemcpyAsync(&tot[0], t->plan_wave_off + nw, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipMemcpyAsync(&tot[1], t->plan_near_off + nw, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    const int64_t F = tot[0], NN = tot[1];
    CAGR_REQUIRE(F >= 0 && NN >= 0, CAGR_ERR_ARG, "adjoint plan: more than 2^31 steps");
    CAGR_HIP(hipMalloc(&t->plan_far_val, std::max<int64_t>(F, 1) * sizeof(float4)));
    CAGR_HIP(hipMalloc(&t->plan_slot_pos, std::max<int64_t>(F, 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->plan_far_node, std::max<int64_t>(F, 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->plan_far_mask, std::max<int64_t>(F, 1) * sizeof(uint64_t)));
    CAGR_HIP(hipMalloc(&t->plan_near_leaf, std::max<int64_t>(NN, 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->plan_near_mask, std::max<int64_t>(NN, 1) * sizeof(uint64_t)));
    if (F > 0 || NN > 0) {
        // the node of every slot; the slots grouped by node (stable: by slot
        // id); the inverse of that order is each slot's node-major position
        uint64_t *keys = nullptr, *ktmp = nullptr;
        int32_t *ids = nullptr, *vtmp = nullptr;
        uint32_t* hist = nullptr;
        struct FreeAll {
            uint64_t **a, **b;
            int32_t **c, **e;
            uint32_t** d;
            ~FreeAll() {
                (void)hipFree(*a); (void)hipFree(*b); (void)hipFree(*c); (void)hipFree(*d);
                (void)hipFree(*e);
            }
        } free_all{&keys, &ktmp, &vtmp, &ids, &hist};
        CAGR_HIP(hipMalloc(&keys, F * sizeof(uint64_t)));
        CAGR_HIP(hipMalloc(&ktmp, F * sizeof(uint64_t)));
        CAGR_HIP(hipMalloc(&ids, F * sizeof(int32_t)));
        CAGR_HIP(hipMalloc(&vtmp, F * sizeof(int32_t)));
        CAGR_HIP(hipMalloc(&hist, sort_hist_len(F) * sizeof(uint32_t)));
        p.wave_off = t->plan_wave_off;
        p.slot_node = keys;
        p.slot_id = ids;
        p.far_node = t->plan_far_node;
        p.far_mask = t->plan_far_mask;
        p.near_off = t->plan_near_off;
        p.near_leaf = t->plan_near_leaf;
        p.near_mask = t->plan_near_mask;
        k_adj32<0, false, FM_FILL><<<blocks, 64, 0, st>>>(p, t->topo, nodes, src, t->pts64, t->pts32);
        CAGR_LAUNCHED();
        int bits = 1;
        while ((int64_t(1) << bits) < m) ++bits;
        if (F > 0) {
            if ((rc = sort_pairs_u64(keys, ids, ktmp, vtmp, hist, F, bits, st))) return rc;
            k_invert<<<(F + kThreads - 1) / kThreads, kThreads, 0, st>>>(ids, F, t->plan_slot_pos);
            CAGR_LAUNCHED();
        }
        CAGR_HIP(hipStreamSynchronize(st));
    }
    t->plan_slots = F;
    return plan_pieces(t, nw, st);
}

}  // namespace cagr
