// Device helpers shared by the traversal kernels (traverse.hip, adjoint.hip).
#pragma once
#include "common.h"

#include <type_traits>

namespace cagr {
namespace dev {

// workgroup -> logical block so that each XCD (workgroups are dealt to the 8
// XCDs round-robin) gets one contiguous, spatially coherent range of blocks
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

__device__ inline double load_as_double(const void* p, int dtype, int64_t i) {
    return dtype == CAGR_F32 ? (double)((const float*)p)[i] : ((const double*)p)[i];
}

// squared distance, each operation rounded (the reference's order:
// d0^2 + d1^2 + d2^2 from the left)
__device__ inline double dist2_rn(double d0, double d1, double d2) {
#pragma clang fp contract(off)
    return (d0 * d0 + d1 * d1) + d2 * d2;
}

// a[i] for a small runtime i without indexing a register array (which would
// spill the array to scratch)
template <typename T>
__device__ inline T sel3(const T* a, int i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

// The band of a near pair's FP32 cutoff test |r2 - w2| <= nb1 w2 + nb0.
// FP32 arithmetic on the pair's float coordinates is off by far less than
// 1e-5 w2 near r = w.  When the tree's points are FP64 values rounded to
// FP32 by up to qr per component (qr = 0 for FP32 points), the query and the
// source add an error of at most 2 qr per component of d, which moves r2 by
// at most 4 sqrt(3) r qr + 12 qr^2 <= 4 qr (w2 + 1) + 64 qr^2 where the test
// can go either way (r within a few qr of w): nb1 = 1e-5 + 4 qr,
// nb0 = 4 qr + 64 qr^2, each rounded up to a float.
struct CutBand {
    float nb1, nb0;
};
inline CutBand cut_band(double qr) {
    const auto up = [](double v) {
        const float f = (float)v;
        return (double)f < v ? nextafterf(f, INFINITY) : f;
    };
    return CutBand{qr > 0 ? up(1e-5 + 4.0 * qr) : 1e-5f, up(4.0 * qr + 64.0 * qr * qr)};
}


// r >= w for a far node decided as the FP64 reference would, from the FP32
// distance^2 to the node's float centre: r2 carries the centre's rounding
// (ce per component, NodeV32 / NodeS32) besides its own, so the band where
// FP32 cannot call it is widened by 4 w ce + 4 ce^2; lanes in the band retest
// with the FP64 centre c (read only then; lanes with use = false never do)
__device__ inline bool inside_centre(float r2, float w2, float wf, float ce, const double* x,
                                     const double* c, double w, bool use = true) {
    const float band = fmaf(4.f * wf, ce, fmaf(4.f * ce, ce, 1e-5f * w2));
    const bool amb = use && fabsf(r2 - w2) <= band;
    bool in = r2 < w2;
    if (__builtin_expect(__ballot(amb) != 0, 0)) {
        if (amb) in = !(sqrt(dist2_rn(x[0] - c[0], x[1] - c[1], x[2] - c[2])) >= w);
    }
    return in;
}


// DPP lane move of a float within rows of 16 lanes (no LDS traffic)
template <int CTRL, int ROWMASK>
__device__ inline float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                                 ROWMASK, 0xf, false));
}

// the sums over the wave's 64 lanes of three values, left in the rows of 16
// lanes: the lane halves are exchanged with v_permlane32_swap (z0 with z1, z2
// with itself) and added, the rows of 16 lanes with v_permlane16_swap, then
// each row is reduced by DPP moves -- every lane of row 0 then holds z0's
// total, of row 1 z2's, of row 2 z1's.  All 64 lanes must be active.
__device__ inline float wave_sum3_rows(float z0, float z1, float z2) {
    auto bits = [](float v) { return __builtin_bit_cast(unsigned, v); };
    auto flt = [](unsigned v) { return __builtin_bit_cast(float, v); };
    const auto a = __builtin_amdgcn_permlane32_swap(bits(z0), bits(z1), false, false);
    const float u = flt(a[0]) + flt(a[1]);   // lanes 0-31: z0[l] + z0[l+32]; 32-63: z1's
    const auto b = __builtin_amdgcn_permlane32_swap(bits(z2), bits(z2), false, false);
    const float v = flt(b[0]) + flt(b[1]);   // every lane l: z2[l % 32] + z2[l % 32 + 32]
    const auto c = __builtin_amdgcn_permlane16_swap(bits(u), bits(v), false, false);
    float w = flt(c[0]) + flt(c[1]);         // rows: u 0+1, v 0+1, u 2+3, v 2+3
    w += dpp_move<0xb1, 0xf>(w);    // quad_perm [1,0,3,2]
    w += dpp_move<0x4e, 0xf>(w);    // quad_perm [2,3,0,1]
    w += dpp_move<0x141, 0xf>(w);   // row_half_mirror
    w += dpp_move<0x140, 0xf>(w);   // row_mirror: every lane holds its row's sum
    return w;
}

// v summed over the lanes that agree in lane mod S (S a power of two <= 64),
// the total in every lane: rotations by S..8 inside rows of 16 lanes (DPP
// row_ror, one vector add each; a rotation by k keeps lane mod k), then the
// row pairs and the wave halves exchanged with v_permlane16_swap and
// v_permlane32_swap.  All 64 lanes must be active.
__device__ inline float group_sum(float v, int S) {
    auto bits = [](float u) { return __builtin_bit_cast(unsigned, u); };
    auto flt = [](unsigned u) { return __builtin_bit_cast(float, u); };
    if (S <= 1) v += dpp_move<0x121, 0xf>(v);   // row_ror:1
    if (S <= 2) v += dpp_move<0x122, 0xf>(v);   // row_ror:2
    if (S <= 4) v += dpp_move<0x124, 0xf>(v);   // row_ror:4
    if (S <= 8) v += dpp_move<0x128, 0xf>(v);   // row_ror:8
    if (S <= 16) {
        const auto c = __builtin_amdgcn_permlane16_swap(bits(v), bits(v), false, false);
        v = flt(c[0]) + flt(c[1]);
    }
    if (S <= 32) {
        const auto c = __builtin_amdgcn_permlane32_swap(bits(v), bits(v), false, false);
        v = flt(c[0]) + flt(c[1]);
    }
    return v;
}

// the same totals as scalars (13 instructions against 21 for three separate
// DPP reductions)
__device__ inline void wave_sum3(float z0, float z1, float z2, float& s0, float& s1, float& s2) {
    auto bits = [](float v) { return __builtin_bit_cast(unsigned, v); };
    auto flt = [](unsigned v) { return __builtin_bit_cast(float, v); };
    const auto a = __builtin_amdgcn_permlane32_swap(bits(z0), bits(z1), false, false);
    const float u = flt(a[0]) + flt(a[1]);   // lanes 0-31: z0[l] + z0[l+32]; 32-63: z1's
    const auto b = __builtin_amdgcn_permlane32_swap(bits(z2), bits(z2), false, false);
    const float v = flt(b[0]) + flt(b[1]);   // every lane l: z2[l % 32] + z2[l % 32 + 32]
    const auto c = __builtin_amdgcn_permlane16_swap(bits(u), bits(v), false, false);
    float w = flt(c[0]) + flt(c[1]);         // rows: u 0+1, v 0+1, u 2+3, v 2+3
    w += dpp_move<0xb1, 0xf>(w);    // quad_perm [1,0,3,2]
    w += dpp_move<0x4e, 0xf>(w);    // quad_perm [2,3,0,1]
    w += dpp_move<0x141, 0xf>(w);   // row_half_mirror
    w += dpp_move<0x140, 0xf>(w);   // row_mirror: every lane holds its row's sum
    s0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, w), 0));
    s2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, w), 16));
    s1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, w), 32));
}

// the same for doubles: each exchange moves the two 32-bit halves, each DPP
// step moves both halves and adds in FP64
__device__ inline void wave_sum3(double z0, double z1, double z2, double& s0, double& s1,
                                 double& s2) {
    auto lo = [](double v) { return (unsigned)(__builtin_bit_cast(uint64_t, v) & 0xffffffffu); };
    auto hi = [](double v) { return (unsigned)(__builtin_bit_cast(uint64_t, v) >> 32); };
    auto dbl = [](unsigned l, unsigned h) {
        return __builtin_bit_cast(double, ((uint64_t)h << 32) | (uint64_t)l);
    };
    const auto al = __builtin_amdgcn_permlane32_swap(lo(z0), lo(z1), false, false);
    const auto ah = __builtin_amdgcn_permlane32_swap(hi(z0), hi(z1), false, false);
    const double u = dbl(al[0], ah[0]) + dbl(al[1], ah[1]);
    const auto bl = __builtin_amdgcn_permlane32_swap(lo(z2), lo(z2), false, false);
    const auto bh = __builtin_amdgcn_permlane32_swap(hi(z2), hi(z2), false, false);
    const double v = dbl(bl[0], bh[0]) + dbl(bl[1], bh[1]);
    const auto cl = __builtin_amdgcn_permlane16_swap(lo(u), lo(v), false, false);
    const auto ch = __builtin_amdgcn_permlane16_swap(hi(u), hi(v), false, false);
    double w = dbl(cl[0], ch[0]) + dbl(cl[1], ch[1]);
    auto move = [&](auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo(w), C, 0xf, 0xf, false);
        const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi(w), C, 0xf, 0xf, false);
        return dbl(l, h);
    };
    w += move(std::integral_constant<int, 0xb1>{});
    w += move(std::integral_constant<int, 0x4e>{});
    w += move(std::integral_constant<int, 0x141>{});
    w += move(std::integral_constant<int, 0x140>{});
    s0 = dbl((unsigned)__builtin_amdgcn_readlane((int)lo(w), 0),
             (unsigned)__builtin_amdgcn_readlane((int)hi(w), 0));
    s2 = dbl((unsigned)__builtin_amdgcn_readlane((int)lo(w), 16),
             (unsigned)__builtin_amdgcn_readlane((int)hi(w), 16));
    s1 = dbl((unsigned)__builtin_amdgcn_readlane((int)lo(w), 32),
             (unsigned)__builtin_amdgcn_readlane((int)hi(w), 32));
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Deepest node depth at which "far" already proves r >= w for every lane of
// the wave (far means dist > open_factor * 2^-depth): for those nodes the
// cutoff test of the far interaction can be skipped.  wmax = the wave's
// largest width.
__device__ inline int far_safe_depth(double open_factor, double wmax) {
    if (!(wmax > 0.0)) return 64;
    const double ratio = open_factor / (wmax * (1.0 + 1e-4));
    if (!(ratio >= 1.0)) return -1;
    return ilogb(ratio);
}

// Is every point of a leaf (bounding sphere lb) certainly more than w away
// from the query xf, with margins far beyond FP32 rounding?  Then the leaf
// needs no cutoff test at all.
__device__ inline bool leaf_surely_outside(const float* xf, float4 lb, float wf) {
    const float a = xf[0] - lb.x, b = xf[1] - lb.y, c = xf[2] - lb.z;
    const float d2 = fmaf(a, a, fmaf(b, b, c * c));
    const float lim = lb.w + wf * 1.0001f + 1e-6f;
    return d2 * (1.f - 1e-6f) > lim * lim;
}

// Make the listed wave-uniform values available at this point: the scalar
// loads of a node step's records are issued together and waited for once,
// instead of each where a branch first uses it (one scalar-memory round trip
// per step instead of several)
#define CAGR_LOADS_HERE(...) asm volatile("" ::__VA_ARGS__)

// Two source records' fields available here: their scalar loads are issued
// together and waited for once (a scalar load is waited for with
// lgkmcnt(0), so the compiler would otherwise wait for the first record
// before issuing the second's loads)
__device__ __forceinline__ void loads_here(const SrcV32& a, const SrcV32& b) {
    CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.mu[0]), "s"(a.mu[1]),
                    "s"(a.mu[2]), "s"(a.e[0]), "s"(a.e[1]), "s"(a.e[2]), "s"(a.q[0]), "s"(a.q[1]),
                    "s"(a.q[2]), "s"(b.y[0]), "s"(b.y[1]), "s"(b.y[2]), "s"(b.mu[0]), "s"(b.mu[1]),
                    "s"(b.mu[2]), "s"(b.e[0]), "s"(b.e[1]), "s"(b.e[2]), "s"(b.q[0]), "s"(b.q[1]),
                    "s"(b.q[2]));
}
template <int K>
__device__ __forceinline__ void loads_here(const SrcS32<K>& a, const SrcS32<K>& b) {
    if constexpr (K == 3) {
        CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.a[0]), "s"(a.a[1]),
                        "s"(a.a[2]), "s"(b.y[0]), "s"(b.y[1]), "s"(b.y[2]), "s"(b.a[0]),
                        "s"(b.a[1]), "s"(b.a[2]));
    } else {
        static_assert(K == 1, "one or three sets");
        CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.a[0]), "s"(b.y[0]),
                        "s"(b.y[1]), "s"(b.y[2]), "s"(b.a[0]));
    }
}

__device__ __forceinline__ void loads_here(const SrcV64& a, const SrcV64& b) {
    CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.mu[0]), "s"(a.mu[1]),
                    "s"(a.mu[2]), "s"(b.y[0]), "s"(b.y[1]), "s"(b.y[2]), "s"(b.mu[0]),
                    "s"(b.mu[1]), "s"(b.mu[2]));
}
template <int K>
__device__ __forceinline__ void loads_here(const SrcS64<K>& a, const SrcS64<K>& b) {
    if constexpr (K == 3) {
        CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.s[0]), "s"(a.s[1]),
                        "s"(a.s[2]), "s"(b.y[0]), "s"(b.y[1]), "s"(b.y[2]), "s"(b.s[0]),
                        "s"(b.s[1]), "s"(b.s[2]));
    } else {
        static_assert(K == 1, "one or three sets");
        CAGR_LOADS_HERE("s"(a.y[0]), "s"(a.y[1]), "s"(a.y[2]), "s"(a.s[0]), "s"(b.y[0]),
                        "s"(b.y[1]), "s"(b.y[2]), "s"(b.s[0]));
    }
}

// r >= w of query x and source y decided as the FP64 reference would: FP32
// unless within band (nb1 w2 + nb0) of w2, then the FP64 test with the FP64
// points (y read only there; lanes with use = false never take it)
__device__ inline bool inside_fp32(float r2, float w2, float band, const double* x,
                                   const double* y, double w, bool use = true) {
    const bool amb = use && fabsf(r2 - w2) <= band;
    bool in = r2 < w2;
    if (__builtin_expect(__ballot(amb) != 0, 0)) {
        if (amb) in = !(sqrt(dist2_rn(x[0] - y[0], x[1] - y[1], x[2] - y[2])) >= w);
    }
    return in;
}

// f(rec, index) for the wave-uniform records arr[b..e) in order, two at a
// time, the two records' loads waited for once
template <typename R, typename F>
__device__ __forceinline__ void for_each_rec(const R* __restrict__ arr, int32_t b, int32_t e,
                                             F&& f) {
    int32_t j = b;
    for (; j + 1 < e; j += 2) {
        const R r0 = arr[j];
        const R r1 = arr[j + 1];
        loads_here(r0, r1);
        __builtin_amdgcn_sched_barrier(0);
        f(r0, j);
        f(r1, j + 1);
    }
    if (j < e) f(arr[j], j);
}

// Element i of a record array addressed by a 32-bit byte offset: a scalar
// load then takes base + SGPR offset (one multiply) instead of 64-bit address
// arithmetic on every node step.  The host guarantees count * sizeof(T) < 2^32.
template <typename T>
__device__ inline const T& rec32(const T* base, int32_t i) {
    return *(const T*)((const char*)base + (uint64_t)((uint32_t)i * (uint32_t)sizeof(T)));
}

// ---- spreading a near leaf over the idle lanes ----
// At a leaf where only na lanes are near, lane L works for slot q = L >> lr
// (the near lane of rank q: the q-th set bit of the near mask) as group
// g = L & (2^lr - 1) of that slot, taking the leaf's points g, g + 2^lr, ...;
// the 2^lr partial sums of a slot are then added in a fixed butterfly and the
// near lane of rank q takes the slot's total.  All 64 lanes must be active in
// these helpers.
struct Spread {
    int lr;      // log2 of the lanes per slot
    int g;       // this lane's group within its slot
    bool part;   // this lane works for a slot (L < na << lr)
    int owner;   // the near lane whose query this lane's slot evaluates
    int rank;    // this lane's rank among the near lanes (meaningful on them)
};

// lanes per slot (log2): as many as fit next to na near lanes, at most 16,
// and not more than the leaf's points; 0 = do not spread (small leaves)
__device__ inline int spread_log2(int na, int npts) {
    int lr = 0;
    if (npts < 4) return 0;
    while (lr < 4 && (na << (lr + 1)) <= 64 && (2 << lr) <= npts) ++lr;
    return lr;
}

template <typename T>
__device__ inline T lane_pull(T v, int src_lane) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_ds_bpermute(src_lane << 2, __builtin_bit_cast(int, v)));
    } else {
        static_assert(sizeof(T) == 8, "32- or 64-bit values only");
        const int64_t b = __builtin_bit_cast(int64_t, v);
        const int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(b & 0xffffffff));
        const int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(b >> 32));
        return __builtin_bit_cast(T, ((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
    }
}

__device__ inline Spread spread_setup(uint64_t nm, bool nearl, int lr) {
    const int lane = (int)(threadIdx.x & 63);
    const int na = __popcll(nm);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi(
        (uint32_t)(nm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
    // near lanes send their id to lane `rank`, the others to lanes na.. (a
    // permutation of the 64 lanes), so lane q < na receives the id of slot q
    const int dest = nearl ? rank : na + (lane - rank);
    const int own = __builtin_amdgcn_ds_permute(dest << 2, lane);
    Spread s;
    s.lr = lr;
    s.g = lane & ((1 << lr) - 1);
    const int q = lane >> lr;
    s.part = q < na;
    s.owner = __builtin_amdgcn_ds_bpermute((s.part ? q : 0) << 2, own);
    s.rank = rank;
    return s;
}

// sum over the 2^lr lanes of a slot, lr <= 4 (every lane of the slot gets the
// same total: each step adds the same two values in either lane).  The steps
// are DPP lane moves inside rows of 16, each one vector add with a moved
// operand: lanes l and l^1, l and l^2 (quad permutations), then the two
// quads of a half row (half-row mirror) and the two halves of a row (row
// mirror) -- a shuffle through LDS would cost an address computation, a
// ds_bpermute and its wait per step
template <int CTRL>
__device__ inline double dpp_move_f64(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf,
                                                              false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int CTRL>
__device__ inline float dpp_step(float v) { return dpp_move<CTRL, 0xf>(v); }
template <int CTRL>
__device__ inline double dpp_step(double v) { return dpp_move_f64<CTRL>(v); }

template <typename T>
__device__ inline T spread_sum(T v, int lr) {
    if (lr >= 1) v += dpp_step<0xb1>(v);    // quad_perm [1,0,3,2]
    if (lr >= 2) v += dpp_step<0x4e>(v);    // quad_perm [2,3,0,1]
    if (lr >= 3) v += dpp_step<0x141>(v);   // row_half_mirror
    if (lr >= 4) v += dpp_step<0x140>(v);   // row_mirror
    return v;
}
// the same for NV values at once (one branch per step for all of them)
template <int NV, typename T>
__device__ inline void spread_sums(T* v, int lr) {
    if (lr >= 1)
        for (int i = 0; i < NV; ++i) v[i] += dpp_step<0xb1>(v[i]);
    if (lr >= 2)
        for (int i = 0; i < NV; ++i) v[i] += dpp_step<0x4e>(v[i]);
    if (lr >= 3)
        for (int i = 0; i < NV; ++i) v[i] += dpp_step<0x141>(v[i]);
    if (lr >= 4)
        for (int i = 0; i < NV; ++i) v[i] += dpp_step<0x140>(v[i]);
}

// lane `lane`'s 64-bit value to every lane (a scalar)
__device__ inline uint64_t lane_bcast_u64(uint64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T>
__device__ inline T lane_bcast(T v, int lane) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
    } else {
        const int64_t b = __builtin_bit_cast(int64_t, v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
        return __builtin_bit_cast(T, ((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
    }
}

}  // namespace dev
}  // namespace cagr
