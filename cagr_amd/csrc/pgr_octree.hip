// GaussRecon's PGR adaptive octree over the samples and its query corners
// (Octree::setTree, gaussrecon_src/Octree.cpp:193-400; the query set of
// main_GaussRecon_cuda.cu:255-291), built with sorts and segment sums.
//
// The reference inserts the samples one by one into a pointer octree.  The
// quantities that decide its final shape are float sums per cell, taken in
// sample order:
//   * centre weights: every sample adds quadratic B-spline weights to the 27
//     cells around its cell at each depth 0..S, S = D - 2
//     (NonLinearUpdateWeightContribution, Octree.cpp:1292-1325);
//   * normal lengths: every sample, at the depth its centre weights give
//     (NonLinearGetSampleDepthAndWeight, :1403-1427), adds a length to the 27
//     cells around it at that depth and, for a fractional depth, at the depth
//     above (:1328-1401, :1461-1497).
// ClipTree (:1499-1528) then keeps the children of a node iff a strict
// descendant has |normal length| > 1e-6, so the final tree is the root plus
// the children of every strict ancestor of such a cell; node creation by
// NeighborKey::setNeighbors only adds zero-valued cells and does not matter.
//
// Here each sum is a segment of records sorted by cell (stable, so sample
// order is kept inside a cell) and one wave adds a segment serially in the
// reference's float / double mix, which makes every sum bit-identical to the
// sequential build.  The tree follows level by level from the deepest cells
// with normals (sort + unique of parents), the leaves come out in preorder by
// a sort on left-aligned Morton keys, and the corners of the leaves in
// gridDataVector order by a sort on the corner keys plus a scan over first
// occurrences (Octree.cpp:312-360, 432-477; ReconOctNode.cpp:95-105).
//
// Every floating-point expression keeps the reference's types and order and is
// compiled without contraction.  One difference remains: the sample depth uses
// logf, which here is log in double rounded to float (correctly rounded) while
// glibc's logf is within 0.82 ulp; the two differ in the last bit for about
// 0.2 % of arguments.  That can change the tree only when a sample's depth lies
// within an ulp of an integer or a cell's normal length within an ulp of 1e-6.
#include <algorithm>
#include <cmath>
#include <new>

#include "common.h"

#pragma clang fp contract(off)

#include "pgr.h"

namespace cagr {
namespace {

constexpr int kT = 256;
inline int64_t nblk(int64_t n) { return (n + kT - 1) / kT; }

constexpr double kSampleScale = 1. / (0.125 * 0.125 + 0.75 * 0.75 + 0.125 * 0.125);
constexpr int kMaxD = 13;                       // cell offsets fit 13 bits per axis
constexpr uint64_t kNoCell = 1ull << 43;        // sorts after every cell key
constexpr int kCellBits = 44;

__host__ __device__ inline uint64_t ckey(int d, uint32_t x, uint32_t y, uint32_t z) {
    return ((uint64_t)d << 39) | ((uint64_t)z << 26) | ((uint64_t)y << 13) | x;
}
__host__ __device__ inline void cdecode(uint64_t k, int& d, uint32_t& x, uint32_t& y, uint32_t& z) {
    d = (int)(k >> 39);
    z = (uint32_t)(k >> 26) & 0x1FFF;
    y = (uint32_t)(k >> 13) & 0x1FFF;
    x = (uint32_t)k & 0x1FFF;
}
// cell keys of one depth, compact (z, y, x) with d bits each
__device__ inline uint64_t tkey(int d, uint32_t x, uint32_t y, uint32_t z) {
    return ((uint64_t)z << (2 * d)) | ((uint64_t)y << d) | x;
}

// cell index of coordinate p in (0, 1] at depth d.  The reference descends
// comparing p with the exact cell centres, going up iff p > centre
// (Cube::CornerIndex, Cube.cpp:46-52); with t = p * 2^(l+1) (exact) the step
// is o' = 2o + (t > 2o + 1), whose closed form is o_d = ceil(p * 2^d) - 1.
__device__ inline uint32_t descend(float p, int d) {
    const int c = (int)ceilf(ldexpf(p, d)) - 1;
    return (uint32_t)max(c, 0);
}

// float centre of cell o at depth d (ReconOctNode::centerAndWidth)
__device__ inline float centre(int d, uint32_t o) {
    const float w = (float)1.0 / (1 << d);
    return (float)((0.5 + (double)o) * w);
}

// B-spline weights of the 3 cells along one axis (Octree.cpp:1302-1307); the
// divisions by the width 2^-d are multiplications by 2^d, which round the same
__device__ inline void spline3(float c, float p, double width, double inv_width, double out[3]) {
    const float cp = c - p;
    double x = (cp - width) * inv_width;
    out[0] = 1.125 + 1.5 * x + 0.5 * x * x;
    x = cp * inv_width;
    out[1] = 0.75 - x * x;
    out[2] = 1.0 - out[1] - out[0];
}

__device__ inline void axis_weights(const float p[3], int d, uint32_t o[3], double dx[3][3]) {
    const double w = (double)((float)1.0 / (1 << d));
    const double iw = (double)(1 << d);
    for (int k = 0; k < 3; ++k) {
        o[k] = descend(p[k], d);
        spline3(centre(d, o[k]), p[k], w, iw, dx[k]);
    }
}

// ------------------------------------------------------------ bounding box
__global__ __launch_bounds__(kT) void k_bbox_part(const float* __restrict__ pts, int64_t n,
                                                  float* __restrict__ part) {
    __shared__ float s[6][kT];
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) mn[k] = mx[k] = pts[k];
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
        for (int k = 0; k < 3; ++k) {
            const float v = pts[3 * i + k];
            if (v > mx[k]) mx[k] = v;
            if (v < mn[k]) mn[k] = v;
        }
    for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = mn[k];
        s[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 3; ++k) {
                const float a = s[k][threadIdx.x + h], b = s[3 + k][threadIdx.x + h];
                if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
                if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_bbox_final(const float* __restrict__ part, int nb,
                                                   float* __restrict__ out) {
    if (threadIdx.x >= 6) return;
    const int k = threadIdx.x;
    float v = part[k];
    for (int b = 1; b < nb; ++b) {
        const float u = part[6 * b + k];
        if (k < 3 ? u < v : u > v) v = u;
    }
    out[k] = v;
}

__global__ __launch_bounds__(kT) void k_normalize(const float* __restrict__ pts, int64_t n,
                                                  float c0, float c1, float c2, float scale,
                                                  float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= 3 * n) return;
    const int k = (int)(i % 3);
    const float c = k == 0 ? c0 : (k == 1 ? c1 : c2);
    out[i] = (pts[i] - c) / scale;
}

// ------------------------------------------------------------ segments
// heads of the runs of equal keys below `none`; flags[m] = 0 closes the scan
__global__ __launch_bounds__(kT) void k_heads(const uint64_t* __restrict__ keys, int64_t m,
                                              uint64_t none, int32_t* __restrict__ flags) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r > m) return;
    if (r == m) {
        flags[r] = 0;
        return;
    }
    const uint64_t k = keys[r];
    flags[r] = (k < none && (r == 0 || keys[r - 1] != k)) ? 1 : 0;
}

// after the scan: start[s] of every run, start[U] = end of the last run,
// ukeys[s] = the run's key
__global__ __launch_bounds__(kT) void k_starts(const uint64_t* __restrict__ keys, int64_t m,
                                               uint64_t none, const int32_t* __restrict__ pos,
                                               int32_t* __restrict__ start,
                                               uint64_t* __restrict__ ukeys) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= m) return;
    const uint64_t k = keys[r];
    if (k >= none) return;
    if (pos[r + 1] != pos[r]) {
        start[pos[r]] = (int32_t)r;
        if (ukeys) ukeys[pos[r]] = k;
    }
    if (r + 1 == m || keys[r + 1] >= none) start[pos[m]] = (int32_t)(r + 1);
}

// serial float sum of a segment in record order, in the reference's
// float-accumulator / double-term form; one wave per segment.  The wave
// evaluates 64 terms at once, then folds them in lane order: three dependent
// instructions per term (widen, add, narrow) on a wave-uniform loop.
template <class F>
__device__ inline float wave_serial_sum(int32_t lo, int32_t hi, F term) {
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
    const int lane = threadIdx.x & 63;
    float acc = 0.f;
    auto fold = [&](int vlo, int vhi, int l) {
        const uint64_t b = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(vlo, l) |
                           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vhi, l) << 32);
        acc = (float)((double)acc + __longlong_as_double((long long)b));
    };
    for (int32_t base = lo; base < hi; base += 64) {
        const int32_t r = base + lane;
        const int64_t v = __double_as_longlong(r < hi ? term(r) : 0.0);
        const int vlo = (int)v, vhi = (int)(v >> 32);
        if (hi - base >= 64) {
#pragma unroll
            for (int l = 0; l < 64; ++l) fold(vlo, vhi, l);
        } else {
            const int cnt = hi - base;
            for (int l = 0; l < cnt; ++l) fold(vlo, vhi, l);
        }
    }
    return acc;
}

// ------------------------------------------------------------ centre weights
__global__ __launch_bounds__(kT) void k_cwc_recs(const float* __restrict__ p, int64_t n, int d,
                                                 uint64_t* __restrict__ keys,
                                                 int32_t* __restrict__ vals) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= 27 * n) return;
    const int64_t i = r / 27;
    const int s = (int)(r % 27);
    const int64_t lim = (int64_t)1 << d;
    const int64_t x = (int64_t)descend(p[3 * i], d) + s / 9 - 1;
    const int64_t y = (int64_t)descend(p[3 * i + 1], d) + (s / 3) % 3 - 1;
    const int64_t z = (int64_t)descend(p[3 * i + 2], d) + s % 3 - 1;
    const bool in = x >= 0 && y >= 0 && z >= 0 && x < lim && y < lim && z < lim;
    keys[r] = in ? tkey(d, (uint32_t)x, (uint32_t)y, (uint32_t)z) : (1ull << (3 * d));
    vals[r] = (int32_t)r;
}

// one entry per cell of every depth: its record run [lo, hi) in the
// concatenated sorted records and its depth
struct Seg {
    int64_t lo, hi;
    int32_t d, pad;
};

__global__ __launch_bounds__(kT) void k_seg_table(const int32_t* __restrict__ start, int32_t U,
                                                  int d, int64_t base, Seg* __restrict__ seg) {
    const int32_t s = (int32_t)((int64_t)blockIdx.x * kT + threadIdx.x);
    if (s < U) seg[s] = Seg{base + start[s], base + start[s + 1], d, 0};
}

// all depths in one launch, the shallow (longest, serial) cells first, so
// their chains run side by side
__global__ __launch_bounds__(kT) void k_cwc_sum(const float* __restrict__ p,
                                                const int32_t* __restrict__ vals, int64_t per_depth,
                                                const Seg* __restrict__ seg, int32_t nseg,
                                                float* __restrict__ cwc) {
    const int32_t s = (int32_t)(((int64_t)blockIdx.x * kT + threadIdx.x) >> 6);
    if (s >= nseg) return;
    const Seg g = seg[s];
    const int d = __builtin_amdgcn_readfirstlane(g.d);
    const int32_t* v = vals + (int64_t)d * per_depth;
    const int64_t off = (int64_t)d * per_depth;
    const float acc = wave_serial_sum((int32_t)(g.lo - off), (int32_t)(g.hi - off), [&](int32_t r) {
        const int32_t rv = v[r];
        const int64_t i = rv / 27;
        const int t = rv % 27;
        const float q[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        uint32_t o[3];
        double dx[3][3];
        axis_weights(q, d, o, dx);
        for (int k = 0; k < 3; ++k) dx[k][0] *= kSampleScale;
        const double dxdy = dx[0][t / 9] * dx[1][(t / 3) % 3] * (double)1.0f;
        return dxdy * dx[2][t % 3];
    });
    if ((threadIdx.x & 63) == 0) cwc[s] = acc;
}

// ------------------------------------------------------------ sample depths, splats
struct Levels {
    const uint64_t* uk[kMaxD + 1];
    const float* cw[kMaxD + 1];
    int32_t U[kMaxD + 1];
};

__device__ inline float lookup(const Levels& L, int d, uint64_t key) {
    const uint64_t* a = L.uk[d];
    int32_t lo = 0, hi = L.U[d];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < L.U[d] && a[lo] == key) ? L.cw[d][lo] : 0.f;
}

// NonLinearGetSampleWeight (Octree.cpp:1429-1458)
__device__ inline float sample_weight(const Levels& L, const float q[3], int d) {
    uint32_t o[3];
    double dx[3][3];
    axis_weights(q, d, o, dx);
    const int64_t lim = (int64_t)1 << d;
    float sum = 0.f;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double dxdy = dx[0][a] * dx[1][b];
            for (int e = 0; e < 3; ++e) {
                const int64_t x = (int64_t)o[0] + a - 1, y = (int64_t)o[1] + b - 1,
                              z = (int64_t)o[2] + e - 1;
                if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
                const float c = lookup(L, d, tkey(d, (uint32_t)x, (uint32_t)y, (uint32_t)z));
                sum += (float)(dxdy * dx[2][e] * (double)c);
            }
        }
    return (float)(1.0 / (double)sum);
}

__device__ inline float logf_rn(float x) { return (float)log((double)x); }

__device__ inline void splat_recs(const float q[3], int d, float len, uint64_t* keys,
                                  double* v64) {
    uint32_t o[3];
    double dx[3][3];
    axis_weights(q, d, o, dx);
    const int64_t lim = (int64_t)1 << d;
    int j = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double dxdy = dx[0][a] * dx[1][b];
            for (int e = 0; e < 3; ++e, ++j) {
                const int64_t x = (int64_t)o[0] + a - 1, y = (int64_t)o[1] + b - 1,
                              z = (int64_t)o[2] + e - 1;
                const bool in = x >= 0 && y >= 0 && z >= 0 && x < lim && y < lim && z < lim;
                keys[j] = in ? ckey(d, (uint32_t)x, (uint32_t)y, (uint32_t)z) : kNoCell;
                v64[j] = (double)len * (dxdy * dx[2][e]);
            }
        }
}

// Only the centre weights of depths >= min_level are read: a sample whose
// depth search climbs above them is deferred (todo[i] = 1, no records
// written) to a second launch with every level; `retry` launches process
// only the samples with todo set.
__global__ __launch_bounds__(kT) void k_splat_recs(const float* __restrict__ p, int64_t n, int S,
                                                   int D, int dmin, double log4, float normal_len,
                                                   Levels L, int min_level, bool retry,
                                                   int32_t* __restrict__ todo,
                                                   int32_t* __restrict__ n_todo,
                                                   uint64_t* __restrict__ keys,
                                                   int32_t* __restrict__ vals,
                                                   double* __restrict__ v64) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    if (retry && !todo[i]) return;
    auto defer = [&]() {
        todo[i] = 1;
        atomicAdd(n_todo, 1);
    };
    const float q[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    const float spn = 1.0f;
    // NonLinearGetSampleDepthAndWeight (Octree.cpp:1403-1427)
    int d = S;
    if (d < min_level) return defer();
    const float weight = 1.0f / sample_weight(L, q, d);
    float depth;
    if (weight >= spn) {
        depth = (float)(d + (double)logf_rn(weight / spn) / log4);
    } else {
        float old_a = weight, new_a = weight;
        while (new_a < spn && d > 0) {
            --d;
            if (d < min_level) return defer();
            old_a = new_a;
            new_a = 1.0f / sample_weight(L, q, d);
        }
        depth = (float)(d + logf_rn(new_a / spn) / logf_rn(new_a / old_a));
    }
    const float alpha = (float)exp2(-2.0 * (double)depth);
    // NonLinearSplatOrientedPoint (Octree.cpp:1356-1400)
    if (depth < dmin) depth = (float)dmin;
    if (depth > D) depth = (float)D;
    int top = (int)ceilf(depth);
    double dx = 1.0 - ((float)top - depth);
    if (top <= dmin) {
        top = dmin;
        dx = 1;
    } else if (top > D) {
        top = D;
        dx = 1;
    }
    uint64_t* kk = keys + 54 * i;
    double* vv = v64 + 54 * i;
    double width = 1.0 / (1 << top);
    const float len = (float)(normal_len * alpha / (float)(width * width * width) * dx);
    splat_recs(q, top, len, kk, vv);
    if (fabs(1.0 - dx) > 1e-6) {
        dx = (float)(1.0 - dx);
        width = 1.0 / (1 << (top - 1));
        const float len2 = (float)((double)(normal_len * alpha) / (width * width * width) * dx);
        splat_recs(q, top - 1, len2, kk + 27, vv + 27);
    } else {
        for (int j = 27; j < 54; ++j) {
            kk[j] = kNoCell;
            vv[j] = 0.0;
        }
    }
    for (int j = 0; j < 54; ++j) vals[54 * i + j] = (int32_t)(54 * i + j);
    if (!retry) todo[i] = 0;
}

__global__ __launch_bounds__(kT) void k_nl_sum(const double* __restrict__ v64,
                                               const int32_t* __restrict__ vals,
                                               const int32_t* __restrict__ start, int32_t U,
                                               int32_t* __restrict__ flag) {
    const int32_t s = (int32_t)(((int64_t)blockIdx.x * kT + threadIdx.x) >> 6);
    if (s > U) return;
    if (s == U) {
        if ((threadIdx.x & 63) == 0) flag[U] = 0;
        return;
    }
    const float acc = wave_serial_sum(start[s], start[s + 1],
                                      [&](int32_t r) { return v64[vals[r]]; });
    if ((threadIdx.x & 63) == 0) flag[s] = fabsf(acc) > 1e-6f ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_compact(const uint64_t* __restrict__ in,
                                                const int32_t* __restrict__ pos, int32_t U,
                                                uint64_t* __restrict__ out) {
    const int32_t s = (int32_t)((int64_t)blockIdx.x * kT + threadIdx.x);
    if (s < U && pos[s + 1] != pos[s]) out[pos[s]] = in[s];
}

// first position of each depth in the depth-sorted cell keys: off[d], d < 17
__global__ __launch_bounds__(64) void k_depth_off(const uint64_t* __restrict__ k, int32_t m,
                                                  int32_t* __restrict__ off) {
    const int d = threadIdx.x;
    if (d > 16) return;
    int32_t lo = 0, hi = m;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((int)(k[mid] >> 39) < d) lo = mid + 1;
        else hi = mid;
    }
    off[d] = lo;
}

// ------------------------------------------------------------ clipped tree
__global__ __launch_bounds__(kT) void k_parents(const uint64_t* __restrict__ a, int32_t na,
                                                const uint64_t* __restrict__ b, int32_t nb,
                                                uint64_t* __restrict__ keys,
                                                int32_t* __restrict__ vals) {
    const int32_t r = (int32_t)((int64_t)blockIdx.x * kT + threadIdx.x);
    if (r >= na + nb) return;
    int d;
    uint32_t x, y, z;
    cdecode(r < na ? a[r] : b[r - na], d, x, y, z);
    keys[r] = ckey(d - 1, x >> 1, y >> 1, z >> 1);
    vals[r] = r;
}

__device__ inline bool contains(const uint64_t* a, int32_t n, uint64_t k) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == k;
}

__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 13; ++b) r |= (uint64_t)((v >> b) & 1u) << (3 * b);
    return r;
}

// children of the internal nodes of one depth: leaves get their preorder key
// (Morton code at depth H, x lowest, then the depth), internal ones kNone
__global__ __launch_bounds__(kT) void k_children(const uint64_t* __restrict__ in_d, int32_t n_d,
                                                 const uint64_t* __restrict__ in_next,
                                                 int32_t n_next, int H, uint64_t none,
                                                 uint64_t* __restrict__ keys,
                                                 int32_t* __restrict__ vals) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= 8 * (int64_t)n_d) return;
    int d;
    uint32_t x, y, z;
    cdecode(in_d[r >> 3], d, x, y, z);
    const int c = (int)(r & 7);
    x = 2 * x + (c & 1);
    y = 2 * y + ((c >> 1) & 1);
    z = 2 * z + ((c >> 2) & 1);
    ++d;
    const uint64_t ck = ckey(d, x, y, z);
    uint64_t key = none;
    if (!contains(in_next, n_next, ck)) {
        const uint64_t m = spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
        key = ((m << (3 * (H - d))) << 4) | (uint64_t)d;
    }
    keys[r] = key;
    vals[r] = (int32_t)r;
}

__global__ __launch_bounds__(kT) void k_leaf_decode(const uint64_t* __restrict__ keys, int64_t nl,
                                                    int H, int32_t* __restrict__ leaves) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= nl) return;
    const uint64_t k = keys[r];
    const int d = (int)(k & 15);
    const uint64_t m = (k >> 4) >> (3 * (H - d));
    uint32_t o[3] = {0, 0, 0};
    for (int b = 0; b < d; ++b)
        for (int a = 0; a < 3; ++a) o[a] |= (uint32_t)((m >> (3 * b + a)) & 1u) << b;
    leaves[4 * r] = d;
    leaves[4 * r + 1] = (int32_t)o[0];
    leaves[4 * r + 2] = (int32_t)o[1];
    leaves[4 * r + 3] = (int32_t)o[2];
}

// ------------------------------------------------------------ corners
// ReconOctNode::getCornerIndex (ReconOctNode.cpp:95-105) on the grid of depth md
__device__ inline int64_t corner_key(const int32_t* lf, int c, int md) {
    const int d = lf[0];
    int64_t off[3];
    for (int k = 0; k < 3; ++k) off[k] = ((int64_t)lf[1 + k] + ((c >> k) & 1)) << (md - d + 1);
    return off[0] | (off[1] << 15) | (off[2] << 30);
}

__global__ __launch_bounds__(kT) void k_corner_keys(const int32_t* __restrict__ leaves, int64_t R,
                                                    int md, uint64_t* __restrict__ keys,
                                                    int32_t* __restrict__ vals,
                                                    int32_t* __restrict__ first) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r > R) return;
    first[r] = 0;
    if (r == R) return;
    keys[r] = (uint64_t)corner_key(leaves + 4 * (r >> 3), (int)(r & 7), md);
    vals[r] = (int32_t)r;
}

__global__ __launch_bounds__(kT) void k_corner_first(const uint64_t* __restrict__ keys,
                                                     const int32_t* __restrict__ vals, int64_t R,
                                                     int32_t* __restrict__ first) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r < R && (r == 0 || keys[r] != keys[r - 1])) first[vals[r]] = 1;
}

// the corner keys in key order: ck[rank] = key, ck_idx[rank] = its
// gridDataVector position (the run's first record is its first occurrence)
__global__ __launch_bounds__(kT) void k_corner_index(const uint64_t* __restrict__ keys,
                                                     const int32_t* __restrict__ vals, int64_t R,
                                                     const int32_t* __restrict__ pos,
                                                     const int32_t* __restrict__ rank,
                                                     int64_t* __restrict__ ck,
                                                     int32_t* __restrict__ ck_idx) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= R || rank[r + 1] == rank[r]) return;
    ck[rank[r]] = (int64_t)keys[r];
    ck_idx[rank[r]] = pos[vals[r]];
}

__global__ __launch_bounds__(kT) void k_corner_write(const int32_t* __restrict__ leaves,
                                                     const int32_t* __restrict__ pos, int64_t R,
                                                     int md, float grid,
                                                     float* __restrict__ corners) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= R || pos[r + 1] == pos[r]) return;
    const int64_t key = corner_key(leaves + 4 * (r >> 3), (int)(r & 7), md);
    float* c = corners + 3 * (int64_t)pos[r];
    c[0] = (float)(key & 0xFFFF) * grid;
    c[1] = (float)((key >> 15) & 0xFFFF) * grid;
    c[2] = (float)(key >> 30) * grid;
}

// ------------------------------------------------------------ host side
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t n) {
        if (n <= cap) return CAGR_OK;
        if (p) CAGR_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        CAGR_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        cap = n;
        return CAGR_OK;
    }
};

// a second stream (default priority) for work that overlaps the main one
struct SideStream {
    hipStream_t main = nullptr, s = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int create(hipStream_t st) {
        main = st;
        CAGR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        CAGR_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        CAGR_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        return CAGR_OK;
    }
    // the side stream waits for everything enqueued on the main stream so far
    hipError_t fork() {
        hipError_t e = hipEventRecord(ev_fork, main);
        return e != hipSuccess ? e : hipStreamWaitEvent(s, ev_fork, 0);
    }
    // the main stream waits for everything enqueued on the side stream
    hipError_t join() {
        hipError_t e = hipEventRecord(ev_join, s);
        return e != hipSuccess ? e : hipStreamWaitEvent(main, ev_join, 0);
    }
    ~SideStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
    }
};

struct Work {
    DBuf<uint64_t> keys, ktmp;
    DBuf<int32_t> vals, vtmp, flags, start, stmp;
    DBuf<uint32_t> hist;
    hipStream_t st;
    int sort(int64_t m, int bits) {
        int rc;
        if ((rc = hist.ensure(sort_hist_len(m)))) return rc;
        return sort_pairs_u64(keys.p, vals.p, ktmp.p, vtmp.p, hist.p, m, bits, st);
    }
    // records, sort scratch and run buffers for up to m records
    int reserve(int64_t m) {
        int rc;
        if ((rc = keys.ensure(m)) || (rc = ktmp.ensure(m)) || (rc = vals.ensure(m)) ||
            (rc = vtmp.ensure(m)) || (rc = hist.ensure(sort_hist_len(m))) ||
            (rc = flags.ensure(m + 1)) || (rc = start.ensure(m + 1)) ||
            (rc = stmp.ensure(scan_tmp_len(m + 1))))
            return rc;
        return CAGR_OK;
    }
    // runs of equal keys below `none` in keys[0, m): start.p[0..U] and, if
    // ukeys, their keys; returns U through *U
    int runs(int64_t m, uint64_t none, uint64_t* ukeys, int32_t* U) {
        int rc;
        if ((rc = flags.ensure(m + 1)) || (rc = start.ensure(m + 1)) ||
            (rc = stmp.ensure(scan_tmp_len(m + 1))))
            return rc;
        k_heads<<<nblk(m + 1), kT, 0, st>>>(keys.p, m, none, flags.p);
        CAGR_LAUNCHED();
        if ((rc = scan_exclusive<int32_t>(flags.p, m + 1, stmp.p, st))) return rc;
        k_starts<<<nblk(m), kT, 0, st>>>(keys.p, m, none, flags.p, start.p, ukeys);
        CAGR_LAUNCHED();
        CAGR_HIP(hipMemcpyAsync(U, flags.p + m, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        return CAGR_OK;
    }
};

int pgr_build(cagr_pgr* g, const float* pts, int64_t n, int D, int dmin, hipStream_t st) {
    int rc;
    const int S = D - 2;
    Work w;
    w.st = st;
    CAGR_HIP(hipMalloc(&g->samples, 3 * n * sizeof(float)));

    // bounding box, scale and centre (Octree.cpp:213-253), then the samples in
    // [0,1]^3 (:256-259)
    {
        const int nb = (int)std::min<int64_t>(1024, nblk(n));
        DBuf<float> part;
        if ((rc = part.ensure(6 * nb + 6))) return rc;
        k_bbox_part<<<nb, kT, 0, st>>>(pts, n, part.p);
        CAGR_LAUNCHED();
        k_bbox_final<<<1, 64, 0, st>>>(part.p, nb, part.p + 6 * nb);
        CAGR_LAUNCHED();
        float bb[6];
        CAGR_HIP(hipMemcpyAsync(bb, part.p + 6 * nb, sizeof(bb), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        const float ext[3] = {bb[3] - bb[0], bb[4] - bb[1], bb[5] - bb[2]};
        g->scale = std::max(std::max(ext[0], ext[1]), ext[2]) * 1.1f;
        for (int k = 0; k < 3; ++k) g->center[k] = (bb[3 + k] + bb[k] - g->scale) / 2;
        for (int k = 0; k < 6; ++k) g->bb[k] = bb[k];
        CAGR_REQUIRE(g->scale > 0.f && std::isfinite(g->scale), CAGR_ERR_ARG,
                     "the samples must span a finite, non-empty box");
        k_normalize<<<nblk(3 * n), kT, 0, st>>>(pts, n, g->center[0], g->center[1], g->center[2],
                                                 g->scale, g->samples);
        CAGR_LAUNCHED();
    }
    const float* p = g->samples;

    // centre weights per depth 0..S: records (cell, sample*27 + slot), stable
    // sort by cell, kept per depth; then one serial sum per cell.  The shallow
    // depths hold the long serial chains (the root's has N terms): their sums
    // start on a second stream as soon as they are sorted and run beside the
    // sorts of the deeper depths, whose sums follow in one launch.
    Levels L;
    DBuf<uint64_t> ukeys;
    DBuf<uint64_t> lk[kMaxD + 1];
    DBuf<int32_t> allv;
    DBuf<int32_t> sd[kMaxD + 1];
    DBuf<Seg> segs[2];
    DBuf<float> cw[2];
    int64_t seg_off[kMaxD + 2] = {0};
    const int64_t per_depth = 27 * n;
    const int early = std::min(3, S + 1);   // depths summed on the side stream
    SideStream side;
    // (every buffer the later passes need is allocated here: a hipFree would
    // wait for the side stream and undo the overlap)
    if ((rc = side.create(st)) || (rc = w.reserve(54 * n)) ||
        (rc = allv.ensure((size_t)(S + 1) * per_depth)) || (rc = ukeys.ensure(54 * n)))
        return rc;
    // seg table + sums of depths [d0, d1) into group g, on stream s
    auto sums = [&](int d0, int d1, int g, hipStream_t s_tab, hipStream_t s_sum) -> int {
        const int64_t cnt = seg_off[d1] - seg_off[d0];
        int rc2;
        if ((rc2 = segs[g].ensure(cnt)) || (rc2 = cw[g].ensure(cnt))) return rc2;
        for (int d = d0; d < d1; ++d) {
            k_seg_table<<<nblk(L.U[d]), kT, 0, s_tab>>>(sd[d].p, L.U[d], d, d * per_depth,
                                                        segs[g].p + seg_off[d] - seg_off[d0]);
            CAGR_LAUNCHED();
            L.cw[d] = cw[g].p + seg_off[d] - seg_off[d0];
        }
        if (s_tab != s_sum) CAGR_HIP(side.fork());
        if (cnt) {
            k_cwc_sum<<<nblk(64 * cnt), kT, 0, s_sum>>>(p, allv.p, per_depth, segs[g].p,
                                                        (int32_t)cnt, cw[g].p);
            CAGR_LAUNCHED();
        }
        return CAGR_OK;
    };
    for (int d = 0; d <= S; ++d) {
        const int64_t m = per_depth;
        k_cwc_recs<<<nblk(m), kT, 0, st>>>(p, n, d, w.keys.p, w.vals.p);
        CAGR_LAUNCHED();
        if ((rc = w.sort(m, 3 * d + 1))) return rc;
        int32_t U = 0;
        if ((rc = w.runs(m, 1ull << (3 * d), ukeys.p, &U))) return rc;
        if ((rc = lk[d].ensure(U)) || (rc = sd[d].ensure((size_t)U + 1))) return rc;
        CAGR_HIP(hipMemcpyAsync(lk[d].p, ukeys.p, U * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                st));
        CAGR_HIP(hipMemcpyAsync(sd[d].p, w.start.p, ((size_t)U + 1) * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, st));
        CAGR_HIP(hipMemcpyAsync(allv.p + d * per_depth, w.vals.p, m * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, st));
        seg_off[d + 1] = seg_off[d] + U;
        CAGR_REQUIRE(seg_off[d + 1] < (int64_t(1) << 31), CAGR_ERR_ARG,
                     "too many cells for 32-bit indices");
        L.uk[d] = lk[d].p;
        L.U[d] = U;
        if (d + 1 == early && (rc = sums(0, early, 0, st, side.s))) return rc;
    }
    if ((rc = sums(early, S + 1, 1, st, st))) return rc;
    for (int d = S + 1; d <= kMaxD; ++d) {
        L.uk[d] = nullptr;
        L.cw[d] = nullptr;
        L.U[d] = 0;
    }

    // sample depths and normal-length splats: 54 records per sample, sorted by
    // cell, serial sum per cell, cells with |sum| > 1e-6 kept
    DBuf<uint64_t> nb_keys;
    DBuf<double> v64;
    DBuf<int32_t> todo, n_todo, flag, ftmp, dcnt;
    int32_t nnb = 0;
    {
        const int64_t m = 54 * n;
        if ((rc = v64.ensure(m))) return rc;
        // first the samples whose depth search stays at depths >= `early`
        // (their weights are ready); the rest once the side stream's
        // shallow sums are joined.  Usually no sample climbs that far, and
        // the long shallow chains then overlap everything below.
        if ((rc = todo.ensure(n)) || (rc = n_todo.ensure(1))) return rc;
        CAGR_HIP(hipMemsetAsync(n_todo.p, 0, sizeof(int32_t), st));
        k_splat_recs<<<nblk(n), kT, 0, st>>>(p, n, S, D, dmin, std::log(4.0), (float)(2 << D), L,
                                              early, false, todo.p, n_todo.p, w.keys.p, w.vals.p,
                                              v64.p);
        CAGR_LAUNCHED();
        int32_t deferred = 0;
        CAGR_HIP(hipMemcpyAsync(&deferred, n_todo.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        if (deferred) {
            CAGR_HIP(side.join());
            k_splat_recs<<<nblk(n), kT, 0, st>>>(p, n, S, D, dmin, std::log(4.0), (float)(2 << D),
                                                  L, 0, true, todo.p, n_todo.p, w.keys.p,
                                                  w.vals.p, v64.p);
            CAGR_LAUNCHED();
        }
        if ((rc = w.sort(m, kCellBits))) return rc;
        if ((rc = ukeys.ensure(m))) return rc;
        int32_t U = 0;
        if ((rc = w.runs(m, kNoCell, ukeys.p, &U))) return rc;
        if ((rc = flag.ensure((size_t)U + 1)) || (rc = ftmp.ensure(scan_tmp_len((int64_t)U + 1))))
            return rc;
        k_nl_sum<<<nblk(64 * ((int64_t)U + 1)), kT, 0, st>>>(v64.p, w.vals.p, w.start.p, U,
                                                             flag.p);
        CAGR_LAUNCHED();
        if ((rc = scan_exclusive<int32_t>(flag.p, (int64_t)U + 1, ftmp.p, st))) return rc;
        CAGR_HIP(hipMemcpyAsync(&nnb, flag.p + U, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        if ((rc = nb_keys.ensure(nnb))) return rc;
        k_compact<<<nblk(U), kT, 0, st>>>(ukeys.p, flag.p, U, nb_keys.p);
        CAGR_LAUNCHED();
    }

    // cells with normals per depth (the keys are depth-major); the tree height
    // is the deepest of them
    int32_t nb_off[17] = {0};
    {
        if ((rc = dcnt.ensure(17))) return rc;
        k_depth_off<<<1, 64, 0, st>>>(nb_keys.p, nnb, dcnt.p);
        CAGR_LAUNCHED();
        CAGR_HIP(hipMemcpyAsync(nb_off, dcnt.p, sizeof(nb_off), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
    }
    int32_t cnt[16];
    int H = 0;
    for (int d = 0; d < 16; ++d) {
        cnt[d] = nb_off[d + 1] - nb_off[d];
        if (cnt[d]) H = d;
    }

    // internal nodes, deepest level first: I[d-1] = unique parents of the
    // cells with normals at depth d and of I[d]
    DBuf<uint64_t> I[kMaxD + 2];
    int32_t nI[kMaxD + 2] = {0};
    for (int d = H; d >= 1; --d) {
        const int32_t m = cnt[d] + nI[d];
        if ((rc = w.reserve(m))) return rc;
        k_parents<<<nblk(m), kT, 0, st>>>(nb_keys.p + nb_off[d], cnt[d], I[d].p, nI[d], w.keys.p,
                                           w.vals.p);
        CAGR_LAUNCHED();
        if ((rc = w.sort(m, 39))) return rc;
        if ((rc = I[d - 1].ensure(m))) return rc;
        if ((rc = w.runs(m, kNoCell, I[d - 1].p, &nI[d - 1]))) return rc;
    }

    // leaves in preorder
    int64_t nl = 1;
    DBuf<int32_t> leaves;
    if (H == 0) {
        if ((rc = leaves.ensure(4))) return rc;
        CAGR_HIP(hipMemsetAsync(leaves.p, 0, 4 * sizeof(int32_t), st));
    } else {
        int64_t R = 0, inner = 0;
        for (int d = 0; d < H; ++d) {
            R += 8 * (int64_t)nI[d];
            if (d >= 1) inner += nI[d];
        }
        nl = R - inner;
        CAGR_REQUIRE(R < (int64_t(1) << 31) && 8 * nl < (int64_t(1) << 31), CAGR_ERR_ARG,
                     "PGR octree too large for 32-bit record indices");
        if ((rc = w.reserve(R))) return rc;
        const int lbits = 3 * H + 4;
        const uint64_t none = 1ull << lbits;
        int64_t base = 0;
        for (int d = 0; d < H; ++d) {
            const int64_t m = 8 * (int64_t)nI[d];
            k_children<<<nblk(m), kT, 0, st>>>(I[d].p, nI[d], I[d + 1].p, nI[d + 1], H, none,
                                                w.keys.p + base, w.vals.p + base);
            CAGR_LAUNCHED();
            base += m;
        }
        if ((rc = w.sort(R, lbits + 1))) return rc;
        if ((rc = leaves.ensure(4 * nl))) return rc;
        k_leaf_decode<<<nblk(nl), kT, 0, st>>>(w.keys.p, nl, H, leaves.p);
        CAGR_LAUNCHED();
    }

    // corners of the leaves in order of first occurrence (gridDataVector)
    const int md = H + 1;
    const int64_t R = 8 * nl;
    if ((rc = w.reserve(R)) || (rc = w.flags.ensure(R + 1)) ||
        (rc = w.stmp.ensure(scan_tmp_len(R + 1))))
        return rc;
    k_corner_keys<<<nblk(R + 1), kT, 0, st>>>(leaves.p, R, md, w.keys.p, w.vals.p, w.flags.p);
    CAGR_LAUNCHED();
    if ((rc = w.sort(R, md + 32))) return rc;
    k_corner_first<<<nblk(R), kT, 0, st>>>(w.keys.p, w.vals.p, R, w.flags.p);
    CAGR_LAUNCHED();
    if ((rc = scan_exclusive<int32_t>(w.flags.p, R + 1, w.stmp.p, st))) return rc;
    int32_t nc = 0;
    CAGR_HIP(hipMemcpyAsync(&nc, w.flags.p + R, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    CAGR_HIP(hipMalloc(&g->corners, 3 * std::max<int64_t>(nc, 1) * sizeof(float)));
    const float grid = (float)(1.0f / std::pow(2.0, md + 1));
    k_corner_write<<<nblk(R), kT, 0, st>>>(leaves.p, w.flags.p, R, md, grid, g->corners);
    CAGR_LAUNCHED();

    // kept for the mesh: the corner keys in key order with their
    // gridDataVector positions, and the internal nodes (depth-major, sorted)
    {
        DBuf<int32_t> rank;
        if ((rc = rank.ensure(R + 1)) || (rc = w.stmp.ensure(scan_tmp_len(R + 1)))) return rc;
        k_heads<<<nblk(R + 1), kT, 0, st>>>(w.keys.p, R, ~0ull, rank.p);
        CAGR_LAUNCHED();
        if ((rc = scan_exclusive<int32_t>(rank.p, R + 1, w.stmp.p, st))) return rc;
        CAGR_HIP(hipMalloc(&g->ck, std::max<int64_t>(nc, 1) * sizeof(int64_t)));
        CAGR_HIP(hipMalloc(&g->ck_idx, std::max<int64_t>(nc, 1) * sizeof(int32_t)));
        k_corner_index<<<nblk(R), kT, 0, st>>>(w.keys.p, w.vals.p, R, w.flags.p, rank.p, g->ck,
                                                g->ck_idx);
        CAGR_LAUNCHED();
        int64_t n_in = 0;
        for (int d = 0; d < H; ++d) n_in += nI[d];
        CAGR_HIP(hipMalloc(&g->in_keys, std::max<int64_t>(n_in, 1) * sizeof(uint64_t)));
        int64_t at = 0;
        for (int d = 0; d < H; ++d) {
            if (nI[d])
                CAGR_HIP(hipMemcpyAsync(g->in_keys + at, I[d].p, nI[d] * sizeof(uint64_t),
                                        hipMemcpyDeviceToDevice, st));
            at += nI[d];
        }
        g->n_in = n_in;
        CAGR_HIP(hipStreamSynchronize(st));
    }

    g->leaves = leaves.p;
    leaves.p = nullptr;
    g->n_leaves = nl;
    g->n_corners = nc;
    g->depth = md;
    return CAGR_OK;
}

}  // namespace
}  // namespace cagr

using namespace cagr;

extern "C" {

void cagr_pgr_destroy(cagr_pgr* g) {
    if (!g) return;
    int cur = 0;
    const bool sw = hipGetDevice(&cur) == hipSuccess && cur != g->device;
    if (sw) (void)hipSetDevice(g->device);
    if (g->samples) (void)hipFree(g->samples);
    if (g->corners) (void)hipFree(g->corners);
    if (g->leaves) (void)hipFree(g->leaves);
    if (g->in_keys) (void)hipFree(g->in_keys);
    if (g->ck) (void)hipFree(g->ck);
    if (g->ck_idx) (void)hipFree(g->ck_idx);
    if (g->mesh_v) (void)hipFree(g->mesh_v);
    if (g->mesh_t) (void)hipFree(g->mesh_t);
    if (sw) (void)hipSetDevice(cur);
    delete g;
}

int cagr_pgr_build(const float* points, int64_t n, int max_depth, int min_depth, void* stream,
                   cagr_pgr** out) {
    CAGR_REQUIRE(out != nullptr, CAGR_ERR_ARG, "null output handle pointer");
    *out = nullptr;
    CAGR_REQUIRE(points != nullptr, CAGR_ERR_ARG, "null points");
    CAGR_REQUIRE(n >= 1 && n <= (int64_t(1) << 25), CAGR_ERR_ARG,
                 "number of samples must be in [1, 2^25]");
    CAGR_REQUIRE(max_depth >= 3 && max_depth <= kMaxD, CAGR_ERR_ARG,
                 "max_depth must be in [3, 13]");
    CAGR_REQUIRE(min_depth >= 0 && min_depth <= max_depth, CAGR_ERR_ARG,
                 "min_depth must be in [0, max_depth]");
    cagr_pgr* g = new (std::nothrow) cagr_pgr;
    CAGR_REQUIRE(g != nullptr, CAGR_ERR_STATE, "out of host memory");
    CAGR_HIP(hipGetDevice(&g->device));
    g->n = n;
    const int rc = pgr_build(g, points, n, max_depth, min_depth, (hipStream_t)stream);
    if (rc) {
        cagr_pgr_destroy(g);
        return rc;
    }
    *out = g;
    return CAGR_OK;
}

int cagr_pgr_info(const cagr_pgr* g, int64_t* n_leaves, int64_t* n_corners, int* depth,
                  float* scale_center) {
    CAGR_REQUIRE(g != nullptr, CAGR_ERR_ARG, "null PGR octree handle");
    if (n_leaves) *n_leaves = g->n_leaves;
    if (n_corners) *n_corners = g->n_corners;
    if (depth) *depth = g->depth;
    if (scale_center) {
        scale_center[0] = g->scale;
        for (int k = 0; k < 3; ++k) scale_center[1 + k] = g->center[k];
    }
    return CAGR_OK;
}

int cagr_pgr_box(const cagr_pgr* g, float* box) {
    CAGR_REQUIRE(g != nullptr && box != nullptr, CAGR_ERR_ARG, "null argument");
    for (int k = 0; k < 6; ++k) box[k] = g->bb[k];
    return CAGR_OK;
}

int cagr_pgr_read(const cagr_pgr* g, float* samples, float* corners, int32_t* leaves,
                  void* stream) {
    CAGR_REQUIRE(g != nullptr, CAGR_ERR_ARG, "null PGR octree handle");
    const hipStream_t st = (hipStream_t)stream;
    if (samples)
        CAGR_HIP(hipMemcpyAsync(samples, g->samples, 3 * g->n * sizeof(float),
                                hipMemcpyDeviceToDevice, st));
    if (corners)
        CAGR_HIP(hipMemcpyAsync(corners, g->corners, 3 * g->n_corners * sizeof(float),
                                hipMemcpyDeviceToDevice, st));
    if (leaves)
        CAGR_HIP(hipMemcpyAsync(leaves, g->leaves, 4 * g->n_leaves * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, st));
    return CAGR_OK;
}

}  // extern "C"
