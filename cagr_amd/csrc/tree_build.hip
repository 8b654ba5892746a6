// Octree build on the GPU (gfx950): exact Morton keys, a stable LSD radix
// sort, and the preorder topology from the sorted keys.  The per-thread
// bodies live in tree_logic.h (shared with the CPU replay test); this file
// holds the kernels, the radix sort and the host orchestration.
//
// Reference behaviour reproduced: wn_treecode_cpu_treeutils.cpp:90-258 and
// wn_treecode_cpu_torch_interface.cpp:41-102 (root box, leaf rule, preorder
// ids, ascending point lists).
#include "common.h"
#include "tree_export.h"
#include <algorithm>
#include <functional>

namespace cagr {

namespace {

constexpr int kThreads = 256;

inline int64_t blocks_for(int64_t n, int per = kThreads) { return (n + per - 1) / per; }

// ---------------------------------------------------------------- keys
// coff (device, [nclouds + 1]) may be null for a single cloud
template <typename T>
__global__ __launch_bounds__(kThreads) void k_keys(const T* __restrict__ pts, int64_t n,
                                                   const int64_t* __restrict__ coff,
                                                   int64_t nclouds, uint64_t* __restrict__ keys,
                                                   int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t cloud = 0;
    if (coff) {   // last c with coff[c] <= i
        int64_t lo = 0, hi = nclouds - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) / 2;
            if (coff[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        cloud = (uint64_t)lo;
    }
    keys[i] = forest_key(cloud, morton_key(cell_index(pts[3 * i]), cell_index(pts[3 * i + 1]),
                                           cell_index(pts[3 * i + 2])));
    idx[i] = (int32_t)i;
}

// node id range [roots[c], roots[c+1]) of every cloud
__global__ void k_roots(const int64_t* __restrict__ coff, int64_t nclouds,
                        const int32_t* __restrict__ off, int32_t* __restrict__ roots) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= nclouds) roots[c] = off[coff[c]];
}

// ---------------------------------------------------------------- radix sort
// LSD, 8-bit digits, tiles of 2048 keys per 256-thread block.  Stability:
// items of a tile are ranked in (round, thread) order; within a wave by a
// 64-lane digit match built from 8 ballots; waves by an LDS prefix; tiles by
// the bin-major scan of the per-tile histograms.
constexpr int kRsItems = 8;
constexpr int kRsTile = kThreads * kRsItems;

__global__ __launch_bounds__(kThreads) void k_rs_hist(const uint64_t* __restrict__ keys,
                                                      int64_t n, int shift,
                                                      uint32_t* __restrict__ hist,
                                                      int64_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        const int64_t i = base + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// ---------------------------------------------------------------- scan
// Exclusive prefix sum over a device array: tile sums (one block per 4096
// elements, coalesced loads), a scan of the tile sums (recursively), then a
// scan of each tile offset by its tile's prefix.  Tiles pass through LDS so
// every global access is coalesced.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanThreads * kScanPer;

constexpr __host__ __device__ int scan_pos(int i) { return i + (i >> 5); }   // skew: no LDS bank clashes

// exclusive scan of one value per thread over the 256-thread block; *total
// receives the block total
template <typename U>
__device__ inline U block_scan_excl(U v, U* wsum, U& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    U x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const U y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    U pre = 0;
    total = 0;
    for (int k = 0; k < kScanThreads / 64; ++k) {
        if (k < w) pre += wsum[k];
        total += wsum[k];
    }
    __syncthreads();
    return pre + x - v;
}

template <typename U>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const U* __restrict__ a, int64_t len,
                                                            U* __restrict__ sums) {
    __shared__ U wsum[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    U s = 0;
    for (int r = 0; r < kScanPer; ++r) {
        const int64_t i = base + r * kScanThreads + threadIdx.x;
        if (i < len) s += a[i];
    }
    U total;
    (void)block_scan_excl(s, wsum, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename U>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(U* __restrict__ a, int64_t len,
                                                             const U* __restrict__ offs) {
    __shared__ U tile[scan_pos(kScanTile)];
    __shared__ U wsum[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const int t = threadIdx.x;
    for (int r = 0; r < kScanPer; ++r) {
        const int j = r * kScanThreads + t;
        tile[scan_pos(j)] = base + j < len ? a[base + j] : (U)0;
    }
    __syncthreads();
    U v[kScanPer];
    U s = 0;
    for (int r = 0; r < kScanPer; ++r) {
        v[r] = tile[scan_pos(t * kScanPer + r)];
        s += v[r];
    }
    U total;
    U run = block_scan_excl(s, wsum, total) + (offs ? offs[blockIdx.x] : (U)0);
    for (int r = 0; r < kScanPer; ++r) {
        tile[scan_pos(t * kScanPer + r)] = run;
        run += v[r];
    }
    __syncthreads();
    for (int r = 0; r < kScanPer; ++r) {
        const int j = r * kScanThreads + t;
        if (base + j < len) a[base + j] = tile[scan_pos(j)];
    }
}

__global__ __launch_bounds__(kThreads) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, int32_t* __restrict__ vout, int64_t n, int shift,
    const uint32_t* __restrict__ offs, int64_t nblocks) {
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t gbase[256];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    run[t] = 0;
    gbase[t] = offs[(int64_t)t * nblocks + blockIdx.x];
    const uint64_t lanes_below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        wcnt[0][t] = 0;
        wcnt[1][t] = 0;
        wcnt[2][t] = 0;
        wcnt[3][t] = 0;
        __syncthreads();
        const int64_t i = base + r * kThreads + t;
        const bool valid = i < n;
        const uint64_t key = valid ? kin[i] : 0ull;
        const int32_t val = valid ? vin[i] : 0;
        const uint32_t dig = (uint32_t)(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lanes_below);
        if (valid && rank == 0) wcnt[w][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t acc = run[t];
        for (int k = 0; k < 4; ++k) {
            const uint32_t c = wcnt[k][t];
            wcnt[k][t] = acc;
            acc += c;
        }
        run[t] = acc;
        __syncthreads();
        if (valid) {
            const uint32_t dst = gbase[dig] + wcnt[w][dig] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- topology
__global__ __launch_bounds__(kThreads) void k_chain(const uint64_t* __restrict__ keys, int64_t n,
                                                    int S, int max_depth, int8_t* __restrict__ ld,
                                                    int8_t* __restrict__ sd,
                                                    int32_t* __restrict__ cnt) {
    const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b < n) chain_body(b, keys, n, S, max_depth, ld, sd, cnt);
    if (b == n) cnt[n] = 0;
}

__global__ __launch_bounds__(kThreads) void k_nodes(const uint64_t* __restrict__ keys, int64_t n,
                                                    const int8_t* __restrict__ ld,
                                                    const int8_t* __restrict__ sd,
                                                    const int32_t* __restrict__ off,
                                                    double open_factor,
                                                    TopoRec* __restrict__ topo) {
    const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b < n) node_body(b, keys, n, ld, sd, off, open_factor, topo);
}

__global__ __launch_bounds__(kThreads) void k_leaf_sort(const TopoRec* __restrict__ topo,
                                                        int64_t m, int32_t* __restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const TopoRec t = topo[i];
    if (topo_leaf(t)) leaf_sort_body(t, perm);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather_points(const T* __restrict__ pts,
                                                            const int32_t* __restrict__ perm,
                                                            int64_t n, double* __restrict__ p64,
                                                            float* __restrict__ p32) {
    const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    const int64_t j = perm[s];
    for (int d = 0; d < 3; ++d) {
        const T v = pts[3 * j + d];
        p64[3 * s + d] = (double)v;
        p32[3 * s + d] = (float)v;
    }
}

__global__ __launch_bounds__(kThreads) void k_leaf_of(const TopoRec* __restrict__ topo, int64_t m,
                                                      int32_t* __restrict__ leaf_of) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const TopoRec t = topo[i];
    if (topo_leaf(t))
        for (int32_t s = t.begin; s < t.end; ++s) leaf_of[s] = (int32_t)i;
}

// bounding sphere of every leaf's points: centre = middle of their bounding
// box (rounded to float), radius = the largest FP64 distance to that centre,
// rounded up; used only to prove that a whole leaf lies outside a width sphere
// Largest |p64 - (double)p32| over the coordinates into *out (as the bits
// of a non-negative double, which order like the values; NaN never wins):
// how far the float copies the FP32 walks take as queries are off
__global__ __launch_bounds__(kThreads) void k_round_err(const double* __restrict__ p64,
                                                        const float* __restrict__ p32, int64_t len,
                                                        unsigned long long* __restrict__ out) {
    double e = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads) {
        const double v = fabs(p64[i] - (double)p32[i]);
        if (v > e) e = v;
    }
    for (int s = 32; s > 0; s >>= 1) {
        const double o = __shfl_xor(e, s, 64);
        if (o > e) e = o;
    }
    if ((threadIdx.x & 63) == 0 && e > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}

// bounding sphere of each leaf's FP64 points, its radius widened by 2 qr
// (the FP32 walks test it with a float query copy off by up to qr per
// component)
__global__ __launch_bounds__(kThreads) void k_leaf_bounds(const TopoRec* __restrict__ topo,
                                                          int64_t m, const double* __restrict__ p64,
                                                          const unsigned long long* __restrict__ qbits,
                                                          float4* __restrict__ lb) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const TopoRec t = topo[i];
    if (!topo_leaf(t)) {
        lb[i] = make_float4(0.f, 0.f, 0.f, 3.0e38f);
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t s = t.begin; s < t.end; ++s)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], p64[3 * s + d]);
            hi[d] = fmax(hi[d], p64[3 * s + d]);
        }
    float c[3];
    for (int d = 0; d < 3; ++d) c[d] = (float)(0.5 * (lo[d] + hi[d]));
    double r2 = 0.0;
    for (int32_t s = t.begin; s < t.end; ++s) {
        double q = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double e = p64[3 * s + d] - (double)c[d];
            q += e * e;
        }
        r2 = fmax(r2, q);
    }
    const double qr = __longlong_as_double((long long)*qbits);
    const float r = (float)((sqrt(r2) + 2.0 * qr) * (1.0 + 1e-6)) + 1e-7f;
    // a leaf with a non-finite point never qualifies
    lb[i] = make_float4(c[0], c[1], c[2], (r2 == r2 && r2 < 1e300) ? r : 3.0e38f);
}

// nodes keyed by depth for a stable sort (preorder within a depth); leaf
// count and deepest leaf, aggregated per block
__global__ __launch_bounds__(kThreads) void k_depth_keys(const TopoRec* __restrict__ topo, int64_t m,
                                                         uint64_t* __restrict__ keys,
                                                         int32_t* __restrict__ vals,
                                                         int32_t* __restrict__ stats) {
    __shared__ int32_t s_leaves, s_depth;
    if (threadIdx.x == 0) { s_leaves = 0; s_depth = 0; }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < m) {
        const TopoRec t = topo[i];
        keys[i] = (uint64_t)topo_depth(t);
        vals[i] = (int32_t)i;
        if (topo_leaf(t)) {
            atomicAdd(&s_leaves, 1);
            atomicMax(&s_depth, topo_depth(t));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&stats[0], s_leaves);
        atomicMax(&stats[1], s_depth);
    }
}

// first index of each depth in the depth-sorted keys: off[d], d in
// [0, kLevels + 1] (off[kLevels + 1] = m, the end of the deepest depth)
__global__ void k_depth_off(const uint64_t* __restrict__ keys, int64_t m, int32_t* __restrict__ off) {
    const int d = threadIdx.x;
    if (d > kLevels + 1) return;
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < (uint64_t)d) lo = mid + 1;
        else hi = mid;
    }
    off[d] = (int32_t)lo;
}

// child positions of the internal nodes (kids_body), one thread per position
// of depth >= 1 (positions i0.. of the depth order)
__global__ __launch_bounds__(kThreads) void k_kids(TopoRec* __restrict__ topo,
                                                   const int32_t* __restrict__ order, int32_t i0,
                                                   int64_t m) {
    const int64_t i = i0 + (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < m) kids_body(i, topo, order, m);
}

// Segments of the upward launches: for root i (the id-sorted list of roots
// at depth dlo) and level l (depth dlo + l), the run of the depth-sorted
// order holding the root's subtree nodes at that depth.
__global__ __launch_bounds__(kThreads) void k_up_segments(const TopoRec* __restrict__ topo,
                                                          const int32_t* __restrict__ order,
                                                          const int32_t* __restrict__ off,
                                                          int32_t nroots, int dlo, int nlev,
                                                          int2* __restrict__ seg) {
    const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= (int64_t)nroots * nlev) return;
    const int32_t i = (int32_t)(k / nlev);
    const int l = (int)(k % nlev);
    const int32_t r = order[off[dlo] + i];
    const int32_t e = topo[r].escape;
    const int d = dlo + l;
    const int32_t* lst = order + off[d];
    const int32_t len = off[d + 1] - off[d];
    auto lower = [&](int32_t v) {
        int32_t lo = 0, hi = len;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (lst[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    seg[k] = make_int2(off[d] + lower(r), off[d] + lower(e));
}

}  // namespace

int64_t scan_tmp_len(int64_t len) {
    int64_t total = 0;
    while (len > kScanTile) {
        len = (len + kScanTile - 1) / kScanTile;
        total += len;
    }
    return total + 1;
}

template <typename U>
int scan_exclusive(U* a, int64_t len, U* tmp, hipStream_t st) {
    if (len <= 0) return CAGR_OK;
    if (len <= kScanTile) {
        k_scan_tiles<U><<<1, kScanThreads, 0, st>>>(a, len, nullptr);
        CAGR_LAUNCHED();
        return CAGR_OK;
    }
    const int64_t nt = (len + kScanTile - 1) / kScanTile;
    k_scan_sums<U><<<nt, kScanThreads, 0, st>>>(a, len, tmp);
    CAGR_LAUNCHED();
    int rc = scan_exclusive<U>(tmp, nt, tmp + nt, st);
    if (rc) return rc;
    k_scan_tiles<U><<<nt, kScanThreads, 0, st>>>(a, len, tmp);
    CAGR_LAUNCHED();
    return CAGR_OK;
}
template int scan_exclusive<int32_t>(int32_t*, int64_t, int32_t*, hipStream_t);
template int scan_exclusive<uint32_t>(uint32_t*, int64_t, uint32_t*, hipStream_t);
template int scan_exclusive<int64_t>(int64_t*, int64_t, int64_t*, hipStream_t);

int64_t sort_hist_len(int64_t n) {
    const int64_t h = 256 * std::max<int64_t>(1, (n + kRsTile - 1) / kRsTile);
    return h + scan_tmp_len(h);
}

// Sorts (keys, vals) by the low `bits` bits of keys, stably.  The result is
// left in keys/vals; *_tmp and hist (sort_hist_len(n) entries) are scratch.
int sort_pairs_u64(uint64_t* keys, int32_t* vals, uint64_t* keys_tmp, int32_t* vals_tmp,
                   uint32_t* hist, int64_t n, int bits, hipStream_t st) {
    if (n <= 1) return CAGR_OK;
    const int64_t nb = (n + kRsTile - 1) / kRsTile;
    uint64_t* ka = keys;
    uint64_t* kb = keys_tmp;
    int32_t* va = vals;
    int32_t* vb = vals_tmp;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8, ++passes) {
        k_rs_hist<<<nb, kThreads, 0, st>>>(ka, n, shift, hist, nb);
        CAGR_LAUNCHED();
        const int rcs = scan_exclusive<uint32_t>(hist, 256 * nb, hist + 256 * nb, st);
        if (rcs) return rcs;
        k_rs_scatter<<<nb, kThreads, 0, st>>>(ka, va, kb, vb, n, shift, hist, nb);
        CAGR_LAUNCHED();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (passes & 1) {
        CAGR_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        CAGR_HIP(hipMemcpyAsync(vals, va, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    }
    return CAGR_OK;
}

// Morton order of n arbitrary points (keys on the tree's root box, same cell
// index as the build): order[i] = index of the i-th point in Morton order.
size_t morton_order_scratch(int64_t n) {
    return 2 * n * sizeof(uint64_t) + n * sizeof(int32_t) + sort_hist_len(n) * sizeof(uint32_t) + 256;
}

int morton_order(const void* pts, int dtype, int64_t n, int32_t* order, void* scratch,
                 hipStream_t st) {
    if (n == 0) return CAGR_OK;
    char* p = (char*)scratch;
    uint64_t* keys = (uint64_t*)p;
    p += n * sizeof(uint64_t);
    uint64_t* ktmp = (uint64_t*)p;
    p += n * sizeof(uint64_t);
    int32_t* vtmp = (int32_t*)p;
    p += n * sizeof(int32_t);
    p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    uint32_t* hist = (uint32_t*)p;
    const int64_t nb = blocks_for(n);
    if (dtype == CAGR_F32)
        k_keys<float><<<nb, kThreads, 0, st>>>((const float*)pts, n, nullptr, 1, keys, order);
    else
        k_keys<double><<<nb, kThreads, 0, st>>>((const double*)pts, n, nullptr, 1, keys, order);
    CAGR_LAUNCHED();
    return sort_pairs_u64(keys, order, ktmp, vtmp, hist, n, kKeyBits, st);
}

int build_tree(cagr_tree* t, const void* points, int64_t n, int dtype, int max_depth,
               int leaf_size, double theta, hipStream_t st, const int64_t* cloud_off,
               int64_t nclouds) {
    t->n = n;
    t->max_depth = max_depth;
    t->leaf_size = leaf_size;
    t->theta = theta;
    t->open_factor = 2.0 / theta;
    t->point_dtype = dtype;
    CAGR_HIP(hipGetDevice(&t->device));
    // the event that orders later calls on this handle across streams
    CAGR_HIP(hipEventCreateWithFlags(&t->done_ev, hipEventDisableTiming));
    // clouds: a plain tree is a forest of one
    t->num_clouds = cloud_off ? nclouds : 1;
    t->cloud_pt.assign(t->num_clouds + 1, 0);
    if (cloud_off)
        for (int64_t c = 0; c <= nclouds; ++c) t->cloud_pt[c] = cloud_off[c];
    else
        t->cloud_pt[1] = n;
    t->cloud_node.assign(t->num_clouds + 1, 0);
    if (n == 0) {   // the reference builds no node at all for an empty cloud
        t->m = 0;
        t->depth_off.assign(kLevels + 2, 0);
        return CAGR_OK;
    }

    // scratch for the build (freed at the end)
    uint64_t *ktmp = nullptr;
    int32_t *vtmp = nullptr, *cnt = nullptr;
    uint32_t* hist = nullptr;
    int8_t *ld = nullptr, *sd = nullptr;
    auto free_scratch = [&]() {
        (void)hipFree(ktmp); (void)hipFree(vtmp); (void)hipFree(cnt); (void)hipFree(hist);
        (void)hipFree(ld); (void)hipFree(sd);
    };
    struct Guard { std::function<void()> f; ~Guard() { f(); } } guard{free_scratch};

    CAGR_HIP(hipMalloc(&t->keys, n * sizeof(uint64_t)));
    CAGR_HIP(hipMalloc(&t->perm, n * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->pts64, n * 3 * sizeof(double)));
    CAGR_HIP(hipMalloc(&t->pts32, n * 3 * sizeof(float)));
    CAGR_HIP(hipMalloc(&ktmp, n * sizeof(uint64_t)));
    CAGR_HIP(hipMalloc(&vtmp, n * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&hist, sort_hist_len(n) * sizeof(uint32_t)));
    CAGR_HIP(hipMalloc(&cnt, (n + 1) * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&ld, n));
    CAGR_HIP(hipMalloc(&sd, n));

    int64_t* coff_d = nullptr;
    struct Guard2 { int64_t** p; ~Guard2() { (void)hipFree(*p); } } guard2{&coff_d};
    int cloud_bits = 0;
    if (cloud_off) {
        CAGR_HIP(hipMalloc(&coff_d, (nclouds + 1) * sizeof(int64_t)));
        CAGR_HIP(hipMemcpyAsync(coff_d, t->cloud_pt.data(), (nclouds + 1) * sizeof(int64_t),
                                hipMemcpyHostToDevice, st));
        while ((int64_t(1) << cloud_bits) < nclouds) ++cloud_bits;
    }
    const int64_t nb = blocks_for(n);
    if (dtype == CAGR_F32)
        k_keys<float><<<nb, kThreads, 0, st>>>((const float*)points, n, coff_d, nclouds, t->keys, t->perm);
    else
        k_keys<double><<<nb, kThreads, 0, st>>>((const double*)points, n, coff_d, nclouds, t->keys, t->perm);
    CAGR_LAUNCHED();
    int rc = sort_pairs_u64(t->keys, t->perm, ktmp, vtmp, hist, n, kKeyBits + cloud_bits, st);
    if (rc) return rc;

    k_chain<<<blocks_for(n + 1), kThreads, 0, st>>>(t->keys, n, leaf_size, max_depth, ld, sd, cnt);
    CAGR_LAUNCHED();
    // the radix histogram buffer is free again and long enough for the scan's tile sums
    rc = scan_exclusive<int32_t>(cnt, n + 1, (int32_t*)hist, st);
    if (rc) return rc;
    int32_t m32 = 0;
    CAGR_HIP(hipMemcpyAsync(&m32, cnt + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    const int64_t m = m32;
    t->m = m;
    CAGR_REQUIRE(m > 0, CAGR_ERR_STATE, "tree build produced no nodes");
    // the traversal kernels address node records by 32-bit byte offsets; the
    // largest record array is 3 scalar-family records (96 bytes) per node
    CAGR_REQUIRE(m * 96 < (int64_t(1) << 32), CAGR_ERR_ARG,
                 "tree has more than 44739242 nodes (about 4.9e8 points): too large for one tree");
    if (cloud_off) {
        int32_t* roots_d = nullptr;
        CAGR_HIP(hipMalloc(&roots_d, (nclouds + 1) * sizeof(int32_t)));
        k_roots<<<(int)((nclouds + 256) / 256), 256, 0, st>>>(coff_d, nclouds, cnt, roots_d);
        std::vector<int32_t> roots(nclouds + 1);
        const hipError_t e1 = hipMemcpyAsync(roots.data(), roots_d, (nclouds + 1) * sizeof(int32_t),
                                             hipMemcpyDeviceToHost, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipFree(roots_d);
        CAGR_HIP(e1);
        CAGR_HIP(e2);
        for (int64_t c = 0; c <= nclouds; ++c) t->cloud_node[c] = roots[c];
    } else {
        t->cloud_node[1] = m;
    }

    CAGR_HIP(hipMalloc(&t->topo, m * sizeof(TopoRec)));
    CAGR_HIP(hipMalloc(&t->order, m * sizeof(int32_t)));
    k_nodes<<<nb, kThreads, 0, st>>>(t->keys, n, ld, sd, cnt, t->open_factor, t->topo);
    CAGR_LAUNCHED();
    k_leaf_sort<<<blocks_for(m), kThreads, 0, st>>>(t->topo, m, t->perm);
    CAGR_LAUNCHED();
    if (dtype == CAGR_F32)
        k_gather_points<float><<<nb, kThreads, 0, st>>>((const float*)points, t->perm, n, t->pts64, t->pts32);
    else
        k_gather_points<double><<<nb, kThreads, 0, st>>>((const double*)points, t->perm, n, t->pts64, t->pts32);
    CAGR_LAUNCHED();

    // node ids by depth (stable: preorder within a depth), the depth offsets,
    // and the segments of the two upward launches
    {
        uint64_t *dk = nullptr, *dkt = nullptr;
        int32_t *dvt = nullptr, *off_d = nullptr, *stats = nullptr;
        uint32_t* dh = nullptr;
        struct FreeD {
            uint64_t **a, **b;
            int32_t **c, **d, **e;
            uint32_t** f;
            ~FreeD() {
                (void)hipFree(*a); (void)hipFree(*b); (void)hipFree(*c); (void)hipFree(*d);
                (void)hipFree(*e); (void)hipFree(*f);
            }
        } free_d{&dk, &dkt, &dvt, &off_d, &stats, &dh};
        CAGR_HIP(hipMalloc(&dk, m * sizeof(uint64_t)));
        CAGR_HIP(hipMalloc(&dkt, m * sizeof(uint64_t)));
        CAGR_HIP(hipMalloc(&dvt, m * sizeof(int32_t)));
        CAGR_HIP(hipMalloc(&dh, sort_hist_len(m) * sizeof(uint32_t)));
        CAGR_HIP(hipMalloc(&off_d, (kLevels + 2) * sizeof(int32_t)));
        CAGR_HIP(hipMalloc(&stats, 2 * sizeof(int32_t)));
        CAGR_HIP(hipMemsetAsync(stats, 0, 2 * sizeof(int32_t), st));
        k_depth_keys<<<blocks_for(m), kThreads, 0, st>>>(t->topo, m, dk, t->order, stats);
        CAGR_LAUNCHED();
        if ((rc = sort_pairs_u64(dk, t->order, dkt, dvt, dh, m, 4, st))) return rc;
        k_depth_off<<<1, 64, 0, st>>>(dk, m, off_d);
        CAGR_LAUNCHED();
        int32_t hoff[kLevels + 2], hst[2];
        CAGR_HIP(hipMemcpyAsync(hoff, off_d, sizeof(hoff), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipMemcpyAsync(hst, stats, sizeof(hst), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        t->depth_off.assign(hoff, hoff + kLevels + 2);
        t->n_leaves = hst[0];
        t->depth = hst[1];
        // the segment kernel reads off[d] and off[d + 1] for every depth up to
        // the deepest: they must be in order and end at m
        bool ok = hoff[0] == 0 && hoff[kLevels + 1] == (int32_t)m && t->depth >= 0 &&
                  t->depth <= kLevels;
        for (int d = 0; d <= kLevels; ++d) ok = ok && hoff[d] <= hoff[d + 1];
        CAGR_REQUIRE(ok, CAGR_ERR_STATE, "tree build: inconsistent depth offsets");
        // the roots are at depth 0, so positions off[1].. have a parent and
        // position off[1] - 1 >= 0 exists
        if (m > hoff[1]) {
            CAGR_REQUIRE(hoff[1] >= 1, CAGR_ERR_STATE, "tree build: no root");
            k_kids<<<blocks_for(m - hoff[1]), kThreads, 0, st>>>(t->topo, t->order, hoff[1], m);
            CAGR_LAUNCHED();
        }
        // bands of at most 3 depths over the internal nodes (depths
        // 0..depth-1), deepest first: at most 1 + 8 + 64 nodes under each
        // band root
        auto count = [&](int d) { return t->depth_off[d + 1] - t->depth_off[d]; };
        t->up_bands.clear();
        int64_t nt = 0;
        for (int dhi = t->depth; dhi > 0;) {
            const int dlo = std::max(0, dhi - 3);
            t->up_bands.push_back(make_int4(dlo, dhi - dlo, count(dlo), (int)nt));
            nt += (int64_t)count(dlo) * (dhi - dlo);
            CAGR_REQUIRE(nt < (int64_t)INT32_MAX, CAGR_ERR_ARG,
                         "tree build: too many upward segments");
            dhi = dlo;
        }
        CAGR_HIP(hipMalloc(&t->up_seg, std::max<int64_t>(nt, 1) * sizeof(int2)));
        for (const int4& b : t->up_bands) {
            k_up_segments<<<blocks_for((int64_t)b.z * b.y), kThreads, 0, st>>>(
                t->topo, t->order, off_d, b.z, b.x, b.y, t->up_seg + b.w);
            CAGR_LAUNCHED();
        }
        CAGR_HIP(hipStreamSynchronize(st));
    }

    // wave work items: per cloud, its node range and its points 64 at a time
    {
        std::vector<int4> items;
        for (int64_t c = 0; c < t->num_clouds; ++c)
            for (int64_t q0 = t->cloud_pt[c]; q0 < t->cloud_pt[c + 1]; q0 += 64) {
                const int64_t qn = std::min<int64_t>(64, t->cloud_pt[c + 1] - q0);
                items.push_back(make_int4((int)t->cloud_node[c], (int)t->cloud_node[c + 1], (int)q0,
                                          (int)qn));
            }
        t->num_items = (int64_t)items.size();
        CAGR_HIP(hipMalloc(&t->items, items.size() * sizeof(int4)));
        CAGR_HIP(hipMalloc(&t->item_counter, kQueues * kCounterStride * sizeof(int32_t)));
        // natural order: equal contiguous chunks (each XCD walks one Morton range)
        for (int c = 0; c <= kQueues; ++c)
            t->item_bound[c] = (int32_t)(t->num_items * c / kQueues);
        CAGR_HIP(hipMemcpyAsync(t->items, items.data(), items.size() * sizeof(int4),
                                hipMemcpyHostToDevice, st));
        CAGR_HIP(hipStreamSynchronize(st));
        int cus = 0;
        CAGR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, t->device));
        t->num_cus = cus > 0 ? cus : 256;
    }

    // workspace sized for every operator family
    CAGR_HIP(hipMalloc(&t->src, n * 64));
    CAGR_HIP(hipMalloc(&t->nodes, m * 128));
    CAGR_HIP(hipMalloc(&t->w64, 3 * m * sizeof(double)));
    CAGR_HIP(hipMalloc(&t->r64, 9 * m * sizeof(double)));
    CAGR_HIP(hipMalloc(&t->m64, 9 * m * sizeof(double)));
    CAGR_HIP(hipMalloc(&t->leaf_of, n * sizeof(int32_t)));
    CAGR_HIP(hipMalloc(&t->adj_node, 3 * m * sizeof(double)));
    CAGR_HIP(hipMalloc(&t->adj_pt, 3 * n * sizeof(double)));
    k_leaf_of<<<blocks_for(m), kThreads, 0, st>>>(t->topo, m, t->leaf_of);
    CAGR_LAUNCHED();
    CAGR_HIP(hipMalloc(&t->lbound, m * sizeof(float4)));
    unsigned long long* qbits = nullptr;
    CAGR_HIP(hipMalloc(&qbits, sizeof(unsigned long long)));
    struct FreeQ {
        unsigned long long* p;
        ~FreeQ() { (void)hipFree(p); }
    } free_q{qbits};
    CAGR_HIP(hipMemsetAsync(qbits, 0, sizeof(unsigned long long), st));
    k_round_err<<<(unsigned)std::min<int64_t>(1024, blocks_for(3 * n)), kThreads, 0, st>>>(
        t->pts64, t->pts32, 3 * n, qbits);
    CAGR_LAUNCHED();
    k_leaf_bounds<<<blocks_for(m), kThreads, 0, st>>>(t->topo, m, t->pts64, qbits, t->lbound);
    CAGR_LAUNCHED();
    unsigned long long qb = 0;
    CAGR_HIP(hipMemcpyAsync(&qb, qbits, sizeof(qb), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    t->qround = __builtin_bit_cast(double, qb);
    return CAGR_OK;
}

}  // namespace cagr
