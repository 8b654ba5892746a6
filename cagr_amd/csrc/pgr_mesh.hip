// Marching cubes on the PGR octree (GaussRecon's Octree::GetMCIsoTriangles
// with fullDepthIso 0, nonLinearFit 1, triangles out; gaussrecon_src/
// Octree.cpp:518-1257), on the GPU, one leaf per thread, with the per-leaf
// logic of mc_logic.h.
//
// Vertices: every leaf with roots lists its 12 edge roots (key of the finest
// edge segment, in the reference's slot order); the valid keys are sorted
// stably, the first occurrence of each key over (leaf preorder, slot) gets
// its vertex number by a scan -- the order the reference's roots map fills
// in -- and that occurrence computes the position.
// Triangles: one pass over the leaves that can emit (persistent threads, one
// workspace each in global memory) gives every leaf its loops and the loops
// their fans and area-minimising flips; each leaf reserves room for its run
// with one atomic add, and a scan of the per-leaf counts then moves every
// run to its place, so the triangle list comes out in the reference's order.
#include <algorithm>

#include "common.h"
#include "mc_logic.h"
#include "pgr.h"

#pragma clang fp contract(off)

namespace cagr {
namespace {

using namespace mc;

constexpr int kT = 256;
inline int64_t nblk(int64_t n) { return (n + kT - 1) / kT; }

__device__ inline Node leaf_node(const int32_t* lv, int64_t r) {
    return Node{lv[4 * r], {lv[4 * r + 1], lv[4 * r + 2], lv[4 * r + 3]}};
}

__global__ __launch_bounds__(kT) void k_mc_roots(View t, const int32_t* __restrict__ leaves,
                                                 int64_t nl, uint64_t* __restrict__ keys,
                                                 int32_t* __restrict__ valid) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r > nl) return;
    if (r == nl) {
        valid[12 * nl] = 0;
        return;
    }
    int64_t rk[12];
    const int cnt = leaf_roots(t, leaf_node(leaves, r), rk, nullptr);
    for (int s = 0; s < 12; ++s) {
        const bool ok = s < cnt && rk[s] >= 0;
        keys[12 * r + s] = ok ? (uint64_t)rk[s] : ~0ull;
        valid[12 * r + s] = ok ? 1 : 0;
    }
}

// the valid root records (key, occurrence), in occurrence order
__global__ __launch_bounds__(kT) void k_mc_compact(const uint64_t* __restrict__ keys,
                                                   const int32_t* __restrict__ pos, int64_t n12,
                                                   uint64_t* __restrict__ ckeys,
                                                   int32_t* __restrict__ cocc) {
    const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (q >= n12 || pos[q + 1] == pos[q]) return;
    ckeys[pos[q]] = keys[q];
    cocc[pos[q]] = (int32_t)q;
}

// heads of the sorted valid keys: first[occ] = 1 at each key's first occurrence
__global__ __launch_bounds__(kT) void k_mc_first(const uint64_t* __restrict__ keys,
                                                 const int32_t* __restrict__ occ, int64_t m,
                                                 int32_t* __restrict__ first,
                                                 int32_t* __restrict__ head) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r > m) return;
    if (r == m) {
        head[m] = 0;
        return;
    }
    const bool valid = keys[r] != ~0ull;
    const bool h = valid && (r == 0 || keys[r - 1] != keys[r]);
    head[r] = h ? 1 : 0;
    if (h) first[occ[r]] = 1;
}

__global__ __launch_bounds__(kT) void k_fill(int32_t* __restrict__ a, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r < n) a[r] = 0;
}

// vertex keys in key order with their vertex numbers
__global__ __launch_bounds__(kT) void k_mc_vkeys(const uint64_t* __restrict__ keys,
                                                 const int32_t* __restrict__ occ, int64_t m,
                                                 const int32_t* __restrict__ hrank,
                                                 const int32_t* __restrict__ pos,
                                                 int64_t* __restrict__ vkeys,
                                                 int32_t* __restrict__ vidx) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= m || hrank[r + 1] == hrank[r]) return;
    vkeys[hrank[r]] = (int64_t)keys[r];
    vidx[hrank[r]] = pos[occ[r]];
}

__global__ __launch_bounds__(kT) void k_mc_vertices(View t, const int32_t* __restrict__ leaves,
                                                    const int32_t* __restrict__ pos, int64_t n12,
                                                    float* __restrict__ verts) {
    const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (q >= n12 || pos[q + 1] == pos[q]) return;
    const int s = (int)(q % 12);
    const int o = s >> 2, i = (s >> 1) & 1, j = s & 1;
    RootInfo ri;
    if (!root_index(t, leaf_node(leaves, q / 12), i | (j << 1) | (o << 2), ri)) return;
    float p[3];
    root_position(t, ri, p);
    float* v = verts + 3 * (int64_t)pos[q];
    v[0] = p[0];
    v[1] = p[1];
    v[2] = p[2];
}

// the leaves' left-aligned Morton keys (ascending: preorder) and every
// node's marching-cubes index from its corner values, once
__global__ __launch_bounds__(kT) void k_leaf_morton(View t, const int32_t* __restrict__ leaves,
                                                    int64_t nl, int H, uint64_t* __restrict__ mk,
                                                    uint8_t* __restrict__ mc) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= nl) return;
    const Node n = leaf_node(leaves, r);
    mk[r] = morton_left(n, H);
    mc[r] = (uint8_t)mc_from_corners(t, n);
}

// open-addressing table of `keys` (slot = key + 1) -> position
__global__ __launch_bounds__(kT) void k_hash_build(const uint64_t* __restrict__ keys, int64_t n,
                                                   uint64_t mask,
                                                   unsigned long long* __restrict__ slot,
                                                   int32_t* __restrict__ val) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= n) return;
    const unsigned long long k = keys[r] + 1;
    for (uint64_t h = hash_mix(keys[r]) & mask;; h = (h + 1) & mask) {
        const unsigned long long prev = atomicCAS(&slot[h], 0ull, k);
        if (prev == 0ull || prev == k) {
            val[h] = (int32_t)r;
            return;
        }
    }
}

__global__ __launch_bounds__(kT) void k_inner_mc(View t, int64_t n_in, uint8_t* __restrict__ mc) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= n_in) return;
    const uint64_t k = t.in_keys[r];
    const Node n{(int)(k >> 39), {(int32_t)(k & 0x1FFF), (int32_t)((k >> 13) & 0x1FFF),
                                  (int32_t)((k >> 26) & 0x1FFF)}};
    mc[r] = (uint8_t)mc_from_corners(t, n);
}

// leaves that can emit triangles: those with roots, and the coarser leaf
// covering each face-neighbour cell of one (GetMCIsoEdges gathers a coarse
// leaf's edges from the finer leaves across its faces); a superset is fine
__global__ __launch_bounds__(kT) void k_mc_active(View t, const int32_t* __restrict__ leaves,
                                                  int64_t nl, int H,
                                                  const uint64_t* __restrict__ mk,
                                                  const int32_t* __restrict__ valid,
                                                  int32_t* __restrict__ active) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= nl) return;
    bool roots = false;
    for (int s = 0; s < 12; ++s) roots = roots || valid[12 * r + s + 1] != valid[12 * r + s];
    if (!roots) return;
    active[r] = 1;
    const Node n = leaf_node(leaves, r);
    const int64_t lim = (int64_t)1 << n.d;
    for (int f = 0; f < 6; ++f) {
        Node c = n;
        c.o[f >> 1] += (f & 1) ? 1 : -1;
        if (c.o[f >> 1] < 0 || c.o[f >> 1] >= lim) continue;
        // the deepest existing ancestor-or-self of the neighbour cell
        while (c.d > 0 && !is_internal(t, parent(c))) c = parent(c);
        if (c.d == n.d) continue;                       // same depth: not coarser
        const int64_t i = t.lf_slot ? hash_find(t.lf_slot, t.lf_val, t.lf_mask, morton_left(c, H))
                                    : lower_bound(mk, nl, morton_left(c, H));
        if (i >= 0 && i < nl) active[i] = 1;
    }
}

__global__ __launch_bounds__(kT) void k_compact_ids(const int32_t* __restrict__ pos, int64_t n,
                                                    int32_t* __restrict__ ids) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r < n && pos[r + 1] != pos[r]) ids[pos[r]] = (int32_t)r;
}

struct DevVIndex {
    const int64_t* vkeys;
    const int32_t* vidx;
    int64_t nvk;
    const float* verts;
    __device__ int operator()(int64_t k) const {
        const int64_t i = lower_bound(vkeys, nvk, k);
        return (i < nvk && vkeys[i] == k) ? vidx[i] : -1;
    }
    __device__ void position(int v, float p[3]) const {
        for (int k = 0; k < 3; ++k) p[k] = verts[3 * (int64_t)v + k];
    }
};

// the triangles of every leaf that can emit, in one pass: each leaf builds
// them in its workspace, reserves room in `tmp` with one atomic add and
// copies them there; counts and reservations are kept per leaf so that
// k_mc_place can move every leaf's run to its place in preorder.  Threads
// persist over the leaves (grid stride), one workspace each; a block is one
// wave, whose 64 workspaces are interleaved lane by lane in `ws`.
__global__ __launch_bounds__(64) void k_mc_tris(View t, const int32_t* __restrict__ leaves,
                                                const int32_t* __restrict__ ids, int64_t na,
                                                DevVIndex vi, unsigned char* __restrict__ ws,
                                                int32_t* __restrict__ counts,
                                                int64_t* __restrict__ base,
                                                unsigned long long* __restrict__ used,
                                                int64_t cap, int32_t* __restrict__ tmp,
                                                int32_t* __restrict__ err) {
    const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int64_t nthreads = (int64_t)gridDim.x * 64;
    static_assert(kStride == 64, "one wave's workspaces are interleaved");
    Work w(ws + (size_t)blockIdx.x * 64 * WorkLayout::bytes, threadIdx.x);
    for (int64_t a = tid; a < na; a += nthreads) {
        int k = 0;
        const int c = leaf_triangles(t, leaf_node(leaves, ids[a]), w, vi,
                                            [&](int x, int y, int z) {
                                                if (k < kMaxEdges) {
                                                    w.out[3 * k] = x;
                                                    w.out[3 * k + 1] = y;
                                                    w.out[3 * k + 2] = z;
                                                }
                                                ++k;
                                            });
        if (c < 0 || c > kMaxEdges) {
            atomicOr(err, 1);
            counts[a] = 0;
            base[a] = 0;
            continue;
        }
        counts[a] = c;
        const int64_t b = c ? (int64_t)atomicAdd(used, (unsigned long long)c) : 0;
        base[a] = b;
        if (b + c <= cap)
            for (int q = 0; q < 3 * c; ++q) tmp[3 * b + q] = w.out[q];
    }
}

// every leaf's run of triangles from its reservation to its place in preorder
__global__ __launch_bounds__(kT) void k_mc_place(const int32_t* __restrict__ counts,
                                                 const int64_t* __restrict__ base,
                                                 const int64_t* __restrict__ offs, int64_t na,
                                                 const int32_t* __restrict__ tmp,
                                                 int32_t* __restrict__ tris) {
    const int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (a >= na) return;
    const int64_t from = 3 * base[a], to = 3 * offs[a];
    for (int q = 0; q < 3 * counts[a]; ++q) tris[to + q] = tmp[from + q];
}

__global__ __launch_bounds__(kT) void k_widen(const int32_t* __restrict__ c, int64_t n,
                                              int64_t* __restrict__ o) {
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r < n) o[r] = c[r];
}

template <class T>
struct Buf {
    T* p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        CAGR_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        return CAGR_OK;
    }
};

int pgr_mesh(cagr_pgr* g, const float* values, float iso, float sw, hipStream_t st) {
    int rc;
    View t{};
    t.md = g->depth;
    t.in_keys = g->in_keys;
    t.n_in = g->n_in;
    t.ck = g->ck;
    t.ck_idx = g->ck_idx;
    t.n_ck = g->n_corners;
    t.values = values;
    t.iso = iso;
    t.smooth_width = sw;
    t.grid = (float)(1.0f / std::pow(2.0, t.md + 1));
    const int64_t nl = g->n_leaves, n12 = 12 * nl;
    CAGR_REQUIRE(n12 < (int64_t(1) << 31), CAGR_ERR_ARG, "too many leaves for the mesh");
    const int H = g->depth - 1;
    Buf<uint64_t> mk;
    Buf<uint8_t> leaf_mc, in_mc;
    if ((rc = mk.alloc(nl)) || (rc = leaf_mc.alloc(nl)) || (rc = in_mc.alloc(g->n_in))) return rc;
    k_leaf_morton<<<nblk(nl), kT, 0, st>>>(t, g->leaves, nl, H, mk.p, leaf_mc.p);
    CAGR_LAUNCHED();
    if (g->n_in) {
        k_inner_mc<<<nblk(g->n_in), kT, 0, st>>>(t, g->n_in, in_mc.p);
        CAGR_LAUNCHED();
    }
    t.in_mc = in_mc.p;
    t.leaf_mk = mk.p;
    t.leaf_mc = leaf_mc.p;
    t.n_leaves = nl;
    // hash tables of the internal nodes and of the leaves' Morton keys
    auto pow2 = [](int64_t n) {
        uint64_t s = 1024;
        while (s < 2 * (uint64_t)n) s <<= 1;
        return s;
    };
    const uint64_t in_size = pow2(g->n_in), lf_size = pow2(nl);
    Buf<unsigned long long> in_slot, lf_slot;
    Buf<int32_t> in_val, lf_val;
    if ((rc = in_slot.alloc(in_size)) || (rc = in_val.alloc(in_size)) ||
        (rc = lf_slot.alloc(lf_size)) || (rc = lf_val.alloc(lf_size)))
        return rc;
    CAGR_HIP(hipMemsetAsync(in_slot.p, 0, in_size * sizeof(unsigned long long), st));
    CAGR_HIP(hipMemsetAsync(lf_slot.p, 0, lf_size * sizeof(unsigned long long), st));
    if (g->n_in) {
        k_hash_build<<<nblk(g->n_in), kT, 0, st>>>(g->in_keys, g->n_in, in_size - 1, in_slot.p,
                                                    in_val.p);
        CAGR_LAUNCHED();
    }
    k_hash_build<<<nblk(nl), kT, 0, st>>>(mk.p, nl, lf_size - 1, lf_slot.p, lf_val.p);
    CAGR_LAUNCHED();
    t.in_slot = (const uint64_t*)in_slot.p;
    t.in_val = in_val.p;
    t.in_mask = in_size - 1;
    t.lf_slot = (const uint64_t*)lf_slot.p;
    t.lf_val = lf_val.p;
    t.lf_mask = lf_size - 1;

    // vertices: root keys per (leaf, slot), the valid ones compacted and
    // sorted stably by key
    Buf<uint64_t> all, keys, ktmp;
    Buf<int32_t> valid, occ, otmp, first, head;
    Buf<uint32_t> hist;
    Buf<int32_t> stmp;
    if ((rc = all.alloc(n12)) || (rc = valid.alloc(n12 + 1)) || (rc = first.alloc(n12 + 1)) ||
        (rc = stmp.alloc(scan_tmp_len(n12 + 1))))
        return rc;
    k_mc_roots<<<nblk(nl + 1), kT, 0, st>>>(t, g->leaves, nl, all.p, valid.p);
    CAGR_LAUNCHED();
    if ((rc = scan_exclusive<int32_t>(valid.p, n12 + 1, stmp.p, st))) return rc;
    int32_t m32 = 0;
    CAGR_HIP(hipMemcpyAsync(&m32, valid.p + n12, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    const int64_t m = m32;
    if (g->mesh_v) CAGR_HIP(hipFree(g->mesh_v));
    if (g->mesh_t) CAGR_HIP(hipFree(g->mesh_t));
    g->mesh_v = nullptr;
    g->mesh_t = nullptr;
    g->n_vertices = g->n_triangles = 0;
    // no leaf edge crosses the iso-value: the empty mesh (the reference's
    // GetMCIsoTriangles adds no point and no polygon then)
    if (m == 0) return CAGR_OK;
    if ((rc = keys.alloc(m)) || (rc = ktmp.alloc(m)) || (rc = occ.alloc(m)) ||
        (rc = otmp.alloc(m)) || (rc = head.alloc(m + 1)) || (rc = hist.alloc(sort_hist_len(m))))
        return rc;
    k_mc_compact<<<nblk(n12), kT, 0, st>>>(all.p, valid.p, n12, keys.p, occ.p);
    CAGR_LAUNCHED();
    if ((rc = sort_pairs_u64(keys.p, occ.p, ktmp.p, otmp.p, hist.p, m, 64, st))) return rc;
    k_fill<<<nblk(n12 + 1), kT, 0, st>>>(first.p, n12 + 1);
    CAGR_LAUNCHED();
    k_mc_first<<<nblk(m + 1), kT, 0, st>>>(keys.p, occ.p, m, first.p, head.p);
    CAGR_LAUNCHED();
    if ((rc = scan_exclusive<int32_t>(first.p, n12 + 1, stmp.p, st))) return rc;
    if ((rc = scan_exclusive<int32_t>(head.p, m + 1, stmp.p, st))) return rc;
    int32_t nv = 0;
    CAGR_HIP(hipMemcpyAsync(&nv, first.p + n12, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    CAGR_HIP(hipMalloc(&g->mesh_v, 3 * std::max<int64_t>(nv, 1) * sizeof(float)));
    k_mc_vertices<<<nblk(n12), kT, 0, st>>>(t, g->leaves, first.p, n12, g->mesh_v);
    CAGR_LAUNCHED();
    Buf<int64_t> vkeys;
    Buf<int32_t> vidx;
    if ((rc = vkeys.alloc(nv)) || (rc = vidx.alloc(nv))) return rc;
    k_mc_vkeys<<<nblk(m), kT, 0, st>>>(keys.p, occ.p, m, head.p, first.p, vkeys.p, vidx.p);
    CAGR_LAUNCHED();

    // triangles: the leaves that can emit (compacted, in preorder), counted,
    // given offsets by a scan, then written
    Buf<int32_t> active;
    if ((rc = active.alloc(nl + 1))) return rc;
    k_fill<<<nblk(nl + 1), kT, 0, st>>>(active.p, nl + 1);
    CAGR_LAUNCHED();
    k_mc_active<<<nblk(nl), kT, 0, st>>>(t, g->leaves, nl, H, mk.p, valid.p, active.p);
    CAGR_LAUNCHED();
    Buf<int32_t> atmp;
    if ((rc = atmp.alloc(scan_tmp_len(nl + 1)))) return rc;
    if ((rc = scan_exclusive<int32_t>(active.p, nl + 1, atmp.p, st))) return rc;
    int32_t na32 = 0;
    CAGR_HIP(hipMemcpyAsync(&na32, active.p + nl, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    const int64_t na = na32;
    if (na == 0) {   // roots without an emitting leaf cannot occur; no triangles then
        CAGR_HIP(hipStreamSynchronize(st));
        g->n_vertices = nv;
        return CAGR_OK;
    }
    Buf<int32_t> ids;
    if ((rc = ids.alloc(na))) return rc;
    k_compact_ids<<<nblk(nl), kT, 0, st>>>(active.p, nl, ids.p);
    CAGR_LAUNCHED();

    // one workspace per thread; enough threads to keep every CU busy
    const int64_t nthreads = std::max<int64_t>(1, std::min<int64_t>(na, 256 * 64 * 12));
    const int64_t nb = (nthreads + 63) / 64;
    Buf<unsigned char> ws;
    Buf<int32_t> cnt, err;
    Buf<int64_t> base, offs;
    Buf<unsigned long long> used;
    if ((rc = ws.alloc((size_t)nb * 64 * WorkLayout::bytes)) || (rc = cnt.alloc(na + 1)) || (rc = err.alloc(1)) ||
        (rc = base.alloc(na)) || (rc = offs.alloc(na + 1)) || (rc = used.alloc(1)))
        return rc;
    const DevVIndex vi{vkeys.p, vidx.p, nv, g->mesh_v};
    // a closed surface has about two triangles per vertex (the 1M-point
    // benchmark mesh: 2.003); a larger mesh runs the pass again with the
    // exact size
    int64_t cap = 2 * (int64_t)nv + (int64_t)nv / 10 + 65536;
    Buf<int32_t> tmp;
    for (int attempt = 0;; ++attempt) {
        if ((rc = tmp.alloc(3 * cap))) return rc;
        k_fill<<<1, kT, 0, st>>>(err.p, 1);
        CAGR_LAUNCHED();
        k_fill<<<1, kT, 0, st>>>((int32_t*)used.p, 2);
        CAGR_LAUNCHED();
        k_mc_tris<<<nb, 64, 0, st>>>(t, g->leaves, ids.p, na, vi, ws.p, cnt.p, base.p, used.p,
                                     cap, tmp.p, err.p);
        CAGR_LAUNCHED();
        unsigned long long total = 0;
        int32_t e = 0;
        CAGR_HIP(hipMemcpyAsync(&total, used.p, sizeof(total), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipMemcpyAsync(&e, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CAGR_HIP(hipStreamSynchronize(st));
        CAGR_REQUIRE(e == 0, CAGR_ERR_STATE,
                     "a leaf exceeded the marching-cubes workspace (128 iso edges or roots)");
        if ((int64_t)total <= cap) break;
        CAGR_REQUIRE(attempt == 0, CAGR_ERR_STATE, "mesh size changed between two passes");
        CAGR_HIP(hipFree(tmp.p));
        tmp.p = nullptr;
        cap = (int64_t)total;
    }
    k_fill<<<1, kT, 0, st>>>(cnt.p + na, 1);
    CAGR_LAUNCHED();
    k_widen<<<nblk(na + 1), kT, 0, st>>>(cnt.p, na + 1, offs.p);
    CAGR_LAUNCHED();
    Buf<int64_t> otmp64;
    if ((rc = otmp64.alloc(scan_tmp_len(na + 1)))) return rc;
    if ((rc = scan_exclusive<int64_t>(offs.p, na + 1, otmp64.p, st))) return rc;
    int64_t nt = 0;
    CAGR_HIP(hipMemcpyAsync(&nt, offs.p + na, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    CAGR_HIP(hipStreamSynchronize(st));
    CAGR_HIP(hipMalloc(&g->mesh_t, 3 * std::max<int64_t>(nt, 1) * sizeof(int32_t)));
    k_mc_place<<<nblk(na), kT, 0, st>>>(cnt.p, base.p, offs.p, na, tmp.p, g->mesh_t);
    CAGR_LAUNCHED();
    CAGR_HIP(hipStreamSynchronize(st));
    g->n_vertices = nv;
    g->n_triangles = nt;
    return CAGR_OK;
}

}  // namespace
}  // namespace cagr

using namespace cagr;

extern "C" {

int cagr_pgr_mesh(cagr_pgr* g, const float* values, float iso, float smooth_width, void* stream,
                  int64_t* n_vertices, int64_t* n_triangles) {
    CAGR_REQUIRE(g != nullptr, CAGR_ERR_ARG, "null PGR octree handle");
    CAGR_REQUIRE(values != nullptr, CAGR_ERR_ARG, "null corner values");
    const int rc = pgr_mesh(g, values, iso, smooth_width, (hipStream_t)stream);
    if (rc) return rc;
    if (n_vertices) *n_vertices = g->n_vertices;
    if (n_triangles) *n_triangles = g->n_triangles;
    return CAGR_OK;
}

int cagr_pgr_mesh_read(const cagr_pgr* g, float* vertices, int32_t* triangles, void* stream) {
    CAGR_REQUIRE(g != nullptr, CAGR_ERR_ARG, "null PGR octree handle");
    const hipStream_t st = (hipStream_t)stream;
    if (vertices && g->n_vertices)
        CAGR_HIP(hipMemcpyAsync(vertices, g->mesh_v, 3 * g->n_vertices * sizeof(float),
                                hipMemcpyDeviceToDevice, st));
    if (triangles && g->n_triangles)
        CAGR_HIP(hipMemcpyAsync(triangles, g->mesh_t, 3 * g->n_triangles * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, st));
    return CAGR_OK;
}

}  // extern "C"
