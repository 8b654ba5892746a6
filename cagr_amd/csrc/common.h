// Shared definitions of the cagr_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "tree_logic.h"
#include "../../include/cagr_amd.h"

namespace cagr {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
#define CAGR_HIP(expr)                                                              \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::cagr::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            return CAGR_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)
#define CAGR_LAUNCHED()                                                             \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess) {                                                     \
            ::cagr::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return CAGR_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)
#define CAGR_REQUIRE(cond, code, msg)                                               \
    do {                                                                            \
        if (!(cond)) {                                                              \
            ::cagr::set_error(msg);                                                 \
            return code;                                                            \
        }                                                                           \
    } while (0)

// ---------------------------------------------------------------- records
// Source (point) records, one per sorted point, rebuilt per call; node
// records, one per node (per set for the scalar family), rebuilt by the
// upward pass.  All are read with wave-uniform (scalar) loads in the
// traversal, so they are packed point-major.

// vector family (attr_dim 3: A, A_cpu, G, A2k, A2 fused, A3)
struct alignas(16) SrcV64 { double y[3]; double mu[3]; };
struct alignas(16) SrcV32 { float y[3]; float mu[3]; float e[3]; float q[3]; };
//   e[k] = exp(-An*y_k/2), q[k] = An/2 * e[k] * mu[k]
struct alignas(16) NodeV64 { double c[3]; double m[3]; };
struct alignas(16) NodeV32 { float cf[3]; float hi2, lo2, ce; float m[3]; float e[3]; float q[3]; float pad; };
//   cf = (float)c of the weighted centre c, hi2 / lo2 / ce its opening and
//   cutoff bounds (as in NodeS32; c itself stays in the upward pass's r64),
//   m = attr sum, e/q as above at c

// scalar family (attr_dim 1 per set: AT, A2Tk, fused sum_k A2Tk)
template <int K> struct alignas(16) SrcS64 { double y[3]; double s[K]; };
template <int K> struct alignas(16) SrcS32 { float y[3]; float a[K]; };
//   a[k] = s_k * exp(-An*y_{dir_k}/2)   (= s_k for AT)
struct alignas(16) NodeS64 { double c[3]; double s; };
struct alignas(16) NodeS32 { float cf[3]; float se; float hi2, lo2, ce; float pad; };
//   se = s * exp(-An*c_{dir}/2) at this set's centre (the product in FP64);
//   The FP32 walk decides far/near on cf = (float)c: an FP32 distance^2
//   above hi2 is far and below lo2 near for the FP64 test too; between them
//   (or for the cutoff, within ce of the width) the lane retests in FP64 with
//   the upward pass's r64 entry (open_band)

// persistent traversal queues: one per XCD, counters a cache line apart
constexpr int kQueues = 8;
constexpr int kCounterStride = 32;

}  // namespace cagr

// ---------------------------------------------------------------- tree handle
struct cagr_tree {
    int device = 0;
    int64_t n = 0;            // points
    int64_t m = 0;            // nodes
    int64_t n_leaves = 0;
    int depth = 0;            // deepest node
    int max_depth = 15;
    int leaf_size = 1;
    double theta = 0.5;
    double open_factor = 4.0; // 2 / theta
    int point_dtype = CAGR_F32;

    // device, owned
    double* pts64 = nullptr;      // [n,3] sorted by (key, leaf-local index)
    float* pts32 = nullptr;       // [n,3] same order
    int32_t* perm = nullptr;      // [n] sorted position -> original index
    uint64_t* keys = nullptr;     // [n] sorted Morton keys
    cagr::TopoRec* topo = nullptr;// [m] preorder
    int32_t* order = nullptr;     // [m] node ids by depth, preorder within a depth
    float4* lbound = nullptr;     // [m] bounding sphere {centre, radius} of each leaf's points
    // largest |x - (double)(float)x| over the points' coordinates (0 for FP32
    // points): how far the FP32 walks' float query copies are off
    double qround = 0.0;
    std::vector<int32_t> depth_off;  // host: order[] offsets of depths 0..16 (size 17)
    // the upward pass: one launch over the leaves (64 consecutive node ids a
    // wave), then one launch per band of at most 3 depths, deepest band
    // first, a wave per node at the band's top depth (the band's internal
    // nodes); up_bands[b] = {top depth, depths, roots, first up_seg entry};
    // up_seg holds each band root's run of order[] per level
    std::vector<int4> up_bands;
    int2* up_seg = nullptr;

    // workspace, owned (sized at build)
    void* src = nullptr;          // n * 64 bytes
    void* nodes = nullptr;        // m * 128 bytes
    double* w64 = nullptr;        // [3, m]
    double* r64 = nullptr;        // [3, m, 3]
    double* m64 = nullptr;        // [3, m, 3]
    int32_t* leaf_of = nullptr;   // [n] leaf id of each sorted point
    double* adj_node = nullptr;   // [m, 3] adjoint far-field accumulators
    double* adj_pt = nullptr;     // [3][n] adjoint near-field accumulators (planar)
    size_t ws_queries = 0;        // grow-only query workspace (explicit-query sorting)
    void* qws = nullptr;

    // clouds of a forest (a plain tree is one cloud): sorted point ranges and
    // node id ranges [cloud_node[c], cloud_node[c+1]) (host)
    int64_t num_clouds = 1;
    std::vector<int64_t> cloud_pt, cloud_node;
    // wave work items of the persistent traversal (device): {root, stop, q0, qn}.
    // The pull order is split into kQueues chunks, one per XCD (blocks are
    // dispatched to the XCDs round robin), each with its own counter; a wave
    // drains its XCD's chunk, then steals from the others.
    int4* items = nullptr;
    int32_t* item_order = nullptr;   // pull order of the items (null: natural order)
    int32_t item_bound[cagr::kQueues + 1] = {};   // chunk c = pull positions [bound[c], bound[c+1])
    int32_t* item_counter = nullptr; // [kQueues * kCounterStride] chunk counters
    int64_t num_items = 0;
    // adjoint plan: the fixed far-step sequence of the FP32 adjoint for the
    // node centres of plan_w; far step k of wave w is slot wave_off[w] + k,
    // stored at slot_pos[slot] in node-major order, so L of a node is the sum
    // of the contiguous run [node_off[node], node_off[node + 1]) of far_val
    double* plan_w = nullptr;          // [n] the plan's point weights (original order)
    int32_t* plan_wave_off = nullptr;  // [waves + 1]
    int32_t* plan_node_off = nullptr;  // [m + 1]
    int32_t* plan_slot_pos = nullptr;  // [slots] node-major position of each slot
    float4* plan_far_val = nullptr;    // [slots] node-major
    // the recorded walk: far steps (node, lanes) wave-major in slot order, and
    // near leaves (leaf, lanes) per wave from near_off
    int32_t* plan_far_node = nullptr;  // [slots]
    uint64_t* plan_far_mask = nullptr; // [slots]
    int32_t* plan_near_off = nullptr;  // [waves + 1]
    int32_t* plan_near_leaf = nullptr; // [near entries]
    uint64_t* plan_near_mask = nullptr;// [near entries]
    // the replay's pieces: the far steps and near entries of a long wave cut
    // into ranges of about equal cost, each range replayed as its own wave
    // {wave, far first, far end, near first, near end, -, -, -} (null: one
    // piece per wave, all of its lists), in dispatch order
    int32_t* plan_pieces = nullptr;
    int64_t plan_num_pieces = 0;
    int64_t plan_slots = -1;           // -1: no plan
    // FP32 node and source records of the plan's centres (an upward pass of
    // plan_w with no attributes), kept after the first planned call with
    // convection factor plan_rec_an (NaN: none kept), and those FP64 centres
    // ([m, 3], the cutoff tests' retests)
    void* plan_nodes = nullptr;
    void* plan_src = nullptr;
    double* plan_c64 = nullptr;
    double plan_rec_an = __builtin_nan("");
    int persistent = 0;              // 1: traversal through the item queue
    int waves_per_block = 4;         // launch shape of the persistent traversal kernels
    int grid_waves = 1;              // waves per block of the regular-grid traversal kernels
    // dispatch order of the regular grid's waves (one wave per block; null:
    // Morton order): each XCD's contiguous range of waves, longest first
    int32_t* tile_order = nullptr;
    // walk pieces of the self-query walks (cagr_split_walks; null: none): a
    // wave whose walk costs more than the target is cut into node ranges,
    // each run as its own wave {wave, first node, end node, partial slot}
    // (slot -1: the wave's only piece, straight into the output); the split
    // waves' pieces write FP64 partial slots [slot][64][3] that a combine
    // launch adds up per split wave {wave, first slot, pieces}
    int4* pieces = nullptr;
    int64_t num_pieces = 0;
    int4* split = nullptr;
    int64_t num_split = 0;
    double* planes = nullptr;
    int num_cus = 256;               // compute units of the device
    int64_t grid_blocks = 0;         // persistent grid size (0 = from the device)

    // order of the calls on this handle: its workspaces (source and node
    // records, accumulators, queues, query scratch) serve one call at a time,
    // so a call enqueued on a stream other than the previous call's first
    // waits for done_ev, recorded on the calling stream at the end of every
    // call (not while the stream is being captured into a graph)
    hipEvent_t done_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool ordered = false;         // done_ev holds the last call's end

    // opt-in launch timing (cagr_profile): event pairs per kind, reused
    bool profiling = false;
    std::vector<hipEvent_t> prof_events[2];
    size_t prof_used[2] = {0, 0};
};

namespace cagr {
// record an event of the current region (begin/end alternate) when profiling
int prof_mark(cagr_tree* t, int kind, hipStream_t st);
// call ordering on one handle: call_begin makes `st` wait for the previous
// call's work when that came on another stream; CallEnd records the end of
// this call (destructor, so every return path records it)
int call_begin(cagr_tree* t, hipStream_t st);
void call_end(cagr_tree* t, hipStream_t st);
struct CallEnd {
    cagr_tree* t;
    hipStream_t st;
    ~CallEnd() { call_end(t, st); }
};
}

namespace cagr {
// implemented in tree_build.hip
int build_tree(cagr_tree* t, const void* points, int64_t n, int dtype, int max_depth,
               int leaf_size, double theta, hipStream_t st, const int64_t* cloud_off = nullptr,
               int64_t nclouds = 1);
int sort_pairs_u64(uint64_t* keys, int32_t* vals, uint64_t* keys_tmp, int32_t* vals_tmp,
                   uint32_t* hist, int64_t n, int bits, hipStream_t st);
int64_t sort_hist_len(int64_t n);
// exclusive prefix sum of a[0..len) in place; tmp holds scan_tmp_len(len) elements
int64_t scan_tmp_len(int64_t len);
template <typename U>
int scan_exclusive(U* a, int64_t len, U* tmp, hipStream_t st);
// implemented in upward.hip
int upward(cagr_tree* t, int family, int precision, int K, int dir0, const void* attrs,
           int dtype, const void* weights, double an, hipStream_t st, int wdtype = -1);
int nodes_from_moments(cagr_tree* t, int family, int precision, int dir0, const void* node_attrs,
                       const void* node_reppoints, int dtype, double an, hipStream_t st);
// implemented in traverse.hip
struct Queries {
    int mode = 0;                 // 0 self, 1 explicit, 2 grid
    int64_t nq = 0;
    const void* q = nullptr;      // explicit [nq,3]
    int qdtype = CAGR_F32;
    const int32_t* qorder = nullptr;
    double glo[3] = {0, 0, 0}, gstep[3] = {0, 0, 0};
    int64_t gdim[3] = {0, 0, 0};
    const void* widths = nullptr; // per query (self: original order), or null
    int wdtype = CAGR_F32;
    double width = 0.0;           // used when widths == null
};
// implemented in adjoint.hip (needs the node records of an upward pass with
// the same fixed weights)
// (planned: FP32 far steps go through the tree's adjoint plan instead of atomics;
// nodes / src / c64: records and their FP64 centres to read instead of the
// tree's own, when not null)
int adjoint(cagr_tree* t, int op, int precision, const void* g, int dtype, const void* widths,
            int continuous, double an, void* out, hipStream_t st, bool planned = false,
            const void* nodes = nullptr, const void* src = nullptr,
            const double* c64 = nullptr);
// builds the adjoint plan for point weights (original order; centres of an
// upward pass with them)
int adjoint_plan(cagr_tree* t, const void* weights, int dtype, hipStream_t st);
void adjoint_plan_free(cagr_tree* t);
// cuts the self-query walks into pieces of about piece_cost times the mean
// wave cost (at most max_pieces a wave), costed by the walk on the FP64
// centres of set 0 of the last upward pass; piece_cost <= 0 removes them
int split_walks(cagr_tree* t, double piece_cost, int max_pieces, hipStream_t st);
void split_walks_free(cagr_tree* t);
int traverse(cagr_tree* t, int op, int precision, const Queries& q, int continuous, double an,
             void* out, int odtype, int64_t* cnt_far, int64_t* cnt_near, int64_t* cnt_visit,
             hipStream_t st);
}  // namespace cagr
