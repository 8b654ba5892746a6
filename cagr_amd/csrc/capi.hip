// C ABI of the cagr_amd library: argument checks, operator -> kernel family
// mapping, and the calls into the build / upward / traversal code.
#include "common.h"
#include "tree_export.h"
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>

// scratch of cagr_apply_queries_noalloc (described there)
constexpr int64_t kNoAllocMaxQueries = int64_t(1) << 21;
constexpr size_t kNoAllocScratchBytes = size_t(50) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace cagr {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int prof_mark(cagr_tree* t, int kind, hipStream_t st) {
    if (!t->profiling) return CAGR_OK;
    auto& ev = t->prof_events[kind];
    size_t& used = t->prof_used[kind];
    if (used == ev.size()) {
        hipEvent_t e;
        CAGR_HIP(hipEventCreate(&e));
        ev.push_back(e);
    }
    CAGR_HIP(hipEventRecord(ev[used++], st));
    return CAGR_OK;
}

int call_begin(cagr_tree* t, hipStream_t st) {
    CAGR_REQUIRE(t->done_ev != nullptr, CAGR_ERR_STATE, "tree handle was not built");
    if (!t->ordered || t->last_stream == st) return CAGR_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    CAGR_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return CAGR_OK;   // the graph's caller orders it
    CAGR_HIP(hipStreamWaitEvent(st, t->done_ev, 0));
    return CAGR_OK;
}

void call_end(cagr_tree* t, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!t->done_ev || hipStreamIsCapturing(st, &cs) != hipSuccess ||
        cs != hipStreamCaptureStatusNone)
        return;   // a captured call is ordered by the graph's caller
    t->ordered = hipEventRecord(t->done_ev, st) == hipSuccess;
    t->last_stream = st;
}

// Morton order of arbitrary query points (implemented in tree_build.hip)
int morton_order(const void* pts, int dtype, int64_t n, int32_t* order, void* scratch,
                 hipStream_t st);
size_t morton_order_scratch(int64_t n);

namespace {

struct OpInfo {
    int family;   // 0 vector, 1 scalar sets
    int K;        // sets
    int dir0;     // convection axis of set 0 (scalar family)
    int in_dim, out_dim;
    bool convects;
};

bool op_info(int op, OpInfo* o) {
    switch (op) {
        case CAGR_OP_A: case CAGR_OP_A_CPU: *o = {0, 1, 0, 3, 1, false}; return true;
        case CAGR_OP_G: *o = {0, 1, 0, 3, 3, false}; return true;
        case CAGR_OP_A21: case CAGR_OP_A22: case CAGR_OP_A23: case CAGR_OP_A3:
            *o = {0, 1, 0, 3, 1, true}; return true;
        case CAGR_OP_A2: *o = {0, 1, 0, 3, 3, true}; return true;
        case CAGR_OP_AT: *o = {1, 1, 0, 1, 3, false}; return true;
        case CAGR_OP_A2T1: case CAGR_OP_A2T2: case CAGR_OP_A2T3:
            *o = {1, 1, op - CAGR_OP_A2T1, 1, 3, true}; return true;
        case CAGR_OP_A2T_SUM: *o = {1, 3, 0, 3, 3, true}; return true;
    }
    return false;
}

int check_common(const cagr_tree* t, int op, int precision, int dtype, OpInfo* oi) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    CAGR_REQUIRE(op_info(op, oi), CAGR_ERR_ARG, "unknown operator code");
    CAGR_REQUIRE(precision == CAGR_PREC_FP64 || precision == CAGR_PREC_FP32C, CAGR_ERR_ARG,
                 "precision must be CAGR_PREC_FP64 or CAGR_PREC_FP32C");
    CAGR_REQUIRE(dtype == CAGR_F32 || dtype == CAGR_F64, CAGR_ERR_ARG,
                 "dtype must be CAGR_F32 or CAGR_F64");
    int dev = -1;
    CAGR_HIP(hipGetDevice(&dev));
    CAGR_REQUIRE(dev == t->device, CAGR_ERR_STATE,
                 "current HIP device differs from the device the tree was built on");
    return CAGR_OK;
}

int run_upward_for(cagr_tree* t, int op, int precision, const OpInfo& oi, const void* attrs,
                   int dtype, const void* weights, double an, hipStream_t st, int wdtype = -1) {
    const double an_rec = oi.convects ? an : 0.0;
    return upward(t, oi.family, precision, oi.K, oi.dir0, attrs, dtype, weights, an_rec, st,
                  wdtype);
}

int apply_impl(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
               const void* weights, const void* widths, int continuous, double an, void* out,
               void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    if (t->n == 0) return CAGR_OK;
    CAGR_REQUIRE(attrs && widths && out, CAGR_ERR_ARG, "null attrs / widths / out");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    if ((rc = prof_mark(t, 1, st))) return rc;
    rc = run_upward_for(t, op, precision, oi, attrs, dtype, weights, an, st);
    if (rc) return rc;
    if ((rc = prof_mark(t, 1, st))) return rc;
    Queries q;
    q.mode = 0;
    q.nq = t->n;
    q.widths = widths;
    q.wdtype = dtype;
    if ((rc = prof_mark(t, 0, st))) return rc;
    rc = traverse(t, op, precision, q, continuous, an, out, dtype, nullptr, nullptr, nullptr, st);
    if (rc) return rc;
    return prof_mark(t, 0, st);
}

}  // namespace
}  // namespace cagr

using namespace cagr;

extern "C" {

const char* cagr_last_error(void) { return g_err.c_str(); }

int cagr_version(void) { return 1; }

int cagr_tree_build(const void* points, int64_t n, int dtype, int max_depth, int leaf_size,
                    double theta, void* stream, cagr_tree** out) {
    CAGR_REQUIRE(out != nullptr, CAGR_ERR_ARG, "null output handle pointer");
    *out = nullptr;
    CAGR_REQUIRE(dtype == CAGR_F32 || dtype == CAGR_F64, CAGR_ERR_ARG,
                 "points dtype must be CAGR_F32 or CAGR_F64");
    CAGR_REQUIRE(n >= 0 && n < (int64_t(1) << 31) - 1, CAGR_ERR_ARG,
                 "number of points must be in [0, 2^31 - 1)");
    CAGR_REQUIRE(n == 0 || points != nullptr, CAGR_ERR_ARG, "null points");
    CAGR_REQUIRE(max_depth >= 0 && max_depth <= kLevels, CAGR_ERR_ARG,
                 "max_depth must be in [0, 15] (ALLOWED_MAX_DEPTH)");
    CAGR_REQUIRE(leaf_size >= 1 && leaf_size <= 64, CAGR_ERR_ARG, "leaf_size must be in [1, 64]");
    CAGR_REQUIRE(theta > 0.0 && theta <= 2.0, CAGR_ERR_ARG, "theta must be in (0, 2]");
    cagr_tree* t = new (std::nothrow) cagr_tree;
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_STATE, "out of host memory");
    const int rc = build_tree(t, points, n, dtype, max_depth, leaf_size, theta, (hipStream_t)stream);
    if (rc) {
        const std::string keep = g_err;
        cagr_tree_destroy(t);
        g_err = keep;
        return rc;
    }
    *out = t;
    return CAGR_OK;
}

int cagr_profile(cagr_tree* t, int enable) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    t->profiling = enable != 0;
    if (enable) t->prof_used[0] = t->prof_used[1] = 0;
    return CAGR_OK;
}

int cagr_profile_read(cagr_tree* t, int kind, int64_t* launches, double* total_ms,
                      double* each_ms, int64_t cap) {
    CAGR_REQUIRE(t != nullptr && (kind == 0 || kind == 1), CAGR_ERR_ARG, "bad argument");
    const auto& ev = t->prof_events[kind];
    const size_t used = t->prof_used[kind];
    double ms = 0.0;
    for (size_t i = 0; i + 1 < used; i += 2) {
        CAGR_HIP(hipEventSynchronize(ev[i + 1]));
        float e = 0.f;
        CAGR_HIP(hipEventElapsedTime(&e, ev[i], ev[i + 1]));
        ms += e;
        if (each_ms && (int64_t)(i / 2) < cap) each_ms[i / 2] = e;
    }
    if (launches) *launches = (int64_t)(used / 2);
    if (total_ms) *total_ms = ms;
    return CAGR_OK;
}

int cagr_forest_build(const void* points, const int64_t* cloud_offsets, int64_t num_clouds,
                      int dtype, int max_depth, int leaf_size, double theta, void* stream,
                      cagr_tree** out) {
    CAGR_REQUIRE(out != nullptr && cloud_offsets != nullptr, CAGR_ERR_ARG, "null argument");
    *out = nullptr;
    CAGR_REQUIRE(num_clouds >= 1 && num_clouds < (int64_t(1) << 18), CAGR_ERR_ARG,
                 "num_clouds must be in [1, 2^18)");
    CAGR_REQUIRE(cloud_offsets[0] == 0, CAGR_ERR_ARG, "cloud_offsets[0] must be 0");
    for (int64_t c = 0; c < num_clouds; ++c)
        CAGR_REQUIRE(cloud_offsets[c + 1] >= cloud_offsets[c], CAGR_ERR_ARG,
                     "cloud_offsets must be non-decreasing");
    const int64_t n = cloud_offsets[num_clouds];
    CAGR_REQUIRE(dtype == CAGR_F32 || dtype == CAGR_F64, CAGR_ERR_ARG, "bad points dtype");
    CAGR_REQUIRE(n >= 0 && n < (int64_t(1) << 31) - 1, CAGR_ERR_ARG, "too many points");
    CAGR_REQUIRE(n == 0 || points != nullptr, CAGR_ERR_ARG, "null points");
    CAGR_REQUIRE(max_depth >= 0 && max_depth <= kLevels, CAGR_ERR_ARG, "max_depth must be in [0, 15]");
    CAGR_REQUIRE(leaf_size >= 1 && leaf_size <= 64, CAGR_ERR_ARG, "leaf_size must be in [1, 64]");
    CAGR_REQUIRE(theta > 0.0 && theta <= 2.0, CAGR_ERR_ARG, "theta must be in (0, 2]");
    cagr_tree* t = new (std::nothrow) cagr_tree;
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_STATE, "out of host memory");
    const int rc = build_tree(t, points, n, dtype, max_depth, leaf_size, theta, (hipStream_t)stream,
                              cloud_offsets, num_clouds);
    if (rc) {
        const std::string keep = g_err;
        cagr_tree_destroy(t);
        g_err = keep;
        return rc;
    }
    *out = t;
    return CAGR_OK;
}

int cagr_forest_roots(const cagr_tree* t, int64_t* node_offsets) {
    CAGR_REQUIRE(t != nullptr && node_offsets != nullptr, CAGR_ERR_ARG, "null argument");
    for (int64_t c = 0; c <= t->num_clouds; ++c) node_offsets[c] = t->cloud_node[c];
    return CAGR_OK;
}

int cagr_set_launch(cagr_tree* t, int persistent, int waves_per_block, int64_t grid_blocks) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    CAGR_REQUIRE(waves_per_block >= 1 && waves_per_block <= 8, CAGR_ERR_ARG,
                 "waves_per_block must be in [1, 8]");
    CAGR_REQUIRE(grid_blocks >= 0, CAGR_ERR_ARG, "grid_blocks must be >= 0");
    CAGR_REQUIRE(persistent || waves_per_block <= 4, CAGR_ERR_ARG,
                 "the regular-grid traversal takes at most 4 waves per block");
    t->persistent = persistent != 0;
    if (persistent) t->waves_per_block = waves_per_block;
    else t->grid_waves = waves_per_block;
    t->grid_blocks = grid_blocks;
    return CAGR_OK;
}

int cagr_split_walks(cagr_tree* t, double piece_cost, int max_pieces, int64_t* counts,
                     void* stream) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    const hipStream_t st = (hipStream_t)stream;
    int rc = call_begin(t, st);
    if (rc) return rc;
    CallEnd end_{t, st};
    rc = split_walks(t, piece_cost, max_pieces, st);
    if (rc) split_walks_free(t);
    if (counts) {
        counts[0] = t->num_pieces;
        counts[1] = t->num_split;
    }
    return rc;
}

int cagr_order_items(cagr_tree* t, const int64_t* query_cost) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    if (!query_cost) split_walks_free(t);
    if (t->num_items == 0) return CAGR_OK;
    if (!query_cost) {   // back to natural order, equal contiguous chunks
        (void)hipFree(t->item_order);
        t->item_order = nullptr;
        (void)hipFree(t->tile_order);
        t->tile_order = nullptr;
        for (int c = 0; c <= kQueues; ++c) t->item_bound[c] = (int32_t)(t->num_items * c / kQueues);
        return CAGR_OK;
    }
    std::vector<int4> items(t->num_items);
    CAGR_HIP(hipMemcpy(items.data(), t->items, t->num_items * sizeof(int4), hipMemcpyDeviceToHost));
    std::vector<int32_t> perm(t->n);
    CAGR_HIP(hipMemcpy(perm.data(), t->perm, t->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> cost(t->num_items, 0);
    for (int64_t i = 0; i < t->num_items; ++i)
        for (int k = 0; k < items[i].w; ++k) cost[i] += query_cost[perm[items[i].z + k]];
    // kQueues contiguous chunks of the natural (Morton, cloud by cloud) order
    // with equal total cost, each pulled in that order (neighbouring items
    // running side by side share the caches; longest first within each chunk
    // measured 1-3 % slower on 64 clouds)
    int64_t total = 0;
    for (int64_t i = 0; i < t->num_items; ++i) total += cost[i];
    int32_t bound[kQueues + 1];
    bound[0] = 0;
    int64_t run = 0, i = 0;
    for (int c = 1; c < kQueues; ++c) {
        const double target = (double)total * c / kQueues;
        while (i < t->num_items && (double)(run + cost[i]) <= target) run += cost[i++];
        bound[c] = (int32_t)i;
    }
    bound[kQueues] = (int32_t)t->num_items;
    (void)hipFree(t->item_order);
    t->item_order = nullptr;
    for (int c = 0; c <= kQueues; ++c) t->item_bound[c] = bound[c];
    if (t->num_clouds == 1) {
        // the regular grid (one wave per block, block b on XCD b mod 8, the
        // XCD's blocks covering one contiguous range of dispatch slots): each
        // XCD's range of waves, longest first
        const int64_t nw = (t->n + 63) / 64;
        std::vector<int64_t> wcost(nw, 0);
        for (int64_t s = 0; s < t->n; ++s) wcost[s / 64] += query_cost[perm[s]];
        std::vector<int32_t> torder(nw);
        const int64_t q = nw / kQueues, r = nw % kQueues;
        int64_t lo = 0;
        for (int x = 0; x < kQueues; ++x) {
            const int64_t len = q + (x < r ? 1 : 0);
            for (int64_t j = lo; j < lo + len; ++j) torder[j] = (int32_t)j;
            std::stable_sort(torder.begin() + lo, torder.begin() + lo + len,
                             [&](int32_t a, int32_t b) { return wcost[a] > wcost[b]; });
            lo += len;
        }
        if (!t->tile_order) CAGR_HIP(hipMalloc(&t->tile_order, std::max<int64_t>(nw, 1) * sizeof(int32_t)));
        CAGR_HIP(hipMemcpy(t->tile_order, torder.data(), nw * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    return CAGR_OK;
}

int cagr_tree_destroy(cagr_tree* t) {
    if (!t) return CAGR_OK;
    adjoint_plan_free(t);
    split_walks_free(t);
    for (auto& ev : t->prof_events)
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    (void)hipFree(t->pts64);
    (void)hipFree(t->pts32);
    (void)hipFree(t->perm);
    (void)hipFree(t->keys);
    (void)hipFree(t->topo);
    (void)hipFree(t->order);
    (void)hipFree(t->up_seg);
    (void)hipFree(t->src);
    (void)hipFree(t->nodes);
    (void)hipFree(t->w64);
    (void)hipFree(t->r64);
    (void)hipFree(t->m64);
    (void)hipFree(t->qws);
    (void)hipFree(t->leaf_of);
    (void)hipFree(t->lbound);
    (void)hipFree(t->items);
    (void)hipFree(t->item_order);
    (void)hipFree(t->tile_order);
    (void)hipFree(t->item_counter);
    (void)hipFree(t->adj_node);
    (void)hipFree(t->adj_pt);
    if (t->done_ev) (void)hipEventDestroy(t->done_ev);
    delete t;
    return CAGR_OK;
}

int cagr_tree_info(const cagr_tree* t, int64_t* num_nodes, int64_t* num_leaves, int64_t* depth,
                   int64_t* list_len) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    if (num_nodes) *num_nodes = t->m;
    if (num_leaves) *num_leaves = t->n_leaves;
    if (depth) *depth = t->depth;
    if (list_len) {
        if (t->m == 0) {
            *list_len = 0;
        } else {
            std::vector<TopoRec> topo(t->m);
            CAGR_HIP(hipMemcpy(topo.data(), t->topo, t->m * sizeof(TopoRec), hipMemcpyDeviceToHost));
            int64_t s = 0;
            for (const auto& r : topo) s += (int64_t)r.end - r.begin;
            *list_len = s;
        }
    }
    return CAGR_OK;
}

int cagr_tree_export(const cagr_tree* t, int64_t* parent, int64_t* children, uint8_t* is_leaf,
                     double* half_w, int64_t* num_points, int64_t* node2point_index,
                     int64_t* node2point_indexstart) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    CAGR_REQUIRE(parent && children && is_leaf && half_w && num_points && node2point_indexstart,
                 CAGR_ERR_ARG, "null export buffer");
    if (t->m == 0) return CAGR_OK;
    std::vector<TopoRec> topo(t->m);
    std::vector<int32_t> perm(t->n);
    std::vector<uint64_t> keys(t->n);
    CAGR_HIP(hipDeviceSynchronize());
    CAGR_HIP(hipMemcpy(topo.data(), t->topo, t->m * sizeof(TopoRec), hipMemcpyDeviceToHost));
    CAGR_HIP(hipMemcpy(perm.data(), t->perm, t->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    CAGR_HIP(hipMemcpy(keys.data(), t->keys, t->n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    export_reference_layout(topo.data(), t->m, perm.data(), keys.data(), parent, children, is_leaf,
                            half_w, num_points, node2point_index, node2point_indexstart);
    return CAGR_OK;
}

int cagr_tree_permutation(const cagr_tree* t, int32_t* perm_host) {
    CAGR_REQUIRE(t != nullptr && perm_host != nullptr, CAGR_ERR_ARG, "null argument");
    if (t->n == 0) return CAGR_OK;
    CAGR_HIP(hipDeviceSynchronize());
    CAGR_HIP(hipMemcpy(perm_host, t->perm, t->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return CAGR_OK;
}

int cagr_upward(cagr_tree* t, const void* attrs, int dtype, int attr_dim, const void* weights,
                double* node_attrs, double* node_reppoints, double* node_weights, void* stream) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    CAGR_REQUIRE(attr_dim == 1 || attr_dim == 3, CAGR_ERR_ARG, "attr_dim must be 1 or 3");
    CAGR_REQUIRE(dtype == CAGR_F32 || dtype == CAGR_F64, CAGR_ERR_ARG, "bad dtype");
    if (t->m == 0) return CAGR_OK;
    CAGR_REQUIRE(attrs && node_attrs && node_reppoints && node_weights, CAGR_ERR_ARG,
                 "null argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = call_begin(t, st);
    if (rc) return rc;
    CallEnd end_{t, st};
    rc = upward(t, attr_dim == 3 ? 0 : 1, CAGR_PREC_FP64, 1, 0, attrs, dtype, weights, 0.0, st);
    if (rc) return rc;
    CAGR_HIP(hipMemcpyAsync(node_attrs, t->m64, t->m * attr_dim * sizeof(double),
                            hipMemcpyDeviceToDevice, st));
    CAGR_HIP(hipMemcpyAsync(node_reppoints, t->r64, t->m * 3 * sizeof(double),
                            hipMemcpyDeviceToDevice, st));
    CAGR_HIP(hipMemcpyAsync(node_weights, t->w64, t->m * sizeof(double), hipMemcpyDeviceToDevice, st));
    return CAGR_OK;
}

int cagr_apply(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
               const void* widths, int continuous, double an, void* out, void* stream) {
    return apply_impl(t, op, precision, attrs, dtype, nullptr, widths, continuous, an, out, stream);
}

int cagr_apply_weighted(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                        const void* weights, const void* widths, int continuous, double an,
                        void* out, void* stream) {
    return apply_impl(t, op, precision, attrs, dtype, weights, widths, continuous, an, out, stream);
}

int cagr_apply_adjoint(cagr_tree* t, int op, int precision, const void* g, int dtype,
                       const void* weights, const void* widths, int continuous, double an,
                       void* out, void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    CAGR_REQUIRE(op == CAGR_OP_A || op == CAGR_OP_A_CPU || op == CAGR_OP_A3 || op == CAGR_OP_A2 ||
                     (op >= CAGR_OP_A21 && op <= CAGR_OP_A23),
                 CAGR_ERR_ARG, "adjoint supports A, A_CPU, A21..A23, A2 and A3");
    CAGR_REQUIRE(t->num_clouds == 1, CAGR_ERR_ARG, "the adjoint needs a single-cloud tree");
    if (t->n == 0) return CAGR_OK;
    CAGR_REQUIRE(g && weights && widths && out, CAGR_ERR_ARG, "null g / weights / widths / out");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    // node centres from the fixed weights (the moments are not needed)
    if ((rc = prof_mark(t, 1, st))) return rc;
    rc = upward(t, 0, precision, 1, 0, nullptr, dtype, weights, oi.convects ? an : 0.0, st);
    if (rc) return rc;
    if ((rc = prof_mark(t, 1, st))) return rc;
    if ((rc = prof_mark(t, 0, st))) return rc;
    rc = adjoint(t, op, precision, g, dtype, widths, continuous, an, out, st);
    if (rc) return rc;
    return prof_mark(t, 0, st);
}

int cagr_adjoint_plan(cagr_tree* t, const void* weights, int dtype, void* stream) {
    CAGR_REQUIRE(t != nullptr, CAGR_ERR_ARG, "null tree handle");
    CAGR_REQUIRE(dtype == CAGR_F32 || dtype == CAGR_F64, CAGR_ERR_ARG, "bad dtype");
    CAGR_REQUIRE(t->num_clouds == 1, CAGR_ERR_ARG, "the adjoint plan needs a single-cloud tree");
    if (t->n == 0) return CAGR_OK;
    CAGR_REQUIRE(weights, CAGR_ERR_ARG, "null weights");
    const hipStream_t st = (hipStream_t)stream;
    int rc = call_begin(t, st);
    if (rc) return rc;
    CallEnd end_{t, st};
    rc = adjoint_plan(t, weights, dtype, st);
    if (rc) adjoint_plan_free(t);
    return rc;
}

int cagr_apply_adjoint_planned(cagr_tree* t, int op, int precision, const void* g, int dtype,
                               const void* widths, int continuous, double an, void* out,
                               void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    CAGR_REQUIRE(op == CAGR_OP_A || op == CAGR_OP_A_CPU || op == CAGR_OP_A3 || op == CAGR_OP_A2 ||
                     (op >= CAGR_OP_A21 && op <= CAGR_OP_A23),
                 CAGR_ERR_ARG, "adjoint supports A, A_CPU, A21..A23, A2 and A3");
    if (t->n == 0) return CAGR_OK;
    CAGR_REQUIRE(t->plan_slots >= 0, CAGR_ERR_STATE, "no adjoint plan: call cagr_adjoint_plan");
    CAGR_REQUIRE(g && widths && out, CAGR_ERR_ARG, "null g / widths / out");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    // the FP32 records of the plan's centres are kept after the first call
    // (per convection factor); FP64 recomputes them
    const double an_rec = oi.convects ? an : 0.0;
    const bool keep = precision == CAGR_PREC_FP32C;
    const bool kept = keep && t->plan_nodes && t->plan_rec_an == an_rec;
    if ((rc = prof_mark(t, 1, st))) return rc;
    if (!kept) {
        rc = upward(t, 0, precision, 1, 0, nullptr, CAGR_F64, t->plan_w, an_rec, st);
        if (rc) return rc;
        if (keep) {
            const size_t nb = (size_t)t->m * sizeof(NodeV32), sb = (size_t)t->n * sizeof(SrcV32);
            const size_t cb = (size_t)t->m * 3 * sizeof(double);
            if (!t->plan_nodes || !t->plan_src || !t->plan_c64) {
                (void)hipFree(t->plan_nodes);
                (void)hipFree(t->plan_src);
                (void)hipFree(t->plan_c64);
                t->plan_nodes = t->plan_src = nullptr;
                t->plan_c64 = nullptr;
                CAGR_HIP(hipMalloc(&t->plan_nodes, nb));
                CAGR_HIP(hipMalloc(&t->plan_src, sb));
                CAGR_HIP(hipMalloc(&t->plan_c64, cb));
            }
            CAGR_HIP(hipMemcpyAsync(t->plan_nodes, t->nodes, nb, hipMemcpyDeviceToDevice, st));
            CAGR_HIP(hipMemcpyAsync(t->plan_src, t->src, sb, hipMemcpyDeviceToDevice, st));
            CAGR_HIP(hipMemcpyAsync(t->plan_c64, t->r64, cb, hipMemcpyDeviceToDevice, st));
            t->plan_rec_an = an_rec;
        }
    }
    if ((rc = prof_mark(t, 1, st))) return rc;
    if ((rc = prof_mark(t, 0, st))) return rc;
    rc = adjoint(t, op, precision, g, dtype, widths, continuous, an, out, st, true,
                 keep ? t->plan_nodes : nullptr, keep ? t->plan_src : nullptr,
                 keep ? t->plan_c64 : nullptr);
    if (rc) return rc;
    return prof_mark(t, 0, st);
}

}  // extern "C"

namespace cagr {
namespace {

__global__ void k_zero_words(uint32_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// query workspace: the Morton order [nq] int32, then the sort's scratch
size_t query_ws_bytes(int64_t nq) { return nq * sizeof(int32_t) + morton_order_scratch(nq); }

// the handle's grow-only query workspace, at least query_ws_bytes(nq)
int grow_query_ws(cagr_tree* t, int64_t nq, hipStream_t st, char** w) {
    const size_t need = query_ws_bytes(nq);
    if (need > t->ws_queries) {
        CAGR_HIP(hipStreamSynchronize(st));
        (void)hipFree(t->qws);
        t->qws = nullptr;
        t->ws_queries = 0;
        CAGR_HIP(hipMalloc(&t->qws, need));
        t->ws_queries = need;
    }
    *w = (char*)t->qws;
    return CAGR_OK;
}

// the shared part of cagr_apply_queries and its no-allocation variant; `ws`
// returns the query workspace (query_ws_bytes(nq) bytes) once the arguments
// have been checked, or fails
template <class Ws>
int apply_queries_impl(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                       const void* queries, int qdtype, int64_t nq, const void* widths,
                       double width, int continuous, double an, void* out, void* stream,
                       Ws ws) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    CAGR_REQUIRE(qdtype == CAGR_F32 || qdtype == CAGR_F64, CAGR_ERR_ARG, "bad query dtype");
    CAGR_REQUIRE(nq >= 0 && nq < (int64_t(1) << 31) - 1, CAGR_ERR_ARG, "bad query count");
    if (nq == 0) return CAGR_OK;
    CAGR_REQUIRE(queries && out, CAGR_ERR_ARG, "null queries / out");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    if (t->n == 0) {   // no sources: every output is zero (a fill kernel, no memset node)
        const int64_t words = nq * oi.out_dim * (dtype == CAGR_F32 ? 1 : 2);
        k_zero_words<<<(words + 255) / 256, 256, 0, st>>>((uint32_t*)out, words);
        CAGR_LAUNCHED();
        return CAGR_OK;
    }
    CAGR_REQUIRE(attrs, CAGR_ERR_ARG, "null attrs");
    char* w = nullptr;
    if ((rc = ws(st, &w))) return rc;
    int32_t* order = (int32_t*)w;
    rc = morton_order(queries, qdtype, nq, order, w + nq * sizeof(int32_t), st);
    if (rc) return rc;
    rc = run_upward_for(t, op, precision, oi, attrs, dtype, nullptr, an, st);
    if (rc) return rc;
    Queries q;
    q.mode = 1;
    q.nq = nq;
    q.q = queries;
    q.qdtype = qdtype;
    q.qorder = order;
    q.widths = widths;
    q.wdtype = dtype;
    q.width = width;
    return traverse(t, op, precision, q, continuous, an, out, dtype, nullptr, nullptr, nullptr, st);
}

// ---------------------------------------------------------------- no-allocation variant
// Scratch of cagr_apply_queries_noalloc: the queries' Morton order and its
// radix-sort buffers, for at most kNoAllocMaxQueries queries (2^21 = a 128^3
// grid; query_ws_bytes(2^21) = 49.1 MiB).  One copy per GPU, held from the
// library's load to the end of the process.  In host code `scratch` names a
// host-side stand-in: its device address comes from hipGetSymbolAddress, per
// device, and only that address is passed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's `scratch` (looked up once per device,
// with that device current)
int scratch_address(char** out) {
    int dev = -1;
    CAGR_HIP(hipGetDevice(&dev));
    CAGR_REQUIRE(dev >= 0 && dev < kMaxDevices, CAGR_ERR_STATE, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CAGR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CAGR_REQUIRE(p != nullptr, CAGR_ERR_HIP, "no device address for the query scratch");
        g_scratch_ptr[dev] = (char*)p;
    }
    *out = g_scratch_ptr[dev];
    return CAGR_OK;
}

}  // namespace
}  // namespace cagr

extern "C" {

int cagr_apply_queries(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                       const void* queries, int qdtype, int64_t nq, const void* widths,
                       double width, int continuous, double an, void* out, void* stream) {
    auto grow = [t, nq](hipStream_t st, char** w) -> int { return grow_query_ws(t, nq, st, w); };
    return apply_queries_impl(t, op, precision, attrs, dtype, queries, qdtype, nq, widths, width,
                              continuous, an, out, stream, grow);
}

int cagr_apply_moments(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                       const void* node_attrs, const void* node_reppoints, const void* queries,
                       int qdtype, int64_t nq, const void* widths, int continuous, double an,
                       void* out, void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    CAGR_REQUIRE(op != CAGR_OP_A2 && op != CAGR_OP_A2T_SUM, CAGR_ERR_ARG,
                 "cagr_apply_moments takes the reference's single operators (not A2 / A2T_SUM)");
    CAGR_REQUIRE(qdtype == CAGR_F32 || qdtype == CAGR_F64, CAGR_ERR_ARG, "bad query dtype");
    if (!queries) nq = t->n;
    CAGR_REQUIRE(nq >= 0 && nq < (int64_t(1) << 31) - 1, CAGR_ERR_ARG, "bad query count");
    if (nq == 0) return CAGR_OK;
    CAGR_REQUIRE(widths && out, CAGR_ERR_ARG, "null widths / out");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    if (t->n == 0) {
        const int64_t words = nq * oi.out_dim * (dtype == CAGR_F32 ? 1 : 2);
        k_zero_words<<<(words + 255) / 256, 256, 0, st>>>((uint32_t*)out, words);
        CAGR_LAUNCHED();
        return CAGR_OK;
    }
    CAGR_REQUIRE(attrs && node_attrs && node_reppoints, CAGR_ERR_ARG,
                 "null attrs / node_attrs / node_reppoints");
    // source records from the point attributes (the upward pass also writes
    // node records of its own moments; they are replaced by the given ones)
    rc = run_upward_for(t, op, precision, oi, attrs, dtype, nullptr, an, st);
    if (rc) return rc;
    rc = nodes_from_moments(t, oi.family, precision, oi.dir0, node_attrs, node_reppoints, dtype,
                            oi.convects ? an : 0.0, st);
    if (rc) return rc;
    Queries q;
    q.nq = nq;
    q.widths = widths;
    q.wdtype = dtype;
    if (!queries) {
        q.mode = 0;
    } else {
        char* w = nullptr;
        if ((rc = grow_query_ws(t, nq, st, &w))) return rc;
        int32_t* order = (int32_t*)w;
        rc = morton_order(queries, qdtype, nq, order, w + nq * sizeof(int32_t), st);
        if (rc) return rc;
        q.mode = 1;
        q.q = queries;
        q.qdtype = qdtype;
        q.qorder = order;
    }
    return traverse(t, op, precision, q, continuous, an, out, dtype, nullptr, nullptr, nullptr, st);
}

int cagr_apply_queries_noalloc(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                               const void* queries, int qdtype, int64_t nq, const void* widths,
                               double width, int continuous, double an, void* out, void* stream) {
    CAGR_REQUIRE(nq <= kNoAllocMaxQueries, CAGR_ERR_ARG,
                 "cagr_apply_queries_noalloc takes at most 2^21 (2097152) queries");
    CAGR_REQUIRE(query_ws_bytes(kNoAllocMaxQueries) <= kNoAllocScratchBytes, CAGR_ERR_STATE,
                 "query scratch smaller than its stated maximum");
    auto fixed = [](hipStream_t, char** w) -> int { return scratch_address(w); };
    return apply_queries_impl(t, op, precision, attrs, dtype, queries, qdtype, nq, widths, width,
                              continuous, an, out, stream, fixed);
}

int cagr_grid_query(cagr_tree* t, int op, int precision, const void* attrs, int dtype,
                    const double* lo, const double* step, const int64_t* dims, double width,
                    int continuous, double an, void* out, void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, precision, dtype, &oi);
    if (rc) return rc;
    CAGR_REQUIRE(op == CAGR_OP_A || op == CAGR_OP_A_CPU || op == CAGR_OP_A3 ||
                     (op >= CAGR_OP_A21 && op <= CAGR_OP_A23),
                 CAGR_ERR_ARG, "grid query supports A, A_CPU, A21..A23 and A3");
    CAGR_REQUIRE(lo && step && dims && out, CAGR_ERR_ARG, "null argument");
    for (int d = 0; d < 3; ++d)
        CAGR_REQUIRE(dims[d] >= 0 && dims[d] < (int64_t(1) << 20), CAGR_ERR_ARG, "bad grid dims");
    const int64_t total = dims[0] * dims[1] * dims[2];
    if (total == 0) return CAGR_OK;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    if (t->n == 0) {
        CAGR_HIP(hipMemsetAsync(out, 0, total * (dtype == CAGR_F32 ? 4 : 8), st));
        return CAGR_OK;
    }
    CAGR_REQUIRE(attrs, CAGR_ERR_ARG, "null attrs");
    rc = run_upward_for(t, op, precision, oi, attrs, dtype, nullptr, an, st);
    if (rc) return rc;
    Queries q;
    q.mode = 2;
    for (int d = 0; d < 3; ++d) {
        q.glo[d] = lo[d];
        q.gstep[d] = step[d];
        q.gdim[d] = dims[d];
    }
    q.width = width;
    return traverse(t, op, precision, q, continuous, an, out, dtype, nullptr, nullptr, nullptr, st);
}

int cagr_count(cagr_tree* t, int op, const void* attrs, int dtype, const void* weights,
               int64_t* n_far, int64_t* n_near, int64_t* n_visit, void* stream) {
    OpInfo oi;
    int rc = check_common(t, op, CAGR_PREC_FP64, dtype, &oi);
    if (rc) return rc;
    if (t->n == 0) return CAGR_OK;
    CAGR_REQUIRE(attrs, CAGR_ERR_ARG, "null attrs");
    CAGR_REQUIRE(n_far || n_near || n_visit, CAGR_ERR_ARG, "no counter output given");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = call_begin(t, st))) return rc;
    CallEnd end_{t, st};
    rc = run_upward_for(t, op, CAGR_PREC_FP64, oi, attrs, dtype, weights, 1.0, st);
    if (rc) return rc;
    Queries q;
    q.mode = 0;
    q.nq = t->n;
    q.width = 0.0;
    return traverse(t, op, CAGR_PREC_FP64, q, 0, 1.0, nullptr, dtype, n_far, n_near, n_visit, st);
}

}  // extern "C"
