/*
 * cagr_amd -- MI355X (gfx950) treecode Gauss-formula operator of CAGR.
 *
 * C ABI.  Plain pointers and sizes only; every array argument is a device
 * pointer unless the comment says "host".  Every function returns a status
 * (CAGR_OK = 0); on failure cagr_last_error() holds a message for the calling
 * thread.  Work is enqueued on the given HIP stream (NULL = legacy default
 * stream) and is asynchronous unless noted.
 *
 * Streams and threads: calls on one tree handle share its workspaces (per-call
 * source and node records, accumulators, queues), so they are ordered: a call
 * enqueued on a stream other than the previous call's makes its stream wait
 * (hipStreamWaitEvent) for the previous call's work first.  Calls on one
 * handle therefore never overlap on the device, whatever the streams; calls on
 * different handles are independent and may run concurrently.  A handle must
 * not be used from two host threads at once (the ordering state is not
 * locked).  Calls captured into a HIP graph are ordered by the graph.
 *
 * Reference interfaces replaced (paths relative to the reference repo root,
 * CUDA-based/ext/wn_treecode/ abbreviated to wnt/):
 *   wnt/wn_treecode_cpu/wn_treecode_cpu_torch_interface.cpp:104-118  build_tree
 *   wnt/wn_treecode_func.py:23-27                                    tree H2D copy
 *   wnt/wn_treecode_cuda/wn_treecode_cuda_torch_interface.cu:41-134  scatter_point_attrs_to_nodes
 *   wnt/wn_treecode_cuda/wn_treecode_cuda_torch_interface.cu:137-777 multiply_by_{A,AT,A3,G,A21,A22,A23,A2T1,A2T2,A2T3}
 *   CUDA-based/ext/gaussrecon_src/main_GaussRecon_cuda.cu:427-466     multiply_by_A3 grid query
 *   CUDA-based/ext/gaussrecon_src/Octree.cpp:77-400                   Octree::setTree (PGR query set)
 *   CUDA-based/ext/gaussrecon_src/Octree.cpp:852-1200                 Octree::GetMCIsoTriangles (mesh)
 */
#ifndef CAGR_AMD_H
#define CAGR_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CAGR_OK 0
#define CAGR_ERR_ARG 1       /* invalid argument (shape, dtype, op, size) */
#define CAGR_ERR_HIP 2       /* HIP runtime error */
#define CAGR_ERR_STATE 3     /* handle used in a state that does not allow it */

/* element types of caller arrays */
#define CAGR_F32 0
#define CAGR_F64 1

/* precision of the operator arithmetic */
#define CAGR_PREC_FP64 0     /* plain FP64 throughout (reference formula) */
#define CAGR_PREC_FP32C 1    /* FP32 interactions summed in FP32 over short runs (one
                                leaf's points, one run of far nodes), each run's
                                partial added into an FP64 running total -- FP64
                                totals, not Kahan/Neumaier compensation; opening
                                tests, cutoff ties and node moments in FP64 */

/* operators.  in = attribute dim per point, out = output dim per query.  */
#define CAGR_OP_A 0          /* in 3, out 1: sum (x-y).mu/r^3 (CUDA sign, cuda_kernels.cu:79-110) */
#define CAGR_OP_A_CPU 1      /* in 3, out 1: negated A (CPU sign, cpu_kernels.cpp:77-103) */
#define CAGR_OP_AT 2         /* in 1, out 3: sum (x-y) s/r^3 (cuda_kernels.cu:228-258) */
#define CAGR_OP_G 3          /* in 3, out 3: sum mu/r^3 - 3 d (d.mu)/r^5 (cuda_kernels.cu:384-414) */
#define CAGR_OP_A21 4        /* in 3, out 1: convection kernel, direction x (cuda_kernels.cu:112-150) */
#define CAGR_OP_A22 5        /* ... direction y (cuda_kernels.cu:152-188) */
#define CAGR_OP_A23 6        /* ... direction z (cuda_kernels.cu:190-224) */
#define CAGR_OP_A2T1 7       /* in 1, out 3: reference transpose, direction x (cuda_kernels.cu:262-300) */
#define CAGR_OP_A2T2 8       /* ... direction y (cuda_kernels.cu:302-339) */
#define CAGR_OP_A2T3 9       /* ... direction z (cuda_kernels.cu:341-379) */
#define CAGR_OP_A3 10        /* in 3, out 1: (A21+A22+A23)/3 (cuda_kernels.cu:694-763) */
#define CAGR_OP_A2 11        /* in 3, out 3: (A21, A22, A23) from one upward pass + one
                                traversal (the three reference calls fused) */
#define CAGR_OP_A2T_SUM 12   /* in 3, out 3: A2T1(s[:,0]) + A2T2(s[:,1]) + A2T3(s[:,2]),
                                each with its own node centres, one traversal
                                (main_CAGR.py:109-111 pattern fused) */

typedef struct cagr_tree cagr_tree;

const char* cagr_last_error(void);
int cagr_version(void);

/* Octree over points [n,3] (dtype CAGR_F32/F64, original order, device).
 * Replaces build_tree (wn_treecode_cpu_torch_interface.cpp:41-118) + the H2D
 * copy of wn_treecode_func.py:23-27; the tree is built on the GPU.
 * max_depth <= 15 (ALLOWED_MAX_DEPTH), 1 <= leaf_size <= 64 (the reference
 * uses 1), theta in (0, 2] (the reference's TREECODE_THRESHOLD 2.0 is
 * theta = 0.5): a node is far from a query iff dist^2(query, node centre) >
 * (2/theta * half_w)^2.  Synchronises the stream (node count is read back).
 * Allocates all device memory the handle needs; later calls on it allocate
 * nothing except cagr_apply_queries with a larger query count than before. */
int cagr_tree_build(const void* points, int64_t n, int dtype, int max_depth, int leaf_size,
                    double theta, void* stream, cagr_tree** out);
int cagr_tree_destroy(cagr_tree* tree);

/* Host outputs: num_nodes M, num_leaves, depth of the deepest node, and the
 * length of the reference node2point_index list. */
int cagr_tree_info(const cagr_tree* tree, int64_t* num_nodes, int64_t* num_leaves,
                   int64_t* depth, int64_t* list_len);

/* The reference's serialised tree (wn_treecode_cpu_treeutils.cpp:202-258),
 * written to HOST buffers: parent [M], children [M,8], is_leaf [M],
 * half_w [M], num_points [M], node2point_index [list_len] (may be NULL),
 * node2point_indexstart [M].  Synchronous. */
int cagr_tree_export(const cagr_tree* tree, int64_t* parent, int64_t* children,
                     uint8_t* is_leaf, double* half_w, int64_t* num_points,
                     int64_t* node2point_index, int64_t* node2point_indexstart);

/* Sorted-to-original point permutation [n] int32 (HOST).  Synchronous. */
int cagr_tree_permutation(const cagr_tree* tree, int32_t* perm_host);

/* Upward pass in FP64 (replaces scatter_point_attrs_to_nodes,
 * wn_treecode_cuda_torch_interface.cu:41-134): attrs [n,attr_dim] (attr_dim 1
 * or 3, original order), weights [n] or NULL for ||attr||_2
 * (wn_treecode_func.py:131).  Outputs (device, FP64, node ids = preorder =
 * the reference's): node_attrs [M,attr_dim], node_reppoints [M,3],
 * node_weights [M]. */
int cagr_upward(cagr_tree* tree, const void* attrs, int dtype, int attr_dim,
                const void* weights, double* node_attrs, double* node_reppoints,
                double* node_weights, void* stream);

/* Operator with the tree's own points as queries (every forward_* of
 * wn_treecode_func.py): upward pass with weights ||attr|| per set, then the
 * traversal.  attrs [n,in] and widths [n] in `dtype`, original order; out
 * [n,out] in `dtype`.  continuous selects the smoothed inner kernel of A, A_cpu,
 * A2k, A2 and A3 (pybind default: false).  an is the convection strength
 * (the reference's An = 1). */
int cagr_apply(cagr_tree* tree, int op, int precision, const void* attrs, int dtype,
               const void* widths, int continuous, double an, void* out, void* stream);

/* As cagr_apply, but with the node centres taken from `weights` [n] (dtype)
 * instead of ||attr||: the operator is then linear in attrs.  weights = NULL
 * is cagr_apply. */
int cagr_apply_weighted(cagr_tree* tree, int op, int precision, const void* attrs, int dtype,
                        const void* weights, const void* widths, int continuous, double an,
                        void* out, void* stream);

/* Exact adjoint of the linear operator of cagr_apply_weighted (node centres
 * fixed by `weights` [n]): out = A^T g, where A is `op` with the tree's points
 * as queries.  op in {A, A_CPU, A21..A23, A3, A2}; g [n,out_dim(op)] and out
 * [n,3] in `dtype`, original order.  No counterpart in the reference, whose
 * A2Tk is the adjoint of the direction-reversed operator (A2Tk(An) =
 * (-A2k(-An))^T), not of A2k; this is the backward of the differentiable
 * form and the adjoint of the CGNR solve. */
int cagr_apply_adjoint(cagr_tree* tree, int op, int precision, const void* g, int dtype,
                       const void* weights, const void* widths, int continuous, double an,
                       void* out, void* stream);

/* Adjoint plan for repeated adjoints with the same node centres (an
 * iterative solve): records, for `weights` [n] (original order, `dtype`), the
 * sequence of far-field steps of every wave of the FP32 adjoint walk, so that
 * cagr_apply_adjoint_planned stores each step's wave sum into a fixed slot
 * and sums every node's slots in a fixed order, with no atomics.  Keeps a
 * copy of the weights; rebuilding replaces the previous plan.  Single-cloud
 * trees.  No counterpart in the reference (its operators recompute
 * everything per call). */
int cagr_adjoint_plan(cagr_tree* tree, const void* weights, int dtype, void* stream);

/* cagr_apply_adjoint with the weights of the tree's adjoint plan
 * (CAGR_ERR_STATE without a plan).  FP32C runs through the plan and is
 * deterministic; FP64 is the plain adjoint. */
int cagr_apply_adjoint_planned(cagr_tree* tree, int op, int precision, const void* g, int dtype,
                               const void* widths, int continuous, double an, void* out,
                               void* stream);

/* The reference's multiply_by_X with the node moments it is handed
 * (wn_treecode_cuda_torch_interface.cu:137-777, wn_treecode_cpu_torch_interface.cpp:
 * 213-378): the traversal reads node_attrs [M,in(op)] and node_reppoints
 * [M,3] (`dtype`, preorder node ids -- e.g. the outputs of
 * scatter_point_attrs_to_nodes) instead of moments of its own upward pass;
 * attrs [n,in(op)] (original order) feed the leaf sums.  queries [nq,3]
 * (qdtype) or NULL for the tree's own points (nq is then ignored);
 * widths [nq] (dtype, per query); out [nq,out(op)] (dtype).  op is one of
 * the reference's single operators (A, A_CPU, AT, G, A21..A23, A2T1..A2T3,
 * A3).  The backend of the _cuda / _cpu compatibility modules. */
int cagr_apply_moments(cagr_tree* tree, int op, int precision, const void* attrs, int dtype,
                       const void* node_attrs, const void* node_reppoints, const void* queries,
                       int qdtype, int64_t nq, const void* widths, int continuous, double an,
                       void* out, void* stream);

/* Operator at external queries [nq,3] (qdtype) -- e.g. the GaussRecon sample
 * isovalue pass (main_GaussRecon_cuda.cu:450-467).  widths [nq] (dtype) or
 * NULL for the scalar `width`.  Queries are Morton-ordered internally for the
 * traversal; out [nq,out] (dtype) is in the caller's order. */
int cagr_apply_queries(cagr_tree* tree, int op, int precision, const void* attrs, int dtype,
                       const void* queries, int qdtype, int64_t nq, const void* widths,
                       double width, int continuous, double an, void* out, void* stream);

/* cagr_apply_queries without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The queries'
 * Morton order and its sort buffers live in a static per-GPU scratch array
 * of 50 MiB (held by every GPU the library runs on until the process ends)
 * instead of the handle's grow-only workspace; nq <= 2^21 (2097152, a 128^3
 * grid), larger counts are rejected with CAGR_ERR_ARG before any device
 * work.  Same arguments, kernels and results as cagr_apply_queries.  Every
 * call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (other handles, other streams, other host threads) race; the scratch keeps
 * the last call's contents and needs no clearing. */
int cagr_apply_queries_noalloc(cagr_tree* tree, int op, int precision, const void* attrs,
                               int dtype, const void* queries, int qdtype, int64_t nq,
                               const void* widths, double width, int continuous, double an,
                               void* out, void* stream);

/* Operator on the regular grid lo + i*step, i in [0,dims) per axis (HOST
 * lo/step/dims), generated in-kernel; out [dims0*dims1*dims2] (dtype),
 * row-major with axis 2 fastest.  op in {A, A_CPU, A21..A23, A3}; one scalar
 * width.  Replaces the GaussRecon corner query (main_GaussRecon_cuda.cu:427-445)
 * on a regular grid. */
int cagr_grid_query(cagr_tree* tree, int op, int precision, const void* attrs, int dtype,
                    const double* lo, const double* step, const int64_t* dims, double width,
                    int continuous, double an, void* out, void* stream);

/* Exact interaction counts of `op`'s traversal with the tree's own points as
 * queries (int64 [n] each, original order; any may be NULL): far-node
 * interactions, leaf-point interactions, nodes visited -- summed over the
 * opening sets of the op. */
int cagr_count(cagr_tree* tree, int op, const void* attrs, int dtype, const void* weights,
               int64_t* n_far, int64_t* n_near, int64_t* n_visit, void* stream);

/* Batched form: one forest over num_clouds independent clouds stored back to
 * back in points [cloud_offsets[num_clouds], 3] (cloud c = rows
 * [cloud_offsets[c], cloud_offsets[c+1]), HOST offsets).  Every cloud gets its
 * own octree (the same tree cagr_tree_build would give it); one upward pass
 * and one persistent traversal launch serve all of them.  The handle works
 * with cagr_apply / cagr_apply_weighted / cagr_count / cagr_upward on the
 * concatenated attrs / widths / out (no counterpart in the reference, which
 * handles one cloud per process). */
int cagr_forest_build(const void* points, const int64_t* cloud_offsets, int64_t num_clouds,
                      int dtype, int max_depth, int leaf_size, double theta, void* stream,
                      cagr_tree** out);
/* Node id ranges of the clouds: node_offsets [num_clouds + 1] (HOST). */
int cagr_forest_roots(const cagr_tree* tree, int64_t* node_offsets);
/* Launch shape of the traversal: persistent = 1 routes single-cloud trees
 * through the persistent work-queue kernel too (forests always use it) with
 * waves_per_block in [1, 8] and grid_blocks (0: 8 waves per SIMD on every
 * CU); persistent = 0 sets the regular grid's waves per block, in [1, 4]. */
int cagr_set_launch(cagr_tree* tree, int persistent, int waves_per_block, int64_t grid_blocks);
/* Balances the work from a per-query cost (HOST int64 [n], original order,
 * e.g. n_far + n_near from cagr_count): the persistent work items' per-XCD
 * chunks (contiguous ranges of the Morton order) get equal cost, and the
 * regular grid's whole waves run longest first within each XCD's share (when
 * the walks are not cut into pieces); NULL restores the natural order and
 * equal chunks (and removes walk pieces). */
int cagr_order_items(cagr_tree* tree, const int64_t* query_cost);
/* Walk pieces (no counterpart in the reference; a load-balancing setting of
 * this library): the self-query walk of a wave (64 consecutive sorted points)
 * whose cost exceeds piece_cost times the mean wave cost is cut into at most
 * max_pieces (in [1, 64]) contiguous preorder node ranges of about equal
 * cost, each run as its own wave, so that the longest walks no longer set a
 * launch's length; the pieces of a wave add up in a fixed order, so results
 * do not depend on their timing (they differ from the uncut walk's only by
 * the grouping of the sums).  Costed by the walk on the FP64 node centres of
 * the most recent upward pass (set 0: e.g. the one of cagr_count), pieces
 * dispatched in Morton order.  Used by the
 * regular-grid self-query traversals of single-cloud trees (one wave per
 * block); piece_cost <= 0 removes them.  counts (HOST int64 [2], may be
 * NULL) receives the number of pieces and of waves cut.  Synchronises the
 * stream. */
int cagr_split_walks(cagr_tree* tree, double piece_cost, int max_pieces, int64_t* counts,
                     void* stream);

/* Opt-in timing of the library's own launches on the caller's stream (no
 * counterpart in the reference, which records no timings).  enable = 1 clears
 * and starts recording a HIP event pair around every traversal launch
 * (kind 0) and every upward pass incl. source records (kind 1) of this handle;
 * enable = 0 stops.  cagr_profile_read synchronises on the recorded events and
 * returns the number of recorded launches, their summed duration in ms and,
 * if each_ms != NULL, the first `cap` durations in launch order (HOST). */
int cagr_profile(cagr_tree* tree, int enable);
int cagr_profile_read(cagr_tree* tree, int kind, int64_t* launches, double* total_ms,
                      double* each_ms, int64_t cap);

/* One iteration's vector work of a conjugate-gradient loop in FP64 (no
 * counterpart in the reference's operator: its solve writes these updates as
 * torch tensor expressions, main_CAGR.py:94-144), on DEVICE vectors of len
 * doubles, after the caller has applied q = A^T A p:
 *   pq = p.q, alpha = *rr / pq (0 if pq = 0), x += alpha p, r -= alpha q,
 *   *rr_new = r.r, p = r + (*rr_new / *rr) p.
 * rr and rr_new: distinct DEVICE scalars; work: DEVICE, cagr_cg_work_len()
 * doubles.  Every reduction runs in a fixed order (the iterates do not depend
 * on scheduling); three launches on the stream, no synchronisation. */
int cagr_cg_step(int64_t len, double* x, double* r, double* p, const double* q,
                 const double* rr, double* rr_new, double* work, int64_t work_len,
                 void* stream);
int64_t cagr_cg_work_len(void);

/* GaussRecon's PGR adaptive octree over oriented samples and its query
 * corners: replaces Octree::setTree(file, d, d_min) and the gridDataVector it
 * leaves (gaussrecon_src/Octree.cpp:77-400, used at
 * main_GaussRecon_cuda.cu:255-291).  points: [n,3] float32, the samples as
 * the reference reads them from its .xyz file (raw coordinates; the normals
 * take no part in the octree).  max_depth = the reference's -d (default 10),
 * in [3, 13]; min_depth = -m (default 1), in [0, max_depth].  Synchronises the
 * stream (sizes are known only after the build). */
typedef struct cagr_pgr cagr_pgr;
int cagr_pgr_build(const float* points, int64_t n, int max_depth, int min_depth, void* stream,
                   cagr_pgr** out);
void cagr_pgr_destroy(cagr_pgr* pgr);
/* sizes (HOST): leaves, corners (= gridDataVector.size()), depth (=
 * Octree::maxDepth after the build: tree height + 1); scale_center[4] =
 * Octree::maxScale and the box origin used for the [0,1]^3 mapping.  Any
 * pointer may be NULL. */
int cagr_pgr_info(const cagr_pgr* pgr, int64_t* n_leaves, int64_t* n_corners, int* depth,
                  float* scale_center);
/* device copies: samples [n,3] in [0,1]^3 (Octree::samplePoints after the
 * mapping, Octree.cpp:256-259); corners [n_corners,3] in [0,1]^3 in
 * gridDataVector order (GridData::coords); leaves [n_leaves,4] int32
 * (depth, x, y, z offsets) in preorder.  Any pointer may be NULL. */
int cagr_pgr_read(const cagr_pgr* pgr, float* samples, float* corners, int32_t* leaves,
                  void* stream);
/* HOST: the samples' bounding box, box[6] = min xyz, max xyz (Octree::bb; the
 * output transform of writePolygon2, Octree.cpp:1259-1267, uses it). */
int cagr_pgr_box(const cagr_pgr* pgr, float* box);
/* Marching cubes on the PGR octree: replaces tree.initLeaf() +
 * tree.GetMCIsoTriangles(isoValue, &mesh, 0, 1, 0, 0)
 * (gaussrecon_src/Octree.cpp:489-496, 852-879, 1064-1200; called at
 * main_GaussRecon_cuda.cu:488-492).  values: [n_corners] float32, the
 * GridData::value of every gridDataVector entry (GaussRecon sets -wn);
 * smooth_width = GridData::smoothWidth (GaussRecon: width / 2).  The mesh
 * stays in the handle (the previous one is replaced); synchronises the
 * stream.  *n_vertices / *n_triangles (HOST) may be NULL. */
int cagr_pgr_mesh(cagr_pgr* pgr, const float* values, float iso, float smooth_width, void* stream,
                  int64_t* n_vertices, int64_t* n_triangles);
/* device copies of the last mesh: vertices [n_vertices,3] in [0,1]^3
 * (CoredVectorMeshData::inCorePoints), triangles [n_triangles,3] int32 (its
 * polygons, in order).  Either pointer may be NULL. */
int cagr_pgr_mesh_read(const cagr_pgr* pgr, float* vertices, int32_t* triangles, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CAGR_AMD_H */
