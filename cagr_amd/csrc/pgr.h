// Handle of a PGR octree (cagr_pgr_build) and of the mesh extracted on it
// (cagr_pgr_mesh); shared by pgr_octree.hip and pgr_mesh.hip.
#pragma once
#include <cstdint>

struct cagr_pgr {
    int device = 0;
    int64_t n = 0;
    int64_t n_leaves = 0;
    int64_t n_corners = 0;
    int depth = 0;             // Octree::maxDepth after the build (tree height + 1)
    float scale = 0.f;         // Octree::maxScale
    float center[3] = {0.f, 0.f, 0.f};
    float bb[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // sample box: min xyz, max xyz
    float* samples = nullptr;  // [n,3] in [0,1]^3
    float* corners = nullptr;  // [n_corners,3] in [0,1]^3, gridDataVector order
    int32_t* leaves = nullptr; // [n_leaves,4] depth, x, y, z in preorder
    // kept for the mesh
    uint64_t* in_keys = nullptr;  // internal nodes, d<<39 | z<<26 | y<<13 | x, ascending
    int64_t n_in = 0;
    int64_t* ck = nullptr;        // corner keys (ReconOctNode::getCornerIndex), ascending
    int32_t* ck_idx = nullptr;    // their gridDataVector positions
    // the last mesh
    float* mesh_v = nullptr;      // [n_vertices,3] in [0,1]^3
    int32_t* mesh_t = nullptr;    // [n_triangles,3]
    int64_t n_vertices = 0;
    int64_t n_triangles = 0;
};
