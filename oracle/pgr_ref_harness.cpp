// Test-only harness around the reference's own PGR octree (GaussRecon).
//
// Built by oracle/Makefile into oracle/_ref/libpgr_ref.so from the reference
// sources where they lie (gaussrecon_src/*.cpp and the vendored ANN sources);
// nothing of the reference is copied here.  Only the tests load it, as the
// checker of the PGR query-set builder.
//
//   pgr_ref_build   Octree::setTree(file, max_depth, min_depth)
//                   (main_GaussRecon_cuda.cu:255-256), then reports the sizes
//   pgr_ref_fetch   the normalised samples in [0,1]^3 (Octree::samplePoints
//                   after readFile, Octree.cpp:256-259) and the grid corners
//                   in gridDataVector order (Octree.cpp:323-360, 432-477)
//   pgr_ref_leaves  depth and offsets of every leaf in preorder
//                   (ReconOctNode::nextLeaf)
//   pgr_ref_mesh    the rest of main_GaussRecon_cuda.cu:477-498 on given grid
//                   values: value / smoothWidth per gridDataVector entry,
//                   initLeaf, GetMCIsoTriangles(iso, mesh, 0, 1, 0, 0) and
//                   writePolygon2 (PLY) when a path is given
//   pgr_ref_mesh_fetch  the mesh's in-core points ([0,1]^3) and its polygons
//   pgr_ref_mc_tables   MarchingCubes::edgeMask / triangles / cornerMap
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "Octree.h"

#include "MarchingCubes.h"

namespace {
Octree* g_tree = nullptr;
CoredVectorMeshData* g_mesh = nullptr;
}

extern "C" int pgr_ref_build(const char* path, int max_depth, int min_depth,
                             int64_t* n_samples, int64_t* n_grid, int* tree_max_depth,
                             int64_t* n_leaves) {
  delete g_tree;
  g_tree = new Octree();
  if (!g_tree->setTree(std::string(path), max_depth, min_depth)) return 1;
  *n_samples = (int64_t)Octree::samplePoints.size();
  *n_grid = (int64_t)g_tree->gridDataVector.size();
  *tree_max_depth = g_tree->maxDepth;
  *n_leaves = g_tree->root.leaves();
  return 0;
}

extern "C" int pgr_ref_fetch(float* samples, float* grid, float* scale_center) {
  if (!g_tree) return 1;
  const auto& sp = Octree::samplePoints;
  for (size_t i = 0; i < sp.size(); ++i) {
    samples[3 * i + 0] = sp[i].x;
    samples[3 * i + 1] = sp[i].y;
    samples[3 * i + 2] = sp[i].z;
  }
  const auto& gv = g_tree->gridDataVector;
  for (size_t i = 0; i < gv.size(); ++i)
    for (int k = 0; k < 3; ++k) grid[3 * i + k] = gv[i]->coords[k];
  scale_center[0] = g_tree->maxScale;
  return 0;
}

extern "C" int pgr_ref_leaves(int32_t* depth_off /* [n_leaves, 4] */) {
  if (!g_tree) return 1;
  int64_t i = 0;
  for (ReconOctNode* n = g_tree->root.nextLeaf(); n; n = g_tree->root.nextLeaf(n), ++i) {
    int d, off[3];
    n->depthAndOffset(d, off);
    depth_off[4 * i + 0] = d;
    depth_off[4 * i + 1] = off[0];
    depth_off[4 * i + 2] = off[1];
    depth_off[4 * i + 3] = off[2];
  }
  return 0;
}

extern "C" int pgr_ref_mesh(const float* values, float iso, float smooth_width,
                            const char* ply_path, int64_t* n_vertices, int64_t* n_polygons,
                            int64_t* n_indices) {
  if (!g_tree) return 1;
  auto& gv = g_tree->gridDataVector;
  for (size_t i = 0; i < gv.size(); ++i) {
    gv[i]->value = values[i];
    gv[i]->smoothWidth = smooth_width;
  }
  delete g_mesh;
  g_mesh = new CoredVectorMeshData();
  g_tree->initLeaf();
  g_tree->GetMCIsoTriangles(iso, g_mesh, 0, 1, 0, 0);
  if (ply_path && ply_path[0]) {
    std::vector<char> name(ply_path, ply_path + std::strlen(ply_path) + 1);
    g_tree->writePolygon2(g_mesh, name.data());
  }
  *n_vertices = (int64_t)g_mesh->inCorePoints.size() + g_mesh->outOfCorePointCount();
  *n_polygons = g_mesh->polygonCount();
  int64_t idx = 0;
  std::vector<CoredVertexIndex> poly;
  g_mesh->resetIterator();
  while (g_mesh->nextPolygon(poly)) idx += (int64_t)poly.size();
  *n_indices = idx;
  return 0;
}

// points [nv,3]; poly_off [np+1]; poly_idx [n_indices] (out-of-core points
// follow the in-core ones, as PlyWritePolygons numbers them)
extern "C" int pgr_ref_mesh_fetch(float* points, int64_t* poly_off, int32_t* poly_idx) {
  if (!g_mesh) return 1;
  const auto& pc = g_mesh->inCorePoints;
  for (size_t i = 0; i < pc.size(); ++i) {
    points[3 * i] = pc[i].x;
    points[3 * i + 1] = pc[i].y;
    points[3 * i + 2] = pc[i].z;
  }
  g_mesh->resetIterator();
  Point q;
  size_t k = pc.size();
  while (g_mesh->nextOutOfCorePoint(q)) {
    points[3 * k] = q.x;
    points[3 * k + 1] = q.y;
    points[3 * k + 2] = q.z;
    ++k;
  }
  std::vector<CoredVertexIndex> poly;
  int64_t np = 0, idx = 0;
  poly_off[0] = 0;
  while (g_mesh->nextPolygon(poly)) {
    for (const auto& v : poly) poly_idx[idx++] = v.inCore ? v.idx : v.idx + (int32_t)pc.size();
    poly_off[++np] = idx;
  }
  return 0;
}

extern "C" void pgr_ref_mc_tables(int32_t* edge_mask, int32_t* triangles, int32_t* corner_map) {
  for (int i = 0; i < 256; ++i) {
    edge_mask[i] = MarchingCubes::edgeMask[i];
    for (int j = 0; j < 16; ++j) triangles[16 * i + j] = MarchingCubes::triangles[i][j];
  }
  for (int c = 0; c < 8; ++c) corner_map[c] = MarchingCubes::cornerMap[c];
}
