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
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "Octree.h"

namespace {
Octree* g_tree = nullptr;
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
