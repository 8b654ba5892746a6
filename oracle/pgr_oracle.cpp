// CPU restatement of GaussRecon's PGR adaptive octree and its query corners.
// TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the HIP
// builder (cagr_pgr_build); no product path loads it.
//
// What the reference computes (gaussrecon_src/Octree.cpp:193-400, the
// PoissonRecon-style sample splatting), restated as cell sums:
//   1. bounding box, scale = 1.1 * longest extent, points mapped into
//      [0,1]^3 (Octree.cpp:207-259), all in float;
//   2. splat depth S = D - 2; every point adds its quadratic B-spline weights
//      to the 27 cells around its cell at each depth 0..S
//      (NonLinearUpdateWeightContribution, :1292-1325): a float sum per cell,
//      accumulated in point order;
//   3. per point, a sample depth from those sums (NonLinearGetSampleDepthAndWeight,
//      :1403-1427), clamped to [min_depth, D]; the point splats a "normal
//      length" into the 27 cells around its cell at the top depth and, when
//      the depth is fractional, at the depth above (:1328-1401, :1461-1497),
//      again float sums in point order;
//   4. ClipTree (:1499-1528) keeps the children of a node iff some strict
//      descendant has |normal length| > 1e-6; the final tree is therefore the
//      root plus the children of every strict ancestor of such a cell;
//   5. corners of every leaf (leaf preorder, corners 0..7) keyed on the grid of
//      depth H + 1, H = tree height (:312-360, ReconOctNode.cpp:95-105), in
//      order of first occurrence (setGridNode, :432-477).
// Node creation by NeighborKey::setNeighbors only ever adds zero-valued cells,
// so the cell sums above fix the tree; the restatement needs no node objects.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <unordered_map>
#include <vector>

namespace {

inline uint64_t cell_key(int d, int64_t x, int64_t y, int64_t z) {
  return ((uint64_t)d << 48) | ((uint64_t)z << 32) | ((uint64_t)y << 16) | (uint64_t)x;
}

// Quadratic B-spline weights of the 3 cells along one axis around a cell of
// float centre c and width w for the float coordinate p (Octree.cpp:1302-1307).
inline void spline3(float c, float p, double width, double out[3]) {
  const float cp = c - p;
  double x = (cp - width) / width;
  out[0] = 1.125 + 1.5 * x + 0.5 * x * x;
  x = cp / width;
  out[1] = 0.75 - x * x;
  out[2] = 1.0 - out[1] - out[0];
}

// float centre and width of the cell (d, o) (ReconOctNode.cpp:331-339)
inline void centre_width(int d, const int64_t o[3], float c[3], float& w) {
  w = (float)1.0 / (1 << d);
  for (int k = 0; k < 3; ++k) c[k] = (float)((0.5 + (double)o[k]) * w);
}

struct State {
  int D = 0, H = 0;
  std::vector<int32_t> leaves;   // [nl, 4] depth, x, y, z in preorder
  std::vector<float> corners;    // [nc, 3]
};
State g;

}  // namespace

extern "C" int pgr_oracle_build(const float* pts, int64_t n, int max_depth, int min_depth,
                                float* samples01, int64_t* n_leaves, int64_t* n_corners,
                                int* tree_depth) {
  if (n <= 0 || max_depth < 3 || max_depth > 13 || min_depth < 0) return 1;
  const int D = max_depth, S = D - 2;
  // 1. bounding box and normalisation (Octree.cpp:213-259)
  float mn[3], mx[3];
  for (int k = 0; k < 3; ++k) mn[k] = mx[k] = pts[k];
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      const float v = pts[3 * i + k];
      if (v > mx[k]) mx[k] = v;
      if (v < mn[k]) mn[k] = v;
    }
  const float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
  const float scale = std::max(std::max(ext[0], ext[1]), ext[2]) * 1.1f;
  float ctr[3];
  for (int k = 0; k < 3; ++k) ctr[k] = (mx[k] + mn[k] - scale) / 2;
  std::vector<float> p(3 * n);
  for (int64_t i = 0; i < 3 * n; ++i) p[i] = (pts[i] - ctr[i % 3]) / scale;
  std::memcpy(samples01, p.data(), 3 * n * sizeof(float));

  // cell of point i at depth d: descend comparing with the exact cell centres
  auto cell_at = [&](int64_t i, int d, int64_t o[3]) {
    o[0] = o[1] = o[2] = 0;
    for (int l = 0; l < d; ++l)
      for (int k = 0; k < 3; ++k) {
        const float c = (float)((0.5 + (double)o[k]) / (double)(1 << l));
        o[k] = 2 * o[k] + (p[3 * i + k] > c ? 1 : 0);
      }
  };

  // 2. centre weights (Octree.cpp:256-288 with :1292-1325), point order
  const double kSampleScale = 1. / (0.125 * 0.125 + 0.75 * 0.75 + 0.125 * 0.125);
  std::unordered_map<uint64_t, float> cwc;
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d <= S; ++d) {
      int64_t o[3];
      cell_at(i, d, o);
      float c[3], w;
      centre_width(d, o, c, w);
      double dx[3][3];
      for (int k = 0; k < 3; ++k) {
        spline3(c[k], p[3 * i + k], (double)w, dx[k]);
        dx[k][0] *= kSampleScale;
      }
      const int64_t lim = (int64_t)1 << d;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double dxdy = dx[0][a] * dx[1][b] * (double)1.0f;
          for (int e = 0; e < 3; ++e) {
            const int64_t x = o[0] + a - 1, y = o[1] + b - 1, z = o[2] + e - 1;
            if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
            float& v = cwc[cell_key(d, x, y, z)];
            v = (float)((double)v + dxdy * dx[2][e]);
          }
        }
    }
  }

  // GetSampleWeight (Octree.cpp:1429-1458)
  auto sample_weight = [&](int64_t i, int d) -> float {
    int64_t o[3];
    cell_at(i, d, o);
    float c[3], w;
    centre_width(d, o, c, w);
    double dx[3][3];
    for (int k = 0; k < 3; ++k) spline3(c[k], p[3 * i + k], (double)w, dx[k]);
    const int64_t lim = (int64_t)1 << d;
    float sum = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double dxdy = dx[0][a] * dx[1][b];
        for (int e = 0; e < 3; ++e) {
          const int64_t x = o[0] + a - 1, y = o[1] + b - 1, z = o[2] + e - 1;
          if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
          sum += (float)(dxdy * dx[2][e] * (double)cwc.at(cell_key(d, x, y, z)));
        }
      }
    return (float)(1.0 / (double)sum);
  };

  // 3. sample depth and the normal-length splats, point order
  const float normal_len = (float)(2 << D);
  const float spn = 1.0f;
  std::unordered_map<uint64_t, float> nl;
  auto splat = [&](int64_t i, int d, float len) {
    int64_t o[3];
    cell_at(i, d, o);
    float c[3], w;
    centre_width(d, o, c, w);
    double dx[3][3];
    for (int k = 0; k < 3; ++k) spline3(c[k], p[3 * i + k], (double)w, dx[k]);
    const int64_t lim = (int64_t)1 << d;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double dxdy = dx[0][a] * dx[1][b];
        for (int e = 0; e < 3; ++e) {
          const int64_t x = o[0] + a - 1, y = o[1] + b - 1, z = o[2] + e - 1;
          if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
          float& v = nl[cell_key(d, x, y, z)];
          v = (float)((double)v + (double)len * (dxdy * dx[2][e]));
        }
      }
  };
  for (int64_t i = 0; i < n; ++i) {
    // NonLinearGetSampleDepthAndWeight (Octree.cpp:1403-1427)
    int d = S;
    float weight = 1.0f / sample_weight(i, d);
    float depth;
    if (weight >= spn) {
      depth = (float)(d + std::log(weight / spn) / std::log(4.0));
    } else {
      float old_a = weight, new_a = weight;
      while (new_a < spn && d > 0) {
        --d;
        old_a = new_a;
        new_a = 1.0f / sample_weight(i, d);
      }
      depth = (float)(d + std::log(new_a / spn) / std::log(new_a / old_a));
    }
    const float alpha = (float)std::pow(4.0, -(double)depth);
    // NonLinearSplatOrientedPoint (Octree.cpp:1356-1400)
    if (depth < min_depth) depth = (float)min_depth;
    if (depth > D) depth = (float)D;
    int top = (int)std::ceil(depth);
    double dx = 1.0 - ((float)top - depth);
    if (top <= min_depth) {
      top = min_depth;
      dx = 1;
    } else if (top > D) {
      top = D;
      dx = 1;
    }
    double width = 1.0 / (1 << top);
    float len = (float)(normal_len * alpha / (float)(width * width * width) * dx);
    splat(i, top, len);
    if (std::fabs(1.0 - dx) > 1e-6) {
      dx = (float)(1.0 - dx);
      width = 1.0 / (1 << (top - 1));
      len = (float)((double)(normal_len * alpha) / (width * width * width) * dx);
      splat(i, top - 1, len);
    }
  }

  // 4. clipped tree: internal nodes = strict ancestors of cells with normals
  std::unordered_map<uint64_t, char> internal;
  int H = 0;
  for (const auto& kv : nl) {
    if (!(std::fabs(kv.second) > 1e-6f)) continue;
    const int d = (int)(kv.first >> 48);
    int64_t x = kv.first & 0xFFFF, y = (kv.first >> 16) & 0xFFFF, z = (kv.first >> 32) & 0xFFFF;
    H = std::max(H, d);
    for (int l = d - 1; l >= 0; --l) {
      x >>= 1; y >>= 1; z >>= 1;
      if (!internal.emplace(cell_key(l, x, y, z), 1).second) break;
    }
  }
  // leaves in preorder: depth-first walk over the internal set
  g = State();
  g.D = D;
  g.H = H;
  std::vector<int32_t>& lv = g.leaves;
  struct Frame { int d; int64_t o[3]; };
  std::vector<Frame> stack{{0, {0, 0, 0}}};
  while (!stack.empty()) {
    const Frame f = stack.back();
    stack.pop_back();
    if (!internal.count(cell_key(f.d, f.o[0], f.o[1], f.o[2]))) {
      lv.insert(lv.end(), {f.d, (int32_t)f.o[0], (int32_t)f.o[1], (int32_t)f.o[2]});
      continue;
    }
    for (int c = 7; c >= 0; --c)
      stack.push_back({f.d + 1, {2 * f.o[0] + (c & 1), 2 * f.o[1] + ((c >> 1) & 1),
                                 2 * f.o[2] + ((c >> 2) & 1)}});
  }
  // 5. corner keys on the grid of depth H + 1 and their coordinates
  const int md = H + 1;
  const float grid = (float)(1.0f / std::pow(2.0, md + 1));
  std::unordered_map<int64_t, char> seen;
  const int64_t nl_count = (int64_t)lv.size() / 4;
  for (int64_t r = 0; r < nl_count; ++r) {
    const int d = lv[4 * r];
    for (int c = 0; c < 8; ++c) {
      int64_t off[3];
      for (int k = 0; k < 3; ++k) off[k] = ((int64_t)lv[4 * r + 1 + k] + ((c >> k) & 1)) << (md - d + 1);
      const int64_t key = off[0] | (off[1] << 15) | (off[2] << 30);
      if (!seen.emplace(key, 1).second) continue;
      g.corners.push_back((float)(key & 0xFFFF) * grid);
      g.corners.push_back((float)((key >> 15) & 0xFFFF) * grid);
      g.corners.push_back((float)(key >> 30) * grid);
    }
  }
  *n_leaves = nl_count;
  *n_corners = (int64_t)g.corners.size() / 3;
  *tree_depth = md;
  return 0;
}

extern "C" int pgr_oracle_fetch(int32_t* leaves, float* corners) {
  std::memcpy(leaves, g.leaves.data(), g.leaves.size() * sizeof(int32_t));
  std::memcpy(corners, g.corners.data(), g.corners.size() * sizeof(float));
  return 0;
}
