// Host-side mesh precompute and BVH build (static meshes, built once).
#pragma once
#include <cstdint>
#include <vector>

#include "gdem_types.h"

namespace gdem {

struct MeshHost {
  std::vector<TriGeo> tri;          // by original triangle id
  std::vector<double> aabb_lo, aabb_hi, centroid;  // [ntri*3]
  std::vector<int> order;           // triangle ids in Morton order of centroids
  std::vector<uint32_t> morton;     // keys in that order
  std::vector<BvhNode> nodes;       // root is node 0
  double gmin[3], gmax[3];
  int n_edges = 0;
};

uint32_t morton_code(uint32_t ix, uint32_t iy, uint32_t iz);

// Static per-cell wall table over the particle cell grid: cell c lists, in
// ascending id, every triangle whose true distance to the cell's centre is at
// most (half cell diagonal + reach) -- a superset of the triangles that can
// pass the candidate test r + skin <= reach for any sphere centred in c.
struct WallCells {
  std::vector<int> off;  // [ncell + 1]
  std::vector<int> tri;  // [off[ncell]]
};
// Returns false when the table would exceed max_entries.
bool build_wall_cells(const MeshHost& m, const double bmin[3], double cell, const int size[3],
                      double reach, size_t max_entries, WallCells* out);

// verts: [ntri*9] facet vertices in file order; tris_per_file may be null.
bool build_mesh(const double* verts, int ntri, const int* tris_per_file, int nfiles, MeshHost* out);

}  // namespace gdem
