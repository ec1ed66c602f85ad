// Host-side triangle precompute and bounding-volume hierarchy.
//
// The mesh is static, so both are built once at scene creation and uploaded.
// The precompute follows load_ascii_stl_double (TriangleMesh.cu:363-752):
// vertices merged per file on a 1e-9 grid, vertex ids sorted per facet, edge
// ids shared across files and offset by nShift = 3*ntri, edge vectors,
// barycentric dot products and inverses, unit normal of e01 x e02 scaled by
// the reciprocal of its length, and the plane offset taken from the last
// vertex read from the file.
//
// The hierarchy is a linear BVH over triangles ordered by the Morton key of
// their centroids (the reference's mortonKeyGen, TriangleMesh.cu:4-21); the
// ordered range is split at its midpoint, as buildBVH does (BVH.cu:741-787).
// Candidate sets do not depend on the topology: a triangle is a candidate iff
// its own box passes the sphere test and the exact test passes, and a parent's
// box always contains its children's boxes exactly.
#include "gdem_mesh.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>
#include <vector>

namespace gdem {

namespace {

struct GridKey {
  long long q[3];
  bool operator==(const GridKey& o) const { return q[0] == o.q[0] && q[1] == o.q[1] && q[2] == o.q[2]; }
};
struct GridHash {
  size_t operator()(const GridKey& k) const {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (int a = 0; a < 3; ++a) h = (h ^ (uint64_t)k.q[a]) * 0x100000001b3ULL;
    return (size_t)h;
  }
};

inline uint32_t spread_bits10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

}  // namespace

uint32_t morton_code(uint32_t ix, uint32_t iy, uint32_t iz) {
  return (spread_bits10(ix) << 2) | (spread_bits10(iy) << 1) | spread_bits10(iz);
}

bool build_mesh(const double* verts, int ntri, const int* tris_per_file, int nfiles, MeshHost* out) {
  out->tri.assign(ntri, TriGeo{});
  out->aabb_lo.assign((size_t)ntri * 3, 0.0);
  out->aabb_hi.assign((size_t)ntri * 3, 0.0);
  out->centroid.assign((size_t)ntri * 3, 0.0);
  for (int a = 0; a < 3; ++a) { out->gmin[a] = 1e300; out->gmax[a] = -1e300; }
  if (ntri == 0) { out->nodes.clear(); return true; }
  const int nshift = 3 * ntri;
  std::vector<double> vx((size_t)nshift), vy((size_t)nshift), vz((size_t)nshift);
  std::map<std::pair<int, int>, int> edge_ids;
  int next_vertex = 0, next_edge = 0, t = 0;
  const int nf = (nfiles <= 0 || !tris_per_file) ? 1 : nfiles;
  for (int f = 0; f < nf; ++f) {
    const int count = (nfiles <= 0 || !tris_per_file) ? ntri : tris_per_file[f];
    std::unordered_map<GridKey, int, GridHash> merged;
    for (int c = 0; c < count; ++c, ++t) {
      if (t >= ntri) return false;
      int id[3];
      const double* fv = verts + (size_t)t * 9;
      for (int k = 0; k < 3; ++k) {
        const double px = fv[3 * k], py = fv[3 * k + 1], pz = fv[3 * k + 2];
        GridKey key{{llround(px / 1e-9), llround(py / 1e-9), llround(pz / 1e-9)}};
        auto ins = merged.emplace(key, next_vertex);
        if (ins.second) {
          vx[next_vertex] = px; vy[next_vertex] = py; vz[next_vertex] = pz;
          ++next_vertex;
        }
        id[k] = ins.first->second;
      }
      std::sort(id, id + 3);
      auto edge = [&](int a, int b) {
        auto ins = edge_ids.emplace(std::make_pair(a, b), next_edge);
        if (ins.second) ++next_edge;
        return ins.first->second + nshift;
      };
      TriGeo& g = out->tri[t];
      g.v0 = id[0]; g.v1 = id[1]; g.v2 = id[2];
      g.e0 = edge(id[0], id[1]);
      g.e1 = edge(id[1], id[2]);
      g.e2 = edge(id[0], id[2]);
      const double P[3][3] = {{vx[id[0]], vy[id[0]], vz[id[0]]},
                              {vx[id[1]], vy[id[1]], vz[id[1]]},
                              {vx[id[2]], vy[id[2]], vz[id[2]]}};
      for (int a = 0; a < 3; ++a) {
        g.p0[a] = P[0][a]; g.p1[a] = P[1][a]; g.p2[a] = P[2][a];
        g.e01[a] = P[1][a] - P[0][a];
        g.e02[a] = P[2][a] - P[0][a];
        g.e12[a] = P[2][a] - P[1][a];
      }
      auto dotp = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
      g.d00 = dotp(g.e01, g.e01);
      g.d00inv = 1. / g.d00;
      g.d01 = dotp(g.e01, g.e02);
      g.d11 = dotp(g.e02, g.e02);
      g.d11inv = 1. / g.d11;
      const double d22 = dotp(g.e12, g.e12);
      g.d22inv = 1. / d22;
      g.denom = 1. / (g.d00 * g.d11 - g.d01 * g.d01);
      double n[3] = {g.e01[1] * g.e02[2] - g.e01[2] * g.e02[1],
                     g.e01[2] * g.e02[0] - g.e01[0] * g.e02[2],
                     g.e01[0] * g.e02[1] - g.e01[1] * g.e02[0]};
      const double s = 1. / sqrt(dotp(n, n));
      for (int a = 0; a < 3; ++a) g.n[a] = s * n[a];
      // plane offset from the raw coordinates of the facet's last vertex line
      g.d = -g.n[0] * fv[6] - g.n[1] * fv[7] - g.n[2] * fv[8];
      for (int a = 0; a < 3; ++a) {
        double lo = P[0][a], hi = P[0][a];
        if (P[1][a] < lo) lo = P[1][a];
        if (P[2][a] < lo) lo = P[2][a];
        if (P[1][a] > hi) hi = P[1][a];
        if (P[2][a] > hi) hi = P[2][a];
        out->aabb_lo[(size_t)t * 3 + a] = lo;
        out->aabb_hi[(size_t)t * 3 + a] = hi;
        if (lo < out->gmin[a]) out->gmin[a] = lo;
        if (hi > out->gmax[a]) out->gmax[a] = hi;
        out->centroid[(size_t)t * 3 + a] = (P[0][a] + P[1][a] + P[2][a]) / 3.0;
      }
    }
  }
  if (t != ntri) return false;
  out->n_edges = next_edge;

  // Morton order of centroids, stable by original index.
  std::vector<uint32_t> key(ntri);
  for (int i = 0; i < ntri; ++i) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const double sc = 1023 / (out->gmax[a] - out->gmin[a]);
      const double v = (out->centroid[(size_t)i * 3 + a] - out->gmin[a]) * sc;
      q[a] = std::isfinite(v) ? (uint32_t)(long long)v : 0u;
    }
    key[i] = morton_code(q[0], q[1], q[2]);
  }
  out->order.resize(ntri);
  for (int i = 0; i < ntri; ++i) out->order[i] = i;
  std::stable_sort(out->order.begin(), out->order.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  out->morton.resize(ntri);
  for (int i = 0; i < ntri; ++i) out->morton[i] = key[out->order[i]];

  // Midpoint split of the Morton-ordered range.  Internal nodes only; a
  // range of one triangle becomes a leaf reference in its parent.
  out->nodes.clear();
  out->nodes.reserve(std::max(1, ntri - 1));
  struct Range { int lo, hi, node, side; };
  auto box_of = [&](int lo, int hi, double* bl, double* bh) {
    for (int a = 0; a < 3; ++a) { bl[a] = 1e300; bh[a] = -1e300; }
    for (int i = lo; i < hi; ++i) {
      const int tt = out->order[i];
      for (int a = 0; a < 3; ++a) {
        bl[a] = std::min(bl[a], out->aabb_lo[(size_t)tt * 3 + a]);
        bh[a] = std::max(bh[a], out->aabb_hi[(size_t)tt * 3 + a]);
      }
    }
  };
  if (ntri == 1) {
    // a single leaf: a root whose both children are the same leaf, which the
    // traversal visits once (the second child is marked empty).
    BvhNode nd{};
    box_of(0, 1, nd.lo[0], nd.hi[0]);
    for (int a = 0; a < 3; ++a) { nd.lo[1][a] = 1e300; nd.hi[1][a] = -1e300; }
    nd.child[0] = ~out->order[0];
    nd.child[1] = ~out->order[0];
    out->nodes.push_back(nd);
    return true;
  }
  std::vector<Range> todo;
  out->nodes.push_back(BvhNode{});
  todo.push_back({0, ntri, 0, -1});
  while (!todo.empty()) {
    Range rg = todo.back();
    todo.pop_back();
    const int mid = (rg.lo + rg.hi) / 2;
    const int lohi[2][2] = {{rg.lo, mid}, {mid, rg.hi}};
    for (int c = 0; c < 2; ++c) {
      const int a = lohi[c][0], b = lohi[c][1];
      box_of(a, b, out->nodes[rg.node].lo[c], out->nodes[rg.node].hi[c]);
      if (b - a == 1) {
        out->nodes[rg.node].child[c] = ~out->order[a];
      } else {
        const int id = (int)out->nodes.size();
        out->nodes.push_back(BvhNode{});
        out->nodes[rg.node].child[c] = id;
        todo.push_back({a, b, id, c});
      }
    }
  }
  // A leaf's box must be exactly its triangle's box (the candidate test).
  return true;
}

}  // namespace gdem
