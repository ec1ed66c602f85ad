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
#include <utility>
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

namespace {

// Squared distance from p to the closest point of triangle (a, b, c), by
// Voronoi-region classification (Ericson, Real-Time Collision Detection,
// 5.1.5).  Used only as a conservative bound.
double point_tri_dist2(const double* p, const double* a, const double* b, const double* c) {
  double ab[3], ac[3], ap[3];
  for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
  auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  auto d2 = [&](const double* q) {
    const double e[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return dot(e, e);
  };
  const double d1 = dot(ab, ap), d2v = dot(ac, ap);
  if (d1 <= 0 && d2v <= 0) return d2(a);
  double bp[3];
  for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return d2(b);
  const double vc = d1 * d4 - d3 * d2v;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    const double q[3] = {a[0] + v * ab[0], a[1] + v * ab[1], a[2] + v * ab[2]};
    return d2(q);
  }
  double cp[3];
  for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return d2(c);
  const double vb = d5 * d2v - d1 * d6;
  if (vb <= 0 && d2v >= 0 && d6 <= 0) {
    const double w = d2v / (d2v - d6);
    const double q[3] = {a[0] + w * ac[0], a[1] + w * ac[1], a[2] + w * ac[2]};
    return d2(q);
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const double q[3] = {b[0] + w * (c[0] - b[0]), b[1] + w * (c[1] - b[1]), b[2] + w * (c[2] - b[2])};
    return d2(q);
  }
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  const double q[3] = {a[0] + ab[0] * v + ac[0] * w, a[1] + ab[1] * v + ac[1] * w, a[2] + ab[2] * v + ac[2] * w};
  return d2(q);
}

}  // namespace

bool build_wall_cells(const MeshHost& m, const double bmin[3], double cell, const int size[3],
                      double reach, size_t max_entries, WallCells* out) {
  const int ntri = (int)m.tri.size();
  const size_t ncell = (size_t)size[0] * size[1] * size[2];
  const double half_diag = 0.5 * cell * std::sqrt(3.0);
  // conservative radius with a relative margin for rounding
  const double rad = (half_diag + reach) * (1.0 + 1e-9) + 1e-12;
  const double rad2 = rad * rad;
  const double inv = 1.0 / cell;
  std::vector<std::pair<int, int>> pairs;  // (cell, triangle), triangles visited in ascending id
  for (int t = 0; t < ntri; ++t) {
    const TriGeo& g = m.tri[t];
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      const double l = m.aabb_lo[(size_t)t * 3 + a] - rad, h = m.aabb_hi[(size_t)t * 3 + a] + rad;
      lo[a] = std::max(0, (int)std::floor((l - bmin[a]) * inv) + 1);
      hi[a] = std::min(size[a] - 1, (int)std::floor((h - bmin[a]) * inv) + 1);
    }
    for (int cz = lo[2]; cz <= hi[2]; ++cz)
      for (int cy = lo[1]; cy <= hi[1]; ++cy)
        for (int cx = lo[0]; cx <= hi[0]; ++cx) {
          // cell cx spans [bmin + (cx-1) cell, bmin + cx cell)
          const double c[3] = {bmin[0] + (cx - 0.5) * cell, bmin[1] + (cy - 0.5) * cell,
                               bmin[2] + (cz - 0.5) * cell};
          const double pd = g.n[0] * c[0] + g.n[1] * c[1] + g.n[2] * c[2] + g.d;
          if (std::fabs(pd) > rad) continue;
          if (point_tri_dist2(c, g.p0, g.p1, g.p2) > rad2) continue;
          pairs.emplace_back((size_t)(size[1] * cz + cy) * size[0] + cx, t);
          if (pairs.size() > max_entries) return false;
        }
  }
  out->off.assign(ncell + 1, 0);
  for (auto& pr : pairs) out->off[(size_t)pr.first + 1]++;
  for (size_t c = 0; c < ncell; ++c) out->off[c + 1] += out->off[c];
  out->tri.assign(pairs.size(), 0);
  std::vector<int> fill(out->off.begin(), out->off.end() - 1);
  for (auto& pr : pairs) out->tri[fill[pr.first]++] = pr.second;  // stays ascending per cell
  return true;
}

}  // namespace gdem
