// CPU oracle for the gdem contact engine -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++/OpenMP, the per-step DEM contact path of
// bi-rugaoishii/gpudem so that the HIP engine can be checked against it and so
// that the reference's CPU path can be timed beside the GPU.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
// library never links or calls it.
//
// Every routine cites the reference file:line it follows.  Arithmetic is kept
// operation-for-operation (built with -ffp-contract=off, matching the
// reference's -fmad=false, makefile:5) because the engine's FP64 path is
// required to be bit-identical to it.
//
// Reference behaviour reproduced on purpose (never fixed silently):
//  * history truncation in update_history(_wall) (cpu_dem.cu:587-634): the
//    swapped-in slot's contact flag is not copied, so the first lost contact
//    drops every later slot as well;
//  * barycentric-sign closest-feature classification (cpu_dem.cu:12-205);
//  * pair forces use k_i and etaconst_i of the owning particle
//    (cpu_dem.cu:953-1011);
//  * 3.14 in the mass, 3.1415 in the damping constant (main.cu:67,
//    ParticleSystem.cu:301);
//  * 10-bit masks in the Morton interleave (BoundingBox.h:125-135).
// Out-of-range neighbour cells (a particle sitting exactly on the upper
// bounding-box face) are skipped here; the reference reads past its table.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>
#include <omp.h>

namespace {

constexpr int kMaxNei = 100;     // hardCodedParameters.h:3
constexpr double kSmall = 1e-15; // hardCodedParameters.h:7

struct V3 {
  double x, y, z;
};
// Vec3.h:23-70, same operation order.
inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 scale3(double s, V3 b) { return V3{s * b.x, s * b.y, s * b.z}; }
inline V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross3(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// BoundingBox.h:125-135
inline uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
inline uint32_t morton3(uint32_t ix, uint32_t iy, uint32_t iz) {
  return (spread10(ix) << 2) | (spread10(iy) << 1) | spread10(iz);
}

// Stable LSD radix sort, 4 x 8-bit passes (BoundingBox.cu:33-104).
void lsd_sort(std::vector<uint32_t>& key, std::vector<int>& idx) {
  const size_t n = key.size();
  std::vector<uint32_t> k2(n);
  std::vector<int> i2(n);
  for (int pass = 0; pass < 4; ++pass) {
    int shift = 8 * pass;
    size_t count[256] = {0};
    for (size_t i = 0; i < n; ++i) count[(key[i] >> shift) & 0xff]++;
    size_t run = 0;
    for (int d = 0; d < 256; ++d) {
      size_t c = count[d];
      count[d] = run;
      run += c;
    }
    for (size_t i = 0; i < n; ++i) {
      size_t pos = count[(key[i] >> shift) & 0xff]++;
      k2[pos] = key[i];
      i2[pos] = idx[i];
    }
    key.swap(k2);
    idx.swap(i2);
  }
}

// ---------------------------------------------------------------------------
// Triangle mesh precompute (TriangleMesh.cu:363-752)
// ---------------------------------------------------------------------------
struct Mesh {
  int nTri = 0, nVert = 0, nShift = 0, nEdge = 0;
  std::vector<double> mx, my, mz;
  std::vector<double> nx, ny, nz, d;
  std::vector<double> e01x, e01y, e01z, e02x, e02y, e02z, e12x, e12y, e12z;
  std::vector<double> d00, d00inv, d01, d11, d11inv, d22, d22inv, denom;
  std::vector<double> minx, maxx, miny, maxy, minz, maxz, cx, cy, cz;
  std::vector<int> i0, i1, i2, te0, te1, te2, sortedIndex;
  std::vector<uint32_t> mortonKey;  // in sorted order after the sort, as the reference leaves it
  double gmin[3] = {1e300, 1e300, 1e300}, gmax[3] = {-1e300, -1e300, -1e300};
  // BVH (BVH.cu:712-787)
  std::vector<double> bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;
  std::vector<int> bleft, bright, btri;
  int nodeCount = 0;
};

struct QKey {
  long long x, y, z;
  bool operator==(const QKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct QKeyHash {
  size_t operator()(const QKey& k) const {
    size_t h = 1469598103934665603ULL;
    h = (h ^ (size_t)k.x) * 1099511628211ULL;
    h = (h ^ (size_t)k.y) * 1099511628211ULL;
    h = (h ^ (size_t)k.z) * 1099511628211ULL;
    return h;
  }
};

// verts: ntri*9 doubles in file order (v0xyz v1xyz v2xyz per facet);
// tris_per_file: triangle count of each STL file (the vertex merge is per
// file, TriangleMesh.cu:466-468; edge ids are global, :489-508).
Mesh* build_mesh(const double* verts, int ntri, const int* tris_per_file, int nfiles) {
  Mesh* m = new Mesh();
  const int T = ntri, Nv = 3 * ntri;
  m->nTri = T;
  m->nVert = Nv;
  m->nShift = Nv;  // TriangleMesh.cu:455
  auto rs = [&](std::vector<double>& a, int n) { a.assign(n, 0.0); };
  rs(m->mx, Nv); rs(m->my, Nv); rs(m->mz, Nv);
  for (auto* a : {&m->nx, &m->ny, &m->nz, &m->d, &m->e01x, &m->e01y, &m->e01z, &m->e02x,
                  &m->e02y, &m->e02z, &m->e12x, &m->e12y, &m->e12z, &m->d00, &m->d00inv,
                  &m->d01, &m->d11, &m->d11inv, &m->d22, &m->d22inv, &m->denom, &m->minx,
                  &m->maxx, &m->miny, &m->maxy, &m->minz, &m->maxz, &m->cx, &m->cy, &m->cz})
    rs(*a, T);
  for (auto* a : {&m->i0, &m->i1, &m->i2, &m->te0, &m->te1, &m->te2, &m->sortedIndex})
    a->assign(T, 0);

  std::map<std::pair<int, int>, int> edges;  // (lo,hi) -> first index
  int vIndex = 0, eIndex = 0, tri = 0;
  for (int f = 0; f < nfiles; ++f) {
    std::unordered_map<QKey, int, QKeyHash> vmap;  // per file
    for (int t = 0; t < tris_per_file[f]; ++t, ++tri) {
      m->sortedIndex[tri] = tri;
      int vid[3];
      double x = 0, y = 0, z = 0;
      for (int c = 0; c < 3; ++c) {
        x = verts[tri * 9 + c * 3 + 0];
        y = verts[tri * 9 + c * 3 + 1];
        z = verts[tri * 9 + c * 3 + 2];
        m->mx[vIndex] = x; m->my[vIndex] = y; m->mz[vIndex] = z;
        QKey q{llround(x / 1e-9), llround(y / 1e-9), llround(z / 1e-9)};
        auto it = vmap.find(q);
        if (it == vmap.end()) {
          vmap.emplace(q, vIndex);
          vid[c] = vIndex++;
        } else {
          vid[c] = it->second;
        }
      }
      // sort3 (TriangleMesh.cu:88-103)
      int v0 = vid[0], v1 = vid[1], v2 = vid[2];
      if (v0 > v1) std::swap(v0, v1);
      if (v1 > v2) std::swap(v1, v2);
      if (v0 > v1) std::swap(v0, v1);
      auto edge_id = [&](int a, int b) {
        auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = edges.find(key);
        if (it != edges.end() && a <= b) return it->second;
        edges.emplace(key, eIndex);
        return eIndex++;
      };
      m->te0[tri] = edge_id(v0, v1) + m->nShift;
      m->te1[tri] = edge_id(v1, v2) + m->nShift;
      m->te2[tri] = edge_id(v0, v2) + m->nShift;
      V3 p0{m->mx[v0], m->my[v0], m->mz[v0]};
      V3 p1{m->mx[v1], m->my[v1], m->mz[v1]};
      V3 p2{m->mx[v2], m->my[v2], m->mz[v2]};
      V3 e01 = sub3(p1, p0), e02 = sub3(p2, p0), e12 = sub3(p2, p1);
      m->e01x[tri] = e01.x; m->e01y[tri] = e01.y; m->e01z[tri] = e01.z;
      m->e02x[tri] = e02.x; m->e02y[tri] = e02.y; m->e02z[tri] = e02.z;
      m->e12x[tri] = e12.x; m->e12y[tri] = e12.y; m->e12z[tri] = e12.z;
      m->d00[tri] = dot3(e01, e01);
      m->d00inv[tri] = 1. / m->d00[tri];
      m->d01[tri] = dot3(e01, e02);
      m->d11[tri] = dot3(e02, e02);
      m->d11inv[tri] = 1. / m->d11[tri];
      m->d22[tri] = dot3(e12, e12);
      m->d22inv[tri] = 1. / m->d22[tri];
      double den = m->d00[tri] * m->d11[tri] - m->d01[tri] * m->d01[tri];
      m->denom[tri] = 1. / den;
      V3 n = cross3(e01, e02);
      n = scale3(1. / sqrt(dot3(n, n)), n);
      m->nx[tri] = n.x; m->ny[tri] = n.y; m->nz[tri] = n.z;
      // plane offset from the last vertex read from the file (TriangleMesh.cu:573)
      m->d[tri] = -n.x * x - n.y * y - n.z * z;
      m->i0[tri] = v0; m->i1[tri] = v1; m->i2[tri] = v2;
      m->cx[tri] = (p0.x + p1.x + p2.x) / 3.0;
      m->cy[tri] = (p0.y + p1.y + p2.y) / 3.0;
      m->cz[tri] = (p0.z + p1.z + p2.z) / 3.0;
      double lo[3] = {p0.x, p0.y, p0.z}, hi[3] = {p0.x, p0.y, p0.z};
      double q1[3] = {p1.x, p1.y, p1.z}, q2[3] = {p2.x, p2.y, p2.z};
      for (int a = 0; a < 3; ++a) {
        if (q1[a] < lo[a]) lo[a] = q1[a];
        if (q2[a] < lo[a]) lo[a] = q2[a];
        if (q1[a] > hi[a]) hi[a] = q1[a];
        if (q2[a] > hi[a]) hi[a] = q2[a];
        if (lo[a] < m->gmin[a]) m->gmin[a] = lo[a];
        if (hi[a] > m->gmax[a]) m->gmax[a] = hi[a];
      }
      m->minx[tri] = lo[0]; m->miny[tri] = lo[1]; m->minz[tri] = lo[2];
      m->maxx[tri] = hi[0]; m->maxy[tri] = hi[1]; m->maxz[tri] = hi[2];
    }
  }
  m->nEdge = eIndex;
  // Morton keys of centroids (TriangleMesh.cu:4-21) then stable sort (:722)
  std::vector<uint32_t> key(T);
  const double scale = 1023;
  double inv[3];
  for (int a = 0; a < 3; ++a) inv[a] = scale / (m->gmax[a] - m->gmin[a]);
  for (int t = 0; t < T; ++t) {
    double c[3] = {m->cx[t], m->cy[t], m->cz[t]};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      double s = (c[a] - m->gmin[a]) * inv[a];
      // a flat axis gives 0*inf = NaN in the reference; x86 truncation yields 0
      q[a] = std::isfinite(s) ? (uint32_t)(long long)s : 0u;
    }
    key[t] = morton3(q[0], q[1], q[2]);
  }
  lsd_sort(key, m->sortedIndex);
  m->mortonKey = key;

  // BVH: median split over the Morton-sorted index range (BVH.cu:741-787),
  // node AABB as in computeNodeAABB (BVH.cu:712-739).
  int maxNodes = 2 * T - 1;
  for (auto* a : {&m->bminx, &m->bminy, &m->bminz, &m->bmaxx, &m->bmaxy, &m->bmaxz})
    a->assign(maxNodes, 0.0);
  m->bleft.assign(maxNodes, -1); m->bright.assign(maxNodes, -1); m->btri.assign(maxNodes, -1);
  struct Task { int s, e, node; };
  std::vector<Task> st;
  int nodeCount = 0;
  st.push_back({0, T, nodeCount++});
  while (!st.empty()) {
    Task t = st.back();
    st.pop_back();
    int s0 = m->sortedIndex[t.s];
    double lo[3] = {m->minx[s0], m->miny[s0], m->minz[s0]};
    double hi[3] = {m->maxx[s0], m->maxy[s0], m->maxz[s0]};
    for (int i = t.s + 1; i < t.e; ++i) {
      int q = m->sortedIndex[i];
      double ql[3] = {m->minx[q], m->miny[q], m->minz[q]};
      double qh[3] = {m->maxx[q], m->maxy[q], m->maxz[q]};
      for (int a = 0; a < 3; ++a) {
        if (ql[a] < lo[a]) lo[a] = ql[a];
        if (qh[a] > hi[a]) hi[a] = qh[a];
      }
    }
    m->bminx[t.node] = lo[0]; m->bminy[t.node] = lo[1]; m->bminz[t.node] = lo[2];
    m->bmaxx[t.node] = hi[0]; m->bmaxy[t.node] = hi[1]; m->bmaxz[t.node] = hi[2];
    if (t.e - t.s == 1) {
      m->bleft[t.node] = -1; m->bright[t.node] = -1;
      m->btri[t.node] = m->sortedIndex[t.s];
      continue;
    }
    int mid = (t.s + t.e) / 2;
    int l = nodeCount++, r = nodeCount++;
    m->bleft[t.node] = l; m->bright[t.node] = r; m->btri[t.node] = -1;
    st.push_back({mid, t.e, r});
    st.push_back({t.s, mid, l});
  }
  m->nodeCount = nodeCount;
  return m;
}

// ---------------------------------------------------------------------------
// Particle system (ParticleSystemMember_common.def, AoS xyz, MAX_NEI rows)
// ---------------------------------------------------------------------------
struct Sim {
  int N = 0;
  double dt = 0, mu = 0, g[3] = {0, 0, 0};
  // grid (BoundingBox.cu:364-478)
  double bmin[3], bmax[3], dx[3], invd[3], skinR, thresh, threshSq;
  int size[3], ncell;
  std::vector<double> x, v, a, f, angv, anga, mom;
  std::vector<double> r, rsq, invr, k, m, sqrtm, invm, moi, invmoi, etac;
  std::vector<double> refx, refy, refz;
  std::vector<int> isActive, pId, cellId, cellx;
  std::vector<int> numCont, numContWall, numNei, numNeiWall;
  std::vector<int> indHis, indHisWall, isContact, isContactWall, neiList, neiListWall, vorENow;
  std::vector<double> dhx, dhy, dhz, dwx, dwy, dwz;
  std::vector<int> pNum, pStart, pList, cellOffset;
  Mesh* mesh = nullptr;
  long long steps = 0, rebuilds = 0;
  int overflow = 0;
  // optional per-contact record (debug / parity): per particle up to 2*kMaxNei entries
  int record = 0;
  std::vector<int> recN, recKey, recKind;
  std::vector<double> recFn, recFt, recDel;
};

// ParticleSystem.cu:202, 294-335 and main.cu:67, evaluated per particle.
void set_material(Sim* s, int i, double r, double density, double E, double res) {
  double m = density * 3.14 * r * r * r * 4. / 3.;
  double k = 4. / 3. * (E * 0.5) * sqrt(r * 0.5) * sqrt(0.05 * r);
  s->r[i] = r;
  s->rsq[i] = r * r;
  s->invr[i] = 1. / r;
  s->k[i] = k;
  s->m[i] = m;
  s->sqrtm[i] = sqrt(m);
  s->invm[i] = 1. / m;
  s->etac[i] = -2. * log(res) * sqrt(k / (3.1415 * 3.1415 + log(res) * log(res)));
  s->moi[i] = 2. / 5. * m * (r * r);
  s->invmoi[i] = 1. / s->moi[i];
}

// cell coordinate (BoundingBox.cu:783-795 / :1089-1104)
inline void cell_of(const Sim* s, int i, int* c) {
  for (int a = 0; a < 3; ++a)
    c[a] = (int)floor((s->x[i * 3 + a] - s->bmin[a]) * s->invd[a]) + 1;
}

// update_pList_fast (BoundingBox.cu:768-837): counting sort of active particles.
void bin_particles(Sim* s) {
#pragma omp parallel for
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    int c[3];
    cell_of(s, i, c);
    s->cellx[i * 3 + 0] = c[0]; s->cellx[i * 3 + 1] = c[1]; s->cellx[i * 3 + 2] = c[2];
    s->cellId[i] = (s->size[1] * c[2] + c[1]) * s->size[0] + c[0];
  }
  // serial counting sort, as in the reference
  std::fill(s->pNum.begin(), s->pNum.end(), 0);
  std::vector<int> used;
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    int c = s->cellId[i];
    if (s->pNum[c] == 0) used.push_back(c);
    s->pNum[c]++;
  }
  int run = 0;
  for (int c : used) { s->pStart[c] = run; run += s->pNum[c]; }
  for (int c : used) s->cellOffset[c] = 0;
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    int c = s->cellId[i];
    s->pList[s->pStart[c] + s->cellOffset[c]] = i;
    s->cellOffset[c] += 1;
  }
}

inline void sort_row(int* row, int n) { std::sort(row, row + n); }

// update_neighborlist (BoundingBox.cu:526-587)
void build_pair_lists(Sim* s) {
  bin_particles(s);
  const double skin = s->skinR;
  int ovf = 0;
#pragma omp parallel for reduction(| : ovf)
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    int bi = i * 3;
    int cx = s->cellx[bi], cy = s->cellx[bi + 1], cz = s->cellx[bi + 2];
    int n = 0;
    for (int sx = -1; sx <= 1; ++sx)
      for (int sy = -1; sy <= 1; ++sy)
        for (int sz = -1; sz <= 1; ++sz) {
          int X = cx + sx, Y = cy + sy, Z = cz + sz;
          if (X < 0 || Y < 0 || Z < 0 || X >= s->size[0] || Y >= s->size[1] || Z >= s->size[2])
            continue;
          int c = (s->size[1] * Z + Y) * s->size[0] + X;
          int beg = s->pStart[c], end = beg + s->pNum[c];
          if (s->pNum[c] == 0) continue;
          for (int q = beg; q < end; ++q) {
            int j = s->pList[q];
            if (i == j || s->isActive[j] != 1) continue;
            int bj = j * 3;
            V3 del{s->x[bi] - s->x[bj], s->x[bi + 1] - s->x[bj + 1], s->x[bi + 2] - s->x[bj + 2]};
            double distsq = dot3(del, del);
            double R = s->r[i] + s->r[j] + skin;
            if (distsq < R * R) {
              if (n < kMaxNei) s->neiList[i * kMaxNei + n] = j;
              else ovf = 1;
              ++n;
            }
          }
        }
    if (n > kMaxNei) n = kMaxNei;
    s->numNei[i] = n;
    sort_row(&s->neiList[i * kMaxNei], n);
    s->refx[i] = s->x[bi];
    s->refy[i] = s->x[bi + 1];
    s->refz[i] = s->x[bi + 2];
  }
  if (ovf) s->overflow |= 1;
}

// update_neighborlist_brute (BoundingBox.cu:480-523) + sort: test helper.
void brute_pair_lists(const Sim* s, int* list, int* num) {
  for (int i = 0; i < s->N; ++i) {
    num[i] = 0;
    if (s->isActive[i] != 1) continue;
    int n = 0;
    for (int j = 0; j < s->N; ++j) {
      if (i == j || s->isActive[j] != 1) continue;
      V3 del{s->x[i * 3] - s->x[j * 3], s->x[i * 3 + 1] - s->x[j * 3 + 1],
             s->x[i * 3 + 2] - s->x[j * 3 + 2]};
      double R = s->r[i] + s->r[j] + s->skinR;
      if (dot3(del, del) < R * R && n < kMaxNei) list[i * kMaxNei + n++] = j;
    }
    num[i] = n;
  }
}

// dist_triangle / dist_triangle_neighbor (cpu_dem.cu:12-205, BVH.cu:44-238):
// the two differ only in the squared radius compared against.
struct TriHit {
  V3 n;
  double dist;
  int hitAt;
};
TriHit tri_distance(const Mesh* M, V3 pos, double Rsq, int j) {
  TriHit res;
  V3 nt{M->nx[j], M->ny[j], M->nz[j]};
  res.n = V3{0., 0., 0.};
  res.dist = 1.e10;
  res.hitAt = -1;
  double dist = 0.;
  dist += nt.x * pos.x;
  dist += nt.y * pos.y;
  dist += nt.z * pos.z;
  dist += M->d[j];
  double distsq = dist * dist;
  if (distsq > Rsq) return res;
  int v0 = M->i0[j], v1 = M->i1[j], v2 = M->i2[j];
  V3 ap{pos.x - M->mx[v0], pos.y - M->my[v0], pos.z - M->mz[v0]};
  double d00 = M->d00[j], d01 = M->d01[j], d11 = M->d11[j], den = M->denom[j];
  double d20 = ap.x * M->e01x[j] + ap.y * M->e01y[j] + ap.z * M->e01z[j];
  double d21 = ap.x * M->e02x[j] + ap.y * M->e02y[j] + ap.z * M->e02z[j];
  double vb = (d11 * d20 - d01 * d21) * den;
  double wb = (d00 * d21 - d01 * d20) * den;
  double ub = 1.0 - vb - wb;
  if (ub > 0. && vb > 0. && wb > 0.) {
    double sg = dist < 0 ? -1.0 : 1.0;
    res.n = scale3(sg, nt);
    res.dist = sg * dist;
    res.hitAt = -1;
    return res;
  }
  auto finish = [&](V3 q, int hit) {
    res.n = sub3(pos, q);
    double dsq = dot3(res.n, res.n);
    if (dsq < Rsq) {
      res.dist = sqrt(dsq);
      res.n = scale3(1. / res.dist, res.n);
    }
    res.hitAt = hit;
    return res;
  };
  V3 P0{M->mx[v0], M->my[v0], M->mz[v0]};
  V3 P1{M->mx[v1], M->my[v1], M->mz[v1]};
  V3 P2{M->mx[v2], M->my[v2], M->mz[v2]};
  if (vb <= 0. && wb <= 0.0) return finish(P0, v0);
  if (ub <= 0. && wb <= 0.0) return finish(P1, v1);
  if (ub <= 0. && vb <= 0.0) return finish(P2, v2);
  if (ub <= 0.0) {
    V3 E{M->e12x[j], M->e12y[j], M->e12z[j]};
    V3 bp = sub3(pos, P1);
    double t = dot3(bp, E) * M->d22inv[j];
    t = fmax(0., fmin(1.0, t));
    return finish(add3(P1, scale3(t, E)), M->te1[j]);
  }
  if (vb <= 0.0) {
    V3 E{M->e02x[j], M->e02y[j], M->e02z[j]};
    double t = d21 * M->d11inv[j];
    t = fmax(0., fmin(1.0, t));
    return finish(add3(P0, scale3(t, E)), M->te2[j]);
  }
  V3 E{M->e01x[j], M->e01y[j], M->e01z[j]};
  double t = d20 * M->d00inv[j];
  t = fmax(0., fmin(1.0, t));
  return finish(add3(P0, scale3(t, E)), M->te0[j]);
}

// sphereAABBOverlapNeighbor (BVH.cu:10-42)
inline bool aabb_hit(const Mesh* M, int node, V3 p, double Rsq) {
  double dx = 0.0, dy = 0.0, dz = 0.0;
  if (p.x < M->bminx[node]) dx = M->bminx[node] - p.x;
  else if (p.x > M->bmaxx[node]) dx = p.x - M->bmaxx[node];
  if (p.y < M->bminy[node]) dy = M->bminy[node] - p.y;
  else if (p.y > M->bmaxy[node]) dy = p.y - M->bmaxy[node];
  if (p.z < M->bminz[node]) dz = M->bminz[node] - p.z;
  else if (p.z > M->bmaxz[node]) dz = p.z - M->bmaxz[node];
  return (dx * dx + dy * dy + dz * dz) <= Rsq;
}

// update_neighborlist_wall (BVH.cu:576-625)
void build_wall_lists(Sim* s) {
  const Mesh* M = s->mesh;
  const double skin = s->skinR;
  int ovf = 0;
#pragma omp parallel for reduction(| : ovf)
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    int n = 0;
    if (M && M->nTri > 0) {
      V3 p{s->x[i * 3], s->x[i * 3 + 1], s->x[i * 3 + 2]};
      double R = s->r[i] + skin;
      double Rsq = R * R;
      int stack[128];
      int sp = 0;
      stack[sp++] = 0;
      while (sp > 0) {
        int node = stack[--sp];
        if (!aabb_hit(M, node, p, Rsq)) continue;
        if (M->btri[node] >= 0) {
          int t = M->btri[node];
          TriHit h = tri_distance(M, p, Rsq, t);
          if (h.dist < s->r[i] + skin) {
            if (n < kMaxNei) s->neiListWall[i * kMaxNei + n] = t;
            else ovf = 1;
            ++n;
          }
        } else {
          stack[sp++] = M->bleft[node];
          stack[sp++] = M->bright[node];
        }
      }
    }
    if (n > kMaxNei) n = kMaxNei;
    s->numNeiWall[i] = n;
    sort_row(&s->neiListWall[i * kMaxNei], n);
  }
  if (ovf) s->overflow |= 2;
}

// Brute force wall candidates: leaf AABB test + exact test over all triangles.
void brute_wall_lists(const Sim* s, int* list, int* num) {
  const Mesh* M = s->mesh;
  for (int i = 0; i < s->N; ++i) {
    num[i] = 0;
    if (s->isActive[i] != 1 || !M) continue;
    V3 p{s->x[i * 3], s->x[i * 3 + 1], s->x[i * 3 + 2]};
    double R = s->r[i] + s->skinR, Rsq = R * R;
    int n = 0;
    for (int t = 0; t < M->nTri; ++t) {
      double dx = 0.0, dy = 0.0, dz = 0.0;
      if (p.x < M->minx[t]) dx = M->minx[t] - p.x; else if (p.x > M->maxx[t]) dx = p.x - M->maxx[t];
      if (p.y < M->miny[t]) dy = M->miny[t] - p.y; else if (p.y > M->maxy[t]) dy = p.y - M->maxy[t];
      if (p.z < M->minz[t]) dz = M->minz[t] - p.z; else if (p.z > M->maxz[t]) dz = p.z - M->maxz[t];
      if (!((dx * dx + dy * dy + dz * dz) <= Rsq)) continue;
      TriHit h = tri_distance(M, p, Rsq, t);
      if (h.dist < s->r[i] + s->skinR && n < kMaxNei) list[i * kMaxNei + n++] = t;
    }
    num[i] = n;
  }
}

struct Contact {
  V3 fn, vn, vt, n;
  double eta;
};

inline void rec(Sim* s, int i, int key, int kind, V3 fn, V3 ft, double del) {
  if (!s->record) return;
  int c = s->recN[i];
  if (c >= 2 * kMaxNei) return;
  int o = i * 2 * kMaxNei + c;
  s->recKey[o] = key;
  s->recKind[o] = kind;
  s->recFn[o * 3] = fn.x; s->recFn[o * 3 + 1] = fn.y; s->recFn[o * 3 + 2] = fn.z;
  s->recFt[o * 3] = ft.x; s->recFt[o * 3 + 1] = ft.y; s->recFt[o * 3 + 2] = ft.z;
  s->recDel[o] = del;
  s->recN[i] = c + 1;
}

// Tangential spring with history, shared by pairs (cpu_dem.cu:783-950) and
// walls (cpu_dem.cu:636-781).  `keys/flags/dx..dz/num` select the pair or the
// wall history of particle i.
V3 tangential(Sim* s, int i, int key, const Contact& c, int* keys, int* flags, double* hx,
              double* hy, double* hz, int* num) {
  const int ci = i * kMaxNei;
  int slot = -1;
  for (int q = 0; q < num[i]; ++q)
    if (keys[ci + q] == key) { slot = q; flags[ci + q] = 1; break; }
  if (slot < 0) {
    slot = num[i];
    if (slot >= kMaxNei) { s->overflow |= 4; slot = kMaxNei - 1; }
    flags[ci + slot] = 1;
    keys[ci + slot] = key;
    num[i] = slot + 1;
    hx[ci + slot] = 0.; hy[ci + slot] = 0.; hz[ci + slot] = 0.;
  }
  const double dt = s->dt, kk = s->k[i];
  V3 dn{hx[ci + slot] + c.vt.x * dt, hy[ci + slot] + c.vt.y * dt, hz[ci + slot] + c.vt.z * dt};
  double dd = dot3(dn, c.n);
  dn = sub3(dn, scale3(dd, c.n));
  V3 ft = scale3(-kk, dn);
  double ftsq = dot3(ft, ft), fnsq = dot3(c.fn, c.fn);
  if (ftsq > (s->mu * s->mu) * fnsq) {
    if (ftsq != 0.) {
      V3 t = scale3(1. / (sqrt(ftsq) + kSmall), ft);
      double fnn = sqrt(fnsq);
      ft = scale3(s->mu * fnn, t);
      dn = scale3(-1. / kk, ft);
    } else {
      ft = V3{0., 0., 0.};
    }
  } else {
    ft.x -= c.eta * c.vt.x;
    ft.y -= c.eta * c.vt.y;
    ft.z -= c.eta * c.vt.z;
  }
  const int bi = i * 3;
  s->f[bi] += ft.x; s->f[bi + 1] += ft.y; s->f[bi + 2] += ft.z;
  V3 mm = scale3(s->r[i], cross3(c.n, ft));
  s->mom[bi] += mm.x; s->mom[bi + 1] += mm.y; s->mom[bi + 2] += mm.z;
  hx[ci + slot] = dn.x; hy[ci + slot] = dn.y; hz[ci + slot] = dn.z;
  return ft;
}

// update_history(_wall) (cpu_dem.cu:587-634), including its truncation.
void prune_history(int i, const int* flags, int* keys, double* hx, double* hy, double* hz, int* num) {
  const int ci = i * kMaxNei;
  for (int q = 0; q < num[i]; ++q) {
    if (flags[ci + q] == 0) {
      int last = num[i] - 1;
      keys[ci + q] = keys[ci + last];
      hx[ci + q] = hx[ci + last]; hy[ci + q] = hy[ci + last]; hz[ci + q] = hz[ci + last];
      hx[ci + last] = 0.; hy[ci + last] = 0.; hz[ci + last] = 0.;
      num[i] -= 1;
      q -= 1;  // flags[ci+q] is not refreshed, exactly as in the reference
    }
  }
}

// particle_collision_verlet + calc_normal_force (cpu_dem.cu:1084-1183, 953-1011)
void pair_forces(Sim* s) {
#pragma omp parallel for
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    const int bi = i * 3, ci = i * kMaxNei;
    for (int q = 0; q < s->numNei[i]; ++q) {
      int j = s->neiList[ci + q];
      if (s->isActive[j] != 1) continue;
      const int bj = j * 3;
      V3 del{s->x[bi] - s->x[bj], s->x[bi + 1] - s->x[bj + 1], s->x[bi + 2] - s->x[bj + 2]};
      double distsq = dot3(del, del);
      double R = s->r[i] + s->r[j];
      if (!(distsq < R * R)) continue;
      double dist = sqrt(distsq);
      double delMag = R - dist;
      V3 n{del.x / dist, del.y / dist, del.z / dist};
      Contact c;
      V3 dl = scale3(-delMag, n);
      V3 vr{s->v[bi] - s->v[bj], s->v[bi + 1] - s->v[bj + 1], s->v[bi + 2] - s->v[bj + 2]};
      double vrn = dot3(vr, n);
      c.vn = scale3(vrn, n);
      double meff = 1. / (s->invm[i] + s->invm[j]);
      double eta = s->etac[i] * sqrt(meff);
      c.eta = eta;
      double kk = s->k[i];
      c.fn.x = -kk * dl.x - eta * c.vn.x;
      c.fn.y = -kk * dl.y - eta * c.vn.y;
      c.fn.z = -kk * dl.z - eta * c.vn.z;
      V3 vrot{s->r[i] * s->angv[bi] + s->r[j] * s->angv[bj],
              s->r[i] * s->angv[bi + 1] + s->r[j] * s->angv[bj + 1],
              s->r[i] * s->angv[bi + 2] + s->r[j] * s->angv[bj + 2]};
      vrot = cross3(vrot, n);
      c.vt = add3(sub3(vr, c.vn), vrot);
      s->f[bi] += c.fn.x; s->f[bi + 1] += c.fn.y; s->f[bi + 2] += c.fn.z;
      c.n = n;
      V3 ft = tangential(s, i, s->pId[j], c, s->indHis.data(), s->isContact.data(), s->dhx.data(),
                         s->dhy.data(), s->dhz.data(), s->numCont.data());
      rec(s, i, s->pId[j], 0, c.fn, ft, delMag);
    }
    prune_history(i, s->isContact.data(), s->indHis.data(), s->dhx.data(), s->dhy.data(),
                  s->dhz.data(), s->numCont.data());
  }
}

// wall_collision_verlet + calc_normal_force_wall (cpu_dem.cu:238-299, 536-585)
void wall_forces(Sim* s) {
  const Mesh* M = s->mesh;
#pragma omp parallel for
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    const int bi = i * 3, ci = i * kMaxNei;
    int nv = 0;
    V3 p{s->x[bi], s->x[bi + 1], s->x[bi + 2]};
    for (int q = 0; q < s->numNeiWall[i]; ++q) {
      int t = s->neiListWall[ci + q];
      TriHit h = tri_distance(M, p, s->rsq[i], t);
      if (!(h.dist < s->r[i])) continue;
      if (h.hitAt != -1) {
        bool dup = false;
        for (int u = 0; u < nv; ++u)
          if (s->vorENow[ci + u] == h.hitAt) { dup = true; break; }
        if (dup) continue;
        if (nv < kMaxNei) s->vorENow[ci + nv] = h.hitAt;
        ++nv;
      }
      double delMag = s->r[i] - h.dist;
      V3 n = h.n;
      Contact c;
      V3 dl = scale3(-delMag, n);
      V3 vr{s->v[bi], s->v[bi + 1], s->v[bi + 2]};
      double vrn = dot3(vr, n);
      c.vn = scale3(vrn, n);
      double eta = s->etac[i] * s->sqrtm[i];
      c.eta = eta;
      double kk = s->k[i];
      c.fn.x = -kk * dl.x - eta * c.vn.x;
      c.fn.y = -kk * dl.y - eta * c.vn.y;
      c.fn.z = -kk * dl.z - eta * c.vn.z;
      V3 vrot{s->r[i] * s->angv[bi], s->r[i] * s->angv[bi + 1], s->r[i] * s->angv[bi + 2]};
      vrot = cross3(vrot, n);
      c.vt = add3(sub3(vr, c.vn), vrot);
      s->f[bi] += c.fn.x; s->f[bi + 1] += c.fn.y; s->f[bi + 2] += c.fn.z;
      c.n = n;
      V3 ft = tangential(s, i, t, c, s->indHisWall.data(), s->isContactWall.data(), s->dwx.data(),
                         s->dwy.data(), s->dwz.data(), s->numContWall.data());
      rec(s, i, t, 1, c.fn, ft, delMag);
    }
    prune_history(i, s->isContactWall.data(), s->indHisWall.data(), s->dwx.data(), s->dwy.data(),
                  s->dwz.data(), s->numContWall.data());
  }
}

// shouldRefreshNeighborList (cpu_dem.cu:1529-1552)
int needs_refresh(const Sim* s) {
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) continue;
    V3 d{s->x[i * 3] - s->refx[i], s->x[i * 3 + 1] - s->refy[i], s->x[i * 3 + 2] - s->refz[i]};
    if (dot3(d, d) > s->threshSq) return 1;
  }
  return 0;
}

// checkOoB (cpu_dem.cu:1575-1589)
void out_of_bounds(Sim* s) {
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] == 0) continue;
    double X = s->x[i * 3], Y = s->x[i * 3 + 1], Z = s->x[i * 3 + 2];
    if (X < s->bmin[0] || X > s->bmax[0] || Y < s->bmin[1] || Y > s->bmax[1] || Z < s->bmin[2] ||
        Z > s->bmax[2])
      s->isActive[i] = 0;
  }
}

void rebuild(Sim* s) {
  build_pair_lists(s);
  build_wall_lists(s);
  s->rebuilds++;
}

// One step.  gpu_order=1 follows device_dem_verlet_verlet (device_dem.cu:1195-1236):
// forces, integrate, OoB, refresh check, rebuild.  gpu_order=0 follows
// cpu_dem_verlet_verlet (cpu_dem.cu:1593-1668) and then main's checkOoB (main.cu:316).
void step(Sim* s, int gpu_order) {
  const int N = s->N;
#pragma omp parallel for
  for (int i = 0; i < N * 3; ++i) s->f[i] = 0.;
#pragma omp parallel for
  for (int i = 0; i < N * 3; ++i) s->mom[i] = 0.;
#pragma omp parallel for
  for (int i = 0; i < N * kMaxNei; ++i) {
    s->isContact[i] = 0;
    s->isContactWall[i] = 0;
  }
  if (s->record) std::fill(s->recN.begin(), s->recN.end(), 0);
  pair_forces(s);
  if (s->mesh && s->mesh->nTri > 0) wall_forces(s);
  const double dt = s->dt;
#pragma omp parallel for
  for (int i = 0; i < N; ++i) {
    if (s->isActive[i] != 1) continue;
    const int bi = i * 3;
    for (int a = 0; a < 3; ++a) s->a[bi + a] = s->f[bi + a] * s->invm[i] + s->g[a];
    for (int a = 0; a < 3; ++a) s->v[bi + a] += s->a[bi + a] * dt;
    for (int a = 0; a < 3; ++a) s->x[bi + a] += s->v[bi + a] * dt;
    for (int a = 0; a < 3; ++a) s->anga[bi + a] = s->mom[bi + a] * s->invmoi[i];
    for (int a = 0; a < 3; ++a) s->angv[bi + a] += s->anga[bi + a] * dt;
  }
  if (gpu_order) {
    out_of_bounds(s);
    if (needs_refresh(s)) rebuild(s);
  } else {
    if (needs_refresh(s)) rebuild(s);
    out_of_bounds(s);
  }
  s->steps++;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI for ctypes (tests and the CPU baseline only)
// ---------------------------------------------------------------------------
extern "C" {

void* orc_mesh_create(const double* verts, int ntri, const int* tris_per_file, int nfiles) {
  if (ntri <= 0) return new Mesh();
  return build_mesh(verts, ntri, tris_per_file, nfiles);
}
void orc_mesh_destroy(void* m) { delete static_cast<Mesh*>(m); }

// Field access for the mesh: returns element count, copies into out if non-null.
long orc_mesh_get(void* mp, const char* name, void* out) {
  Mesh* m = static_cast<Mesh*>(mp);
  auto cpd = [&](const std::vector<double>& a) {
    if (out) std::memcpy(out, a.data(), a.size() * sizeof(double));
    return (long)a.size();
  };
  auto cpi = [&](const std::vector<int>& a) {
    if (out) std::memcpy(out, a.data(), a.size() * sizeof(int));
    return (long)a.size();
  };
  std::string n(name);
#define D(f) if (n == #f) return cpd(m->f);
#define I(f) if (n == #f) return cpi(m->f);
  D(mx) D(my) D(mz) D(nx) D(ny) D(nz) D(d) D(e01x) D(e01y) D(e01z) D(e02x) D(e02y) D(e02z)
  D(e12x) D(e12y) D(e12z) D(d00) D(d00inv) D(d01) D(d11) D(d11inv) D(d22) D(d22inv) D(denom)
  D(minx) D(maxx) D(miny) D(maxy) D(minz) D(maxz) D(cx) D(cy) D(cz)
  I(i0) I(i1) I(i2) I(te0) I(te1) I(te2) I(sortedIndex)
  D(bminx) D(bminy) D(bminz) D(bmaxx) D(bmaxy) D(bmaxz) I(bleft) I(bright) I(btri)
#undef D
#undef I
  if (n == "mortonKey") {
    if (out) std::memcpy(out, m->mortonKey.data(), m->mortonKey.size() * 4);
    return (long)m->mortonKey.size();
  }
  if (n == "gmin") { if (out) std::memcpy(out, m->gmin, 24); return 3; }
  if (n == "gmax") { if (out) std::memcpy(out, m->gmax, 24); return 3; }
  if (n == "nEdge") return m->nEdge;
  if (n == "nodeCount") return m->nodeCount;
  if (n == "nTri") return m->nTri;
  return -1;
}

// params: [dt, mu, gx, gy, gz, minx, maxx, miny, maxy, minz, maxz,
//          cell dx, skin, refresh threshold, density, E, CoR]
void* orc_sim_create(int N, const double* x, const double* v, const double* angv,
                     const double* r, const int* active, void* mesh, const double* P) {
  Sim* s = new Sim();
  s->N = N;
  s->dt = P[0]; s->mu = P[1];
  s->g[0] = P[2]; s->g[1] = P[3]; s->g[2] = P[4];
  s->bmin[0] = P[5]; s->bmax[0] = P[6]; s->bmin[1] = P[7]; s->bmax[1] = P[8];
  s->bmin[2] = P[9]; s->bmax[2] = P[10];
  for (int a = 0; a < 3; ++a) { s->dx[a] = P[11]; s->invd[a] = 1. / P[11]; }
  s->skinR = P[12];
  s->thresh = P[13];
  s->threshSq = s->thresh * s->thresh;
  // grid sizes (BoundingBox.cu:394-396)
  s->ncell = 1;
  for (int a = 0; a < 3; ++a) {
    s->size[a] = (int)ceil((s->bmax[a] - s->bmin[a]) / s->dx[a]) + 2;
    s->ncell *= s->size[a];
  }
  s->mesh = static_cast<Mesh*>(mesh);
  auto zd = [&](std::vector<double>& a, size_t n) { a.assign(n, 0.0); };
  auto zi = [&](std::vector<int>& a, size_t n, int val) { a.assign(n, val); };
  s->x.assign(x, x + 3 * N);
  s->v.assign(v, v + 3 * N);
  s->angv.assign(angv, angv + 3 * N);
  zd(s->a, 3 * N); zd(s->f, 3 * N); zd(s->anga, 3 * N); zd(s->mom, 3 * N);
  for (auto* a : {&s->r, &s->rsq, &s->invr, &s->k, &s->m, &s->sqrtm, &s->invm, &s->moi,
                  &s->invmoi, &s->etac, &s->refx, &s->refy, &s->refz})
    zd(*a, N);
  for (int i = 0; i < N; ++i) set_material(s, i, r[i], P[14], P[15], P[16]);
  s->isActive.assign(active, active + N);
  s->pId.resize(N);
  for (int i = 0; i < N; ++i) s->pId[i] = i;
  zi(s->cellId, N, -1); zi(s->cellx, 3 * N, 0);
  zi(s->numCont, N, 0); zi(s->numContWall, N, 0); zi(s->numNei, N, 0); zi(s->numNeiWall, N, 0);
  const size_t R = (size_t)N * kMaxNei;
  zi(s->indHis, R, -1); zi(s->indHisWall, R, -1); zi(s->isContact, R, -1);
  zi(s->isContactWall, R, -1); zi(s->neiList, R, 0); zi(s->neiListWall, R, 0);
  zi(s->vorENow, R, 0);
  for (auto* a : {&s->dhx, &s->dhy, &s->dhz, &s->dwx, &s->dwy, &s->dwz}) zd(*a, R);
  zi(s->pNum, s->ncell, 0); zi(s->pStart, s->ncell, 0); zi(s->pList, N, 0);
  zi(s->cellOffset, s->ncell, 0);
  // initial lists are built on the host at set-up (main.cu:151, 196)
  rebuild(s);
  s->rebuilds = 0;
  return s;
}

void orc_sim_destroy(void* s) { delete static_cast<Sim*>(s); }

void orc_sim_step(void* sp, int nsteps, int gpu_order) {
  Sim* s = static_cast<Sim*>(sp);
  for (int k = 0; k < nsteps; ++k) step(s, gpu_order);
}

void orc_sim_rebuild(void* sp) { rebuild(static_cast<Sim*>(sp)); }

void orc_sim_set_record(void* sp, int on) {
  Sim* s = static_cast<Sim*>(sp);
  s->record = on;
  if (on) {
    size_t R = (size_t)s->N * 2 * kMaxNei;
    s->recN.assign(s->N, 0);
    s->recKey.assign(R, 0);
    s->recKind.assign(R, 0);
    s->recFn.assign(3 * R, 0.);
    s->recFt.assign(3 * R, 0.);
    s->recDel.assign(R, 0.);
  }
}

long orc_sim_get(void* sp, const char* name, void* out) {
  Sim* s = static_cast<Sim*>(sp);
  auto cpd = [&](const std::vector<double>& a) {
    if (out) std::memcpy(out, a.data(), a.size() * sizeof(double));
    return (long)a.size();
  };
  auto cpi = [&](const std::vector<int>& a) {
    if (out) std::memcpy(out, a.data(), a.size() * sizeof(int));
    return (long)a.size();
  };
  std::string n(name);
#define D(f) if (n == #f) return cpd(s->f);
#define I(f) if (n == #f) return cpi(s->f);
  D(x) D(v) D(a) D(f) D(angv) D(anga) D(mom) D(r) D(rsq) D(invr) D(k) D(m) D(sqrtm) D(invm)
  D(moi) D(invmoi) D(etac) D(refx) D(refy) D(refz) D(dhx) D(dhy) D(dhz) D(dwx) D(dwy) D(dwz)
  D(recFn) D(recFt) D(recDel)
  I(isActive) I(pId) I(cellId) I(cellx) I(numCont) I(numContWall) I(numNei) I(numNeiWall)
  I(indHis) I(indHisWall) I(neiList) I(neiListWall) I(pNum) I(pStart) I(pList)
  I(recN) I(recKey) I(recKind)
#undef D
#undef I
  if (n == "steps") return (long)s->steps;
  if (n == "rebuilds") return (long)s->rebuilds;
  if (n == "overflow") return s->overflow;
  if (n == "size") { if (out) std::memcpy(out, s->size, 12); return 3; }
  if (n == "ncell") return s->ncell;
  return -1;
}

// Overwrite a state field (used to plant states in tests).
long orc_sim_set(void* sp, const char* name, const void* in) {
  Sim* s = static_cast<Sim*>(sp);
  std::string n(name);
  auto sd = [&](std::vector<double>& a) { std::memcpy(a.data(), in, a.size() * 8); return (long)a.size(); };
  auto si = [&](std::vector<int>& a) { std::memcpy(a.data(), in, a.size() * 4); return (long)a.size(); };
  if (n == "x") return sd(s->x);
  if (n == "v") return sd(s->v);
  if (n == "angv") return sd(s->angv);
  if (n == "isActive") return si(s->isActive);
  return -1;
}

void orc_sim_brute_lists(void* sp, int* pair, int* npair, int* wall, int* nwall) {
  Sim* s = static_cast<Sim*>(sp);
  brute_pair_lists(s, pair, npair);
  brute_wall_lists(s, wall, nwall);
}

// Morton keys of the current cell coordinates of active particles
// (dk_morton_key, BoundingBox.cu:1007-1023); inactive entries get 0xffffffff.
void orc_sim_morton(void* sp, uint32_t* key) {
  Sim* s = static_cast<Sim*>(sp);
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) { key[i] = 0xffffffffu; continue; }
    int c[3];
    cell_of(s, i, c);
    key[i] = morton3((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
  }
}

// Morton keys of the cell coordinates stored by the last binning (the
// positions of the last list rebuild); inactive entries get 0xffffffff.
void orc_sim_morton_built(void* sp, uint32_t* key) {
  Sim* s = static_cast<Sim*>(sp);
  for (int i = 0; i < s->N; ++i) {
    if (s->isActive[i] != 1) { key[i] = 0xffffffffu; continue; }
    const int* c = &s->cellx[i * 3];
    key[i] = morton3((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
  }
}

uint32_t orc_morton3(uint32_t ix, uint32_t iy, uint32_t iz) { return morton3(ix, iy, iz); }

void orc_radix_sort(uint32_t* key, int* idx, int n) {
  std::vector<uint32_t> k(key, key + n);
  std::vector<int> i(idx, idx + n);
  lsd_sort(k, i);
  std::memcpy(key, k.data(), 4 * (size_t)n);
  std::memcpy(idx, i.data(), 4 * (size_t)n);
}

// dist_triangle for one point: out = [nx, ny, nz, dist], returns hitAt.
int orc_tri_distance(void* mp, int tri, const double* p, double Rsq, double* out) {
  TriHit h = tri_distance(static_cast<Mesh*>(mp), V3{p[0], p[1], p[2]}, Rsq, tri);
  out[0] = h.n.x; out[1] = h.n.y; out[2] = h.n.z; out[3] = h.dist;
  return h.hitAt;
}

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
