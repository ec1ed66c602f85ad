// gfx950 kernels of the gdem contact engine.
//
// One thread owns one particle (internal Morton order).  Particle state is
// double-buffered (read buffer `cur`, write buffer `cur^1`), so forces and
// integration run in one fused pass and no force/torque arrays or per-step
// memsets exist.  Candidate lists and contact histories are slot-major
// ([slot][N]) so lane i of a wave reads consecutive words.  A list rebuild is
// decided on the device (refresh flag) and every rebuild kernel exits at once
// when the flag is clear, so a step needs no host synchronisation.
//
// FP64 arithmetic follows the reference operation for operation (the build
// uses -ffp-contract=off): pair contact d_particle_collision_verlet +
// d_calc_normal_force + d_calc_tangential_force (device_dem.cu:916-981,
// 793-847, 622-740), wall contact d_wall_collision_verlet +
// d_calc_normal_force_wall + d_calc_tangential_force_wall + d_dist_triangle
// (device_dem.cu:288-499, 13-205), history update d_update_history(_wall)
// (device_dem.cu:743-790), integration k_integrate (device_dem.cu:1065-1091),
// dk_checkOoB (:1116-1125) and k_shouldRefreshNeighborList (:1095-1112).
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "gdem_kernels.h"

namespace gdem {

// ---------------------------------------------------------------------------
// small vector helpers (Vec3.h:23-70 operation order)
// ---------------------------------------------------------------------------
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
__device__ __forceinline__ T dot(V3<T> a, V3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
__device__ __forceinline__ V3<T> scal(T s, V3<T> b) {
  return V3<T>{s * b.x, s * b.y, s * b.z};
}
template <class T>
__device__ __forceinline__ V3<T> vsub(V3<T> a, V3<T> b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
__device__ __forceinline__ V3<T> vadd(V3<T> a, V3<T> b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
__device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
// morton3D (BoundingBox.h:125-135)
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (spread10(x) << 2) | (spread10(y) << 1) | spread10(z);
}

// Non-temporal (streaming) accesses for data a step touches once, so the L2
// keeps the neighbour positions and velocities that tiles gather.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_nt(T v, T* p) {
  __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ double4 ld_nt(const double4* p) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_double4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st_nt(double4 a, double4* p) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  const v4 v = {a.x, a.y, a.z, a.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4*>(p));
}
__device__ __forceinline__ void st_nt(float4 a, float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = {a.x, a.y, a.z, a.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4*>(p));
}

// Last-workgroup detection: returns true in exactly one thread (thread 0 of
// the last workgroup to finish).  No waiting is involved.  Callers use it only
// to flip control words that every workgroup read at its start, so no memory
// fence is needed (a device-scope fence would write back the L2 of the XCD
// once per workgroup); results reach the next kernel through the kernel
// boundary.
__device__ __forceinline__ bool last_block(int* counter) {
  __shared__ int is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (prev == (int)gridDim.x - 1);
  }
  __syncthreads();
  return is_last && threadIdx.x == 0;
}

// ---------------------------------------------------------------------------
// sphere-triangle distance (d_dist_triangle, device_dem.cu:13-205; the list
// builder's twin d_dist_triangle_neighbor, BVH.cu:297-491, differs only in the
// squared radius it compares against)
// ---------------------------------------------------------------------------
struct TriHit {
  V3<double> n;
  double dist;
  int hit;
};

__device__ __forceinline__ TriHit tri_distance(const TriGeo* __restrict__ tg, double px, double py,
                                               double pz, double Rsq) {
  TriHit res;
  res.n = V3<double>{0., 0., 0.};
  res.dist = 1.e10;
  res.hit = -1;
  const V3<double> nt{tg->n[0], tg->n[1], tg->n[2]};
  double dist = 0.;
  dist += nt.x * px;
  dist += nt.y * py;
  dist += nt.z * pz;
  dist += tg->d;
  double distsq = dist * dist;
  if (distsq > Rsq) return res;
  const V3<double> pos{px, py, pz};
  const V3<double> P0{tg->p0[0], tg->p0[1], tg->p0[2]};
  const V3<double> ap = vsub(pos, P0);
  const double d00 = tg->d00, d01 = tg->d01, d11 = tg->d11, den = tg->denom;
  const double d20 = ap.x * tg->e01[0] + ap.y * tg->e01[1] + ap.z * tg->e01[2];
  const double d21 = ap.x * tg->e02[0] + ap.y * tg->e02[1] + ap.z * tg->e02[2];
  const double v = (d11 * d20 - d01 * d21) * den;
  const double w = (d00 * d21 - d01 * d20) * den;
  const double u = 1.0 - v - w;
  if (u > 0. && v > 0. && w > 0.) {
    const double sg = dist < 0 ? -1.0 : 1.0;
    res.n = scal(sg, nt);
    res.dist = sg * dist;
    res.hit = -1;
    return res;
  }
  V3<double> q;
  int hit;
  if (v <= 0. && w <= 0.0) {
    q = P0;
    hit = tg->v0;
  } else if (u <= 0. && w <= 0.0) {
    q = V3<double>{tg->p1[0], tg->p1[1], tg->p1[2]};
    hit = tg->v1;
  } else if (u <= 0. && v <= 0.0) {
    q = V3<double>{tg->p2[0], tg->p2[1], tg->p2[2]};
    hit = tg->v2;
  } else if (u <= 0.0) {
    const V3<double> P1{tg->p1[0], tg->p1[1], tg->p1[2]};
    const V3<double> E{tg->e12[0], tg->e12[1], tg->e12[2]};
    double t = dot(vsub(pos, P1), E) * tg->d22inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P1, scal(t, E));
    hit = tg->e1;
  } else if (v <= 0.0) {
    const V3<double> E{tg->e02[0], tg->e02[1], tg->e02[2]};
    double t = d21 * tg->d11inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P0, scal(t, E));
    hit = tg->e2;
  } else {
    const V3<double> E{tg->e01[0], tg->e01[1], tg->e01[2]};
    double t = d20 * tg->d00inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P0, scal(t, E));
    hit = tg->e0;
  }
  res.n = vsub(pos, q);
  distsq = dot(res.n, res.n);
  if (distsq < Rsq) {
    res.dist = sqrt(distsq);
    res.n = scal(1. / res.dist, res.n);
  }
  res.hit = hit;
  return res;
}

// ---------------------------------------------------------------------------
// contact-history row of one particle
// ---------------------------------------------------------------------------
// Global layout is slot-major: keys [cap][N] (int) and deltas [cap][N]
// (4-vector, w unused), so lane i of a wave touches consecutive words.
// HistLds keeps the live keys in an LDS column of the thread (CAP >= the
// scene's capacity): they are loaded once per step with independent loads,
// and the per-contact search costs no memory round trip.  HistScan searches
// the keys in memory and serves capacities beyond the LDS budget.
template <class H, int CAP>
struct HistLds {
  using H4 = typename Vec4<H>::type;
  int* gkey;
  H4* gd;
  int s, N, cap;
  int n;           // live slots
  uint64_t marks;  // slots in contact this step
  int* k;          // this thread's LDS column: k[q * blockDim.x]
  __device__ __forceinline__ void load(int nh) {
    n = nh;
    marks = 0;
#pragma unroll
    for (int q = 0; q < CAP; ++q)
      if (q < nh) k[q * 256] = ld_nt(gkey + (size_t)q * N + s);
  }
  __device__ __forceinline__ int find(int key) const {
    for (int q = 0; q < n; ++q)
      if (k[q * 256] == key) return q;
    return -1;
  }
  __device__ __forceinline__ void remember(int key) { k[n * 256] = key; }
};

template <class H>
struct HistScan {
  using H4 = typename Vec4<H>::type;
  int* gkey;
  H4* gd;
  int s, N, cap;
  int n;
  uint64_t marks;
  __device__ __forceinline__ void load(int nh) {
    n = nh;
    marks = 0;
  }
  __device__ __forceinline__ int find(int key) const {
    for (int q = 0; q < n; ++q)
      if (gkey[(size_t)q * N + s] == key) return q;
    return -1;
  }
  __device__ __forceinline__ void remember(int) {}
};

// Tangential spring with Coulomb cap and history (d_calc_tangential_force,
// device_dem.cu:622-740; the wall twin :403-499 is identical with the wall
// history).  The slot of `key` is found or appended (d_calc_tangential_force
// :630-653); T is the force type and FP64 keeps the reference's operations.
template <class T, class Row>
__device__ __forceinline__ V3<T> tangential_at(Row& row, uint32_t& err, uint32_t ovf_bit, int key, int slot,
                                               V3<T> old, V3<T> fn, V3<T> vt, V3<T> n, T eta, T k, T mu,
                                               T dt) {
  using H4 = typename Row::H4;
  if (slot < 0) {
    // new contact: appended in list order (old is zero)
    if (row.n < row.cap) {
      slot = row.n;
      st_nt(key, row.gkey + (size_t)slot * row.N + row.s);
      row.remember(key);
      ++row.n;
    } else {
      err |= ovf_bit;
    }
  }
  if (slot >= 0) row.marks |= (uint64_t)1 << slot;
  V3<T> dn{old.x + vt.x * dt, old.y + vt.y * dt, old.z + vt.z * dt};
  const T dd = dot(dn, n);
  dn = vsub(dn, scal(dd, n));
  V3<T> ft = scal(-k, dn);
  const T ftsq = dot(ft, ft);
  const T fnsq = dot(fn, fn);
  if (ftsq > (mu * mu) * fnsq) {
    if (ftsq != T(0)) {
      const V3<T> t = scal(T(1) / (sqrt(ftsq) + T(1e-15)), ft);
      const T fnn = sqrt(fnsq);
      ft = scal(mu * fnn, t);
      dn = scal(T(-1) / k, ft);
    } else {
      ft = V3<T>{T(0), T(0), T(0)};
    }
  } else {
    ft.x -= eta * vt.x;
    ft.y -= eta * vt.y;
    ft.z -= eta * vt.z;
  }
  if (slot >= 0) {
    H4 o;
    o.x = (decltype(o.x))dn.x;
    o.y = (decltype(o.y))dn.y;
    o.z = (decltype(o.z))dn.z;
    o.w = 0;
    st_nt(o, row.gd + (size_t)slot * row.N + row.s);
  }
  return ft;
}

// Slot lookup + history read, then tangential_at.
template <class T, class Row>
__device__ __forceinline__ V3<T> tangential(Row& row, uint32_t& err, uint32_t ovf_bit, int key, V3<T> fn,
                                            V3<T> vt, V3<T> n, T eta, T k, T mu, T dt) {
  using H4 = typename Row::H4;
  const int slot = row.find(key);
  V3<T> old{T(0), T(0), T(0)};
  if (slot >= 0) {
    const H4 d = ld_nt(row.gd + (size_t)slot * row.N + row.s);
    old = V3<T>{(T)d.x, (T)d.y, (T)d.z};
  }
  return tangential_at<T>(row, err, ovf_bit, key, slot, old, fn, vt, n, eta, k, mu, dt);
}

// Slots from the first one without a contact this step are dropped
// (d_update_history, device_dem.cu:768-790: the moved-in slot's flag is not
// copied, so the removal repeats at the same index until the row ends there).
__device__ __forceinline__ int truncate_history(int nh, uint64_t marks) {
  const uint64_t free_slots = ~marks;
  const int first = free_slots ? (__ffsll((unsigned long long)free_slots) - 1) : 64;
  return nh < first ? nh : first;
}

template <class T>
struct Rec {
  int cap;
  int* count;
  int* key;
  int* kind;
  double* fn;
  double* ft;
  double* delta;
};

// The hardware hands workgroup b to XCD b % 8.  Giving each XCD a contiguous
// run of tiles keeps Morton-adjacent particles -- the neighbours a tile
// gathers -- inside one XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int X = 8;
  if (nb < X) return b;
  const int per = nb / X, rem = nb % X;
  const int x = b % X, i = b / X;
  return x * per + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------------------
// fused contact forces + integration for one step
// ---------------------------------------------------------------------------
// Touching pairs evaluated per round of pass 2 (their loads overlap).
#ifndef GDEM_PAIRS_PER_ROUND
#define GDEM_PAIRS_PER_ROUND 1
#endif
constexpr int PAIRS_PER_ROUND = GDEM_PAIRS_PER_ROUND;

// PRow/WRow: history row types for pairs and walls.  Pair candidates are
// screened first (indices and positions fetched eight at a time so their
// latencies overlap, FP64 overlap test), then the touching ones are evaluated
// in list order.  The occupancy bound of 4 waves per SIMD costs the FP64 form
// a few spilled registers and pays for itself (measured on MI355X).
template <int PC, int WC>
struct StepLds {
  int pkey[PC > 0 ? PC : 1][256];
  int wkey[WC > 0 ? WC : 1][256];
  int ve[8][256];
};

template <class Row>
__device__ __forceinline__ void bind_lds(Row&, int*) {}
template <class H, int CAP>
__device__ __forceinline__ void bind_lds(HistLds<H, CAP>& r, int* col) { r.k = col; }

template <class Row>
struct RowCap {
  static constexpr int value = 0;
};
template <class H, int CAP>
struct RowCap<HistLds<H, CAP>> {
  static constexpr int value = CAP;
};

template <class T, class H, bool REC, class PRow, class WRow>
__global__ __launch_bounds__(256, 4) void k_step(SimParams P, Bufs<H> B, Rec<T> rec) {
  using H4 = typename Vec4<H>::type;
  __shared__ StepLds<RowCap<PRow>::value, RowCap<WRow>::value> lds;
  const int tx = threadIdx.x;
  const int s = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int cur = __builtin_amdgcn_readfirstlane(B.ctl[CTL_CUR]);
  const int hc = __builtin_amdgcn_readfirstlane(B.ctl[CTL_HCUR]);
  const int N = P.N;
  const double4* __restrict__ pin = B.pos[cur];
  const double4* __restrict__ vin = B.vel[cur];
  const double4* __restrict__ win = B.angv[cur];
  uint32_t err = 0;
  bool refresh = false;
  if (s < N) {
    const double4 p = pin[s];
    const double4 v = vin[s];
    const double4 w = win[s];
    if (p.w < 0.) {
      st_nt(p, B.pos[cur ^ 1] + s);
      st_nt(v, B.vel[cur ^ 1] + s);
      st_nt(w, B.angv[cur ^ 1] + s);
    } else {
      const H4 mt = ld_nt(B.mat[hc] + s);
      const T k = (T)mt.x, etac = (T)mt.y, sqrtm = (T)mt.z;
      const double invmoi = (double)mt.w;
      const double ri = p.w, invm = v.w;
      const T riT = (T)ri;
      const T mu = (T)P.mu, dtT = (T)P.dt;
      const V3<T> vi{(T)v.x, (T)v.y, (T)v.z};
      const V3<T> wi{(T)w.x, (T)w.y, (T)w.z};
      V3<T> f{T(0), T(0), T(0)}, mom{T(0), T(0), T(0)};
      int nrec = 0;
      const int self_id = (int)w.w;
      const int nnb = ld_nt(B.nnb + s);
      const int nwl = ld_nt(B.nwl + s);
      const double4 rf = ld_nt(B.ref + s);  // position at the last list build

      // ---- particle pairs -------------------------------------------------
      {
        PRow row;
        bind_lds(row, &lds.pkey[0][tx]);
        row.gkey = B.hkey[hc];
        row.gd = B.hd[hc];
        row.s = s;
        row.N = N;
        row.cap = P.hist_cap;
        row.load(ld_nt(B.nh[hc] + s));
        // pass 1: which candidates overlap (FP64 test of the reference, list
        // order); candidates are fetched eight at a time so latencies overlap
        uint64_t hits = 0;
        for (int q0 = 0; q0 < nnb; q0 += 8) {
          int jv[8];
          double4 cv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) jv[u] = (q0 + u < nnb) ? B.nbr[(size_t)(q0 + u) * N + s].x : -1;
#pragma unroll
          for (int u = 0; u < 8; ++u) cv[u] = jv[u] >= 0 ? pin[jv[u]] : make_double4(0., 0., 0., -1.);
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const double4 pj = cv[u];
            if (pj.w < 0.) continue;
            const double dx = p.x - pj.x, dy = p.y - pj.y, dz = p.z - pj.z;
            const double distsq = dx * dx + dy * dy + dz * dz;
            const double R = ri + pj.w;
            if (distsq < R * R) hits |= (uint64_t)1 << (q0 + u);
          }
        }
        // pass 2: the touching pairs in list order, two at a time.  Their
        // history slots are looked up first (new slots are still appended in
        // list order), so position, velocity, spin and history of both are
        // fetched in one round.
        while (hits) {
          const int qa = __ffsll((unsigned long long)hits) - 1;
          hits &= hits - 1;
          const int qb = (PAIRS_PER_ROUND > 1 && hits) ? __ffsll((unsigned long long)hits) - 1 : -1;
          if (qb >= 0) hits &= hits - 1;
          int2 e0 = B.nbr[(size_t)qa * N + s];
          int2 e1 = qb >= 0 ? B.nbr[(size_t)qb * N + s] : make_int2(-1, -1);
          int s0 = row.find(e0.y);
          int s1 = e1.x >= 0 ? row.find(e1.y) : -1;
          double4 pj0 = pin[e0.x], vj0 = vin[e0.x], wj0 = win[e0.x];
          H4 d0 = s0 >= 0 ? ld_nt(row.gd + (size_t)s0 * N + s) : H4{};
          double4 pj1 = make_double4(0., 0., 0., 0.), vj1 = pj1, wj1 = pj1;
          H4 d1{};
          if (e1.x >= 0) {
            pj1 = pin[e1.x];
            vj1 = vin[e1.x];
            wj1 = win[e1.x];
            if (s1 >= 0) d1 = ld_nt(row.gd + (size_t)s1 * N + s);
          }
          for (int u = 0; u < 2; ++u) {
            if (u == 1) {
              if (e1.x < 0) break;
              e0 = e1;
              s0 = s1;
              pj0 = pj1;
              vj0 = vj1;
              wj0 = wj1;
              d0 = d1;
            }
            const double4 pj = pj0, vj = vj0, wj = wj0;
            const double dx = p.x - pj.x, dy = p.y - pj.y, dz = p.z - pj.z;
            const double distsq = dx * dx + dy * dy + dz * dz;
            const double R = ri + pj.w;
            const double dist = sqrt(distsq);
            const double delMagD = R - dist;
            V3<T> n;
            if constexpr (sizeof(T) == 8) {
              n = V3<T>{dx / dist, dy / dist, dz / dist};
            } else {
              const T inv = T(1) / (T)dist;
              n = V3<T>{(T)dx * inv, (T)dy * inv, (T)dz * inv};
            }
            const T delMag = (T)delMagD;
            const V3<T> del = scal(-delMag, n);
            const V3<T> vr{vi.x - (T)vj.x, vi.y - (T)vj.y, vi.z - (T)vj.z};
            const T vrn = dot(vr, n);
            const V3<T> vn = scal(vrn, n);
            const T meff = T(1) / ((T)invm + (T)vj.w);
            const T eta = etac * sqrt(meff);
            V3<T> fn;
            fn.x = -k * del.x - eta * vn.x;
            fn.y = -k * del.y - eta * vn.y;
            fn.z = -k * del.z - eta * vn.z;
            const T rj = (T)pj.w;
            V3<T> vrot{riT * wi.x + rj * (T)wj.x, riT * wi.y + rj * (T)wj.y, riT * wi.z + rj * (T)wj.z};
            vrot = cross(vrot, n);
            const V3<T> vt = vadd(vsub(vr, vn), vrot);
            f.x += fn.x;
            f.y += fn.y;
            f.z += fn.z;
            const int key = e0.y;  // original id of j
            const V3<T> ft = tangential_at<T>(row, err, GDEM_ERR_HIST_OVERFLOW_BIT, key, s0,
                                                 V3<T>{(T)d0.x, (T)d0.y, (T)d0.z}, fn, vt, n, eta, k, mu, dtT);
            f.x += ft.x;
            f.y += ft.y;
            f.z += ft.z;
            const V3<T> mm = scal(riT, cross(n, ft));
            mom.x += mm.x;
            mom.y += mm.y;
            mom.z += mm.z;
            if constexpr (REC) {
              if (nrec < rec.cap) {
                const size_t o = (size_t)self_id * rec.cap + nrec;
                rec.key[o] = key;
                rec.kind[o] = 0;
                rec.fn[o * 3] = fn.x; rec.fn[o * 3 + 1] = fn.y; rec.fn[o * 3 + 2] = fn.z;
                rec.ft[o * 3] = ft.x; rec.ft[o * 3 + 1] = ft.y; rec.ft[o * 3 + 2] = ft.z;
                rec.delta[o] = delMagD;
                ++nrec;
              }
            }
          }
        }
        st_nt(truncate_history(row.n, row.marks), B.nh[hc] + s);
      }

      // ---- walls ------------------------------------------------------------
      {
        WRow row;
        bind_lds(row, &lds.wkey[0][tx]);
        row.gkey = B.wkey[hc];
        row.gd = B.wd[hc];
        row.s = s;
        row.N = N;
        row.cap = P.whist_cap;
        row.load(ld_nt(B.nw[hc] + s));
        int* ve = &lds.ve[0][tx];  // vertex/edge features touched this step, ve[u * 256]
        int nve = 0;
        const double rsq = ri * ri;
        const T eta = etac * sqrtm;
        for (int q = 0; q < nwl; ++q) {
          const int t = ld_nt(B.wl + (size_t)q * N + s);
          const TriHit h = tri_distance(B.tri + t, p.x, p.y, p.z, rsq);
          if (!(h.dist < ri)) continue;
          if (h.hit != -1) {
            bool dup = false;
            const int seen = nve < 8 ? nve : 8;
            for (int u = 0; u < seen; ++u)
              if (ve[u * 256] == h.hit) {
                dup = true;
                break;
              }
            if (dup) continue;
            if (nve < 8) ve[nve * 256] = h.hit;
            else err |= GDEM_ERR_VE_OVERFLOW_BIT;
            ++nve;
          }
          const double delMagD = ri - h.dist;
          const T delMag = (T)delMagD;
          const V3<T> n{(T)h.n.x, (T)h.n.y, (T)h.n.z};
          const V3<T> del = scal(-delMag, n);
          const T vrn = dot(vi, n);
          const V3<T> vn = scal(vrn, n);
          V3<T> fn;
          fn.x = -k * del.x - eta * vn.x;
          fn.y = -k * del.y - eta * vn.y;
          fn.z = -k * del.z - eta * vn.z;
          V3<T> vrot{riT * wi.x, riT * wi.y, riT * wi.z};
          vrot = cross(vrot, n);
          const V3<T> vt = vadd(vsub(vi, vn), vrot);
          f.x += fn.x;
          f.y += fn.y;
          f.z += fn.z;
          const V3<T> ft = tangential<T>(row, err, GDEM_ERR_WHIST_OVERFLOW_BIT, t, fn, vt, n, eta, k, mu, dtT);
          f.x += ft.x;
          f.y += ft.y;
          f.z += ft.z;
          const V3<T> mm = scal(riT, cross(n, ft));
          mom.x += mm.x;
          mom.y += mm.y;
          mom.z += mm.z;
          if constexpr (REC) {
            if (nrec < rec.cap) {
              const size_t o = (size_t)self_id * rec.cap + nrec;
              rec.key[o] = t;
              rec.kind[o] = 1;
              rec.fn[o * 3] = fn.x; rec.fn[o * 3 + 1] = fn.y; rec.fn[o * 3 + 2] = fn.z;
              rec.ft[o * 3] = ft.x; rec.ft[o * 3 + 1] = ft.y; rec.ft[o * 3 + 2] = ft.z;
              rec.delta[o] = delMagD;
              ++nrec;
            }
          }
        }
        st_nt(truncate_history(row.n, row.marks), B.nw[hc] + s);
      }
      if constexpr (REC) rec.count[self_id] = nrec;

      // ---- integration (k_integrate), out-of-bounds, refresh check ----------
      const double dt = P.dt;
      const double ax = (double)f.x * invm + P.g[0];
      const double ay = (double)f.y * invm + P.g[1];
      const double az = (double)f.z * invm + P.g[2];
      double4 vo, po, wo;
      vo.x = v.x + ax * dt;
      vo.y = v.y + ay * dt;
      vo.z = v.z + az * dt;
      vo.w = v.w;
      po.x = p.x + vo.x * dt;
      po.y = p.y + vo.y * dt;
      po.z = p.z + vo.z * dt;
      const double bx = (double)mom.x * invmoi;
      const double by = (double)mom.y * invmoi;
      const double bz = (double)mom.z * invmoi;
      wo.x = w.x + bx * dt;
      wo.y = w.y + by * dt;
      wo.z = w.z + bz * dt;
      wo.w = w.w;
      const bool out = po.x < P.bmin[0] || po.x > P.bmax[0] || po.y < P.bmin[1] ||
                       po.y > P.bmax[1] || po.z < P.bmin[2] || po.z > P.bmax[2];
      po.w = out ? -ri : ri;
      if (!out) {
        const double ex = po.x - rf.x, ey = po.y - rf.y, ez = po.z - rf.z;
        refresh = (ex * ex + ey * ey + ez * ez) > P.threshSq;
      }
      if (!(isfinite(po.x) && isfinite(po.y) && isfinite(po.z) && isfinite(vo.x) &&
            isfinite(vo.y) && isfinite(vo.z)))
        err |= GDEM_ERR_NAN_BIT;
      st_nt(po, B.pos[cur ^ 1] + s);
      st_nt(vo, B.vel[cur ^ 1] + s);
      st_nt(wo, B.angv[cur ^ 1] + s);
    }
  }
  if (__ballot(refresh) != 0 && __lane_id() == 0) B.ctl[CTL_REFRESH] = 1;
  if (err) atomicOr((unsigned int*)&B.ctl[CTL_ERR], err);
  if (last_block(&B.ctl[CTL_DONE_STEP])) {
    B.ctl[CTL_DONE_STEP] = 0;
    atomicExch(&B.ctl[CTL_CUR], cur ^ 1);
  }
}

// ---------------------------------------------------------------------------
// list rebuild (each phase is a per-element device function used both by the
// launch chain and by the single cooperative rebuild kernel)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rebuild_due(const int* ctl) {
  return __builtin_amdgcn_readfirstlane(ctl[CTL_REFRESH]) != 0;
}

// Cell coordinates (d_calcCellId, BoundingBox.cu:1089-1104) and sort keys.
// Also clears the table entry of the cell this particle occupied at the last
// rebuild, so the new table holds exactly the occupied cells.
template <class H>
__device__ __forceinline__ void keys_one(const SimParams& P, const Bufs<H>& B, int s, int cur) {
  const int old = B.cell_of[s];
  if (old >= 0) B.cell_tab[old] = make_int2(-1, -1);
  const double4 p = B.pos[cur][s];
  uint32_t key = 0xffffffffu;
  int cid = -1;
  if (p.w >= 0.) {
    const int cx = (int)floor((p.x - P.bmin[0]) * P.invd[0]) + 1;
    const int cy = (int)floor((p.y - P.bmin[1]) * P.invd[1]) + 1;
    const int cz = (int)floor((p.z - P.bmin[2]) * P.invd[2]) + 1;
    cid = (P.size[1] * cz + cy) * P.size[0] + cx;
    key = P.use_morton ? morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) : (uint32_t)cid;
  }
  B.key[0][s] = key;
  B.idx[0][s] = s;
  B.cell_tmp[s] = cid;
}

// ---- stable LSD radix sort, 8-bit digits ----------------------------------
// Workgroup b owns the contiguous chunk [b*chunk, (b+1)*chunk); per-(digit,
// chunk) counts are scanned digit-major so equal digits keep chunk order, and
// inside a chunk keys are ranked in arrival order with wave ballots.
constexpr int RS_THREADS = 256;

__device__ __forceinline__ void rs_count_block(int n, int chunk, int shift, const uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ hist, int b, int nb) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int beg = b * chunk;
  const int end = min(n, beg + chunk);
  for (int i = beg + threadIdx.x; i < end; i += RS_THREADS) atomicAdd(&cnt[(keys[i] >> shift) & 0xffu], 1u);
  __syncthreads();
  hist[(size_t)threadIdx.x * nb + b] = cnt[threadIdx.x];
  __syncthreads();
}

// Exclusive scan of m counters in place by one workgroup of RS_THREADS
// threads; returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t* __restrict__ v, int m) {
  __shared__ uint32_t part[RS_THREADS];
  const int per = (m + RS_THREADS - 1) / RS_THREADS;
  const int beg = threadIdx.x * per;
  const int end = min(m, beg + per);
  uint32_t sum = 0;
  for (int i = beg; i < end; ++i) sum += v[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < RS_THREADS; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (int i = beg; i < end; ++i) {
    const uint32_t c = v[i];
    v[i] = run;
    run += c;
  }
  const uint32_t total = part[RS_THREADS - 1];
  __syncthreads();
  return total;
}

// Stable scatter of workgroup b's chunk.  base[d] must hold the global
// position of the first key with digit d in this chunk.
__device__ __forceinline__ void rs_scatter_block(int n, int chunk, int shift, const uint32_t* __restrict__ kin,
                                                 const int* __restrict__ iin, uint32_t* __restrict__ kout,
                                                 int* __restrict__ iout, uint32_t* base, int b) {
  constexpr int NW = RS_THREADS / 64;
  __shared__ uint32_t wcnt[NW][256];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  for (int w = 0; w < NW; ++w) wcnt[w][tid] = 0;
  __syncthreads();
  const int beg = b * chunk;
  const int end = min(n, beg + chunk);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r0 = beg; r0 < end; r0 += RS_THREADS) {
    const int i = r0 + tid;
    const bool valid = i < end;
    const uint32_t k = valid ? kin[i] : 0u;
    const int idx = valid ? iin[i] : 0;
    const uint32_t d = (k >> shift) & 0xffu;
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bal = __ballot((d >> bit) & 1u);
      match &= ((d >> bit) & 1u) ? bal : ~bal;
    }
    const int rank = __popcll(match & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(match);
    __syncthreads();
    uint32_t before = 0;
    if (valid)
      for (int w = 0; w < wave; ++w) before += wcnt[w][d];
    const uint32_t pos = base[d] + before + (uint32_t)rank;
    __syncthreads();
    {
      uint32_t tot = 0;
      for (int w = 0; w < NW; ++w) {
        tot += wcnt[w][tid];
        wcnt[w][tid] = 0;
      }
      base[tid] += tot;
    }
    if (valid) {
      kout[pos] = k;
      iout[pos] = idx;
    }
    __syncthreads();
  }
}

// Gather every per-particle array into sorted order (d_swap_ps,
// BoundingBox.cu:928-1005, but including history rows, stiffness and ids).
template <class H>
__device__ __forceinline__ void permute_one(const SimParams& P, const Bufs<H>& B, int s, int cur, int hc) {
  const int N = P.N;
  const int src = B.idx[0][s];
  B.pos[cur ^ 1][s] = B.pos[cur][src];
  B.vel[cur ^ 1][s] = B.vel[cur][src];
  B.angv[cur ^ 1][s] = B.angv[cur][src];
  B.mat[hc ^ 1][s] = B.mat[hc][src];
  B.pid[hc ^ 1][s] = B.pid[hc][src];
  B.cell_of[s] = B.cell_tmp[src];
  const int nh = B.nh[hc][src];
  B.nh[hc ^ 1][s] = nh;
  for (int q = 0; q < nh; ++q) {
    B.hkey[hc ^ 1][(size_t)q * N + s] = B.hkey[hc][(size_t)q * N + src];
    B.hd[hc ^ 1][(size_t)q * N + s] = B.hd[hc][(size_t)q * N + src];
  }
  const int nw = B.nw[hc][src];
  B.nw[hc ^ 1][s] = nw;
  for (int q = 0; q < nw; ++q) {
    B.wkey[hc ^ 1][(size_t)q * N + s] = B.wkey[hc][(size_t)q * N + src];
    B.wd[hc ^ 1][(size_t)q * N + s] = B.wd[hc][(size_t)q * N + src];
  }
}

// Cell table from the sorted keys: first/end of each occupied cell (the
// entries of cells occupied at the last rebuild were cleared by keys_one).
template <class H>
__device__ __forceinline__ void cells_one(const SimParams& P, const Bufs<H>& B, int s) {
  const int N = P.N;
  const uint32_t* key = B.key[0];
  const uint32_t k = key[s];
  if (k == 0xffffffffu) {
    if (s == 0) B.ctl[CTL_NACT] = 0;
    return;
  }
  const int c = B.cell_of[s];
  if (s == 0 || key[s - 1] != k) B.cell_tab[c].x = s;
  const bool last = (s == N - 1) || key[s + 1] != k;
  if (last) B.cell_tab[c].y = s + 1;
  if ((s == N - 1) || key[s + 1] == 0xffffffffu) B.ctl[CTL_NACT] = s + 1;
}

// Verlet pair list (k_update_neighborlist_endsort, BoundingBox.cu:1106-1162):
// all active j != i in the 27 surrounding cells with d.d < (r_i+r_j+skin)^2,
// ordered by original id.  Also records the reference position.
constexpr int NB_MAX = 64;

template <class H>
__device__ __forceinline__ void pairlist_one(const SimParams& P, const Bufs<H>& B, int s, int cur, int hc,
                                             int nact) {
  const int N = P.N;
  const double4* __restrict__ pos = B.pos[cur];
  const int* __restrict__ pid = B.pid[hc];
  const double4 p = pos[s];
  B.ref[s] = p;
  if (s >= nact) {
    B.nnb[s] = 0;
    return;
  }
  const int c = B.cell_of[s];
  const int sx = P.size[0], sy = P.size[1], sz = P.size[2];
  const int cx = c % sx;
  const int cy = (c / sx) % sy;
  const int cz = c / (sx * sy);
  int cand[NB_MAX];
  int cpid[NB_MAX];
  int n = 0;
  bool ovf = false;
  const int cap = P.nb_cap < NB_MAX ? P.nb_cap : NB_MAX;
  for (int dz = -1; dz <= 1; ++dz) {
    const int Z = cz + dz;
    if (Z < 0 || Z >= sz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int Y = cy + dy;
      if (Y < 0 || Y >= sy) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int X = cx + dx;
        if (X < 0 || X >= sx) continue;
        const int cc = (sy * Z + Y) * sx + X;
        const int2 ce = B.cell_tab[cc];
        if (ce.x < 0) continue;
        const int st = ce.x, en = ce.y;
        for (int q = st; q < en; ++q) {
          if (q == s) continue;
          const double4 pj = pos[q];
          const double ddx = p.x - pj.x, ddy = p.y - pj.y, ddz = p.z - pj.z;
          const double distsq = ddx * ddx + ddy * ddy + ddz * ddz;
          const double R = p.w + pj.w + P.skin;
          if (distsq < R * R) {
            if (n >= cap) {
              ovf = true;
              continue;
            }
            const int id = pid[q];
            int at = n;
            while (at > 0 && cpid[at - 1] > id) {
              cpid[at] = cpid[at - 1];
              cand[at] = cand[at - 1];
              --at;
            }
            cpid[at] = id;
            cand[at] = q;
            ++n;
          }
        }
      }
    }
  }
  for (int q = 0; q < n; ++q) B.nbr[(size_t)q * N + s] = make_int2(cand[q], cpid[q]);
  B.nnb[s] = n;
  if (ovf) atomicOr((unsigned int*)&B.ctl[CTL_ERR], GDEM_ERR_NB_OVERFLOW_BIT);
}

// Wall candidates by BVH traversal (k_update_neighborlist_wall,
// BVH.cu:493-540): leaf box test against (r+skin)^2, exact test
// dist < r + skin, ascending triangle ids.
constexpr int WL_MAX = 48;

template <class H>
__device__ __forceinline__ void walllist_one(const SimParams& P, const Bufs<H>& B, int s, int cur, int nact) {
  const int N = P.N;
  if (s >= nact || P.ntri == 0) {
    B.nwl[s] = 0;
    return;
  }
  const double4 p = B.pos[cur][s];
  const double R = p.w + P.skin;
  const double Rsq = R * R;
  const int cap = P.wall_cap < WL_MAX ? P.wall_cap : WL_MAX;
  int list[WL_MAX];
  int n = 0;
  bool ovf = false;
  // sphereAABBOverlapNeighbor (BVH.cu:542-574) then the exact test
  auto consider = [&](int t, const double* lo, const double* hi) {
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (p.x < lo[0]) ex = lo[0] - p.x;
    else if (p.x > hi[0]) ex = p.x - hi[0];
    if (p.y < lo[1]) ey = lo[1] - p.y;
    else if (p.y > hi[1]) ey = p.y - hi[1];
    if (p.z < lo[2]) ez = lo[2] - p.z;
    else if (p.z > hi[2]) ez = p.z - hi[2];
    return (ex * ex + ey * ey + ez * ez) <= Rsq;
  };
  if (P.wall_table) {
    // static per-cell table: ascending triangle ids already
    const int c = B.cell_of[s];
    const int b1 = B.wc_off[c + 1];
    for (int q = B.wc_off[c]; q < b1; ++q) {
      const int t = B.wc_tri[q];
      const double* bx = B.tbox + (size_t)t * 6;
      if (!consider(t, bx, bx + 3)) continue;
      const TriHit h = tri_distance(B.tri + t, p.x, p.y, p.z, Rsq);
      if (h.dist < p.w + P.skin) {
        if (n >= cap) {
          ovf = true;
          continue;
        }
        list[n++] = t;
      }
    }
  } else {
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const BvhNode* nd = B.bvh + stack[--sp];
      const int nchild = (nd->child[0] == nd->child[1]) ? 1 : 2;
      for (int c = 0; c < nchild; ++c) {
        if (!consider(-1, nd->lo[c], nd->hi[c])) continue;
        const int ch = nd->child[c];
        if (ch >= 0) {
          if (sp < 64) stack[sp++] = ch;
          else ovf = true;
        } else {
          const int t = ~ch;
          const TriHit h = tri_distance(B.tri + t, p.x, p.y, p.z, Rsq);
          if (h.dist < p.w + P.skin) {
            if (n >= cap) {
              ovf = true;
              continue;
            }
            int at = n;
            while (at > 0 && list[at - 1] > t) {
              list[at] = list[at - 1];
              --at;
            }
            list[at] = t;
            ++n;
          }
        }
      }
    }
  }
  for (int q = 0; q < n; ++q) B.wl[(size_t)q * N + s] = list[q];
  B.nwl[s] = n;
  if (ovf) atomicOr((unsigned int*)&B.ctl[CTL_ERR], GDEM_ERR_WALL_OVERFLOW_BIT);
}

// ---- the launch chain (one kernel per phase; used under hipGraph capture) ----
template <class H>
__global__ __launch_bounds__(256) void k_keys(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < P.N) keys_one(P, B, s, B.ctl[CTL_CUR]);
}

template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_count(int n, int chunk, int shift,
                                                         const uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ hist, const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  rs_count_block(n, chunk, shift, keys, hist, blockIdx.x, gridDim.x);
}

// Exclusive scan of the whole digit-major count table by one workgroup.
template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan(int m, uint32_t* __restrict__ hist, const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  block_exclusive_scan(hist, m);
}

template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(int n, int chunk, int shift,
                                                           const uint32_t* __restrict__ kin,
                                                           const int* __restrict__ iin,
                                                           uint32_t* __restrict__ kout,
                                                           int* __restrict__ iout,
                                                           const uint32_t* __restrict__ hist,
                                                           const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  __shared__ uint32_t base[256];
  base[threadIdx.x] = hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x];
  __syncthreads();
  rs_scatter_block(n, chunk, shift, kin, iin, kout, iout, base, blockIdx.x);
}

template <class H>
__global__ __launch_bounds__(256) void k_permute(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  if (s < P.N) permute_one(P, B, s, cur, hc);
  if (last_block(&B.ctl[CTL_DONE_PERM])) {
    B.ctl[CTL_DONE_PERM] = 0;
    atomicExch(&B.ctl[CTL_CUR], cur ^ 1);
    atomicExch(&B.ctl[CTL_HCUR], hc ^ 1);
  }
}

template <class H>
__global__ __launch_bounds__(256) void k_cells(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < P.N) cells_one(P, B, s);
}

// Pair and wall lists of the sorted particles, at full occupancy (both the
// chain and the cooperative rebuild end here).  The last workgroup clears the
// refresh flag, which ends the rebuild.
template <class H>
__global__ __launch_bounds__(256) void k_lists(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < P.N) {
    const int cur = B.ctl[CTL_CUR], hc = B.ctl[CTL_HCUR], nact = B.ctl[CTL_NACT];
    pairlist_one(P, B, s, cur, hc, nact);
    walllist_one(P, B, s, cur, nact);
  }
  if (last_block(&B.ctl[CTL_DONE_WALL])) {
    B.ctl[CTL_DONE_WALL] = 0;
    B.ctl[CTL_REBUILDS] += 1;
    atomicExch(&B.ctl[CTL_REFRESH], 0);
  }
}

// ---- the global rebuild phases in one cooperative kernel -------------------
// Keys, the four radix passes, the permutation and the cell table need grid
// barriers; they run in one cooperative kernel launched every step with a
// co-resident grid, which returns at once (before any barrier) when the
// refresh flag is clear.  Every workgroup reads the control words before the
// first barrier; workgroup 0 flips the buffers at the end.  The lists follow
// in k_lists.
template <class H>
__global__ __launch_bounds__(256) void k_rebuild_coop(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  const int N = P.N;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int stride = G * blockDim.x;
  const int first = b * blockDim.x + threadIdx.x;
  uint32_t* hist = B.rs_hist;
  uint32_t* dtot = B.rs_hist + (size_t)256 * G;
  for (int s = first; s < N; s += stride) keys_one(P, B, s, cur);
  grid.sync();
  __shared__ uint32_t base[256];
  __shared__ uint32_t dpre[256];
  for (int pass = 0; pass < 4; ++pass) {
    const int a = pass & 1;
    const int shift = 8 * pass;
    rs_count_block(N, P.rs_chunk, shift, B.key[a], hist, b, G);
    grid.sync();
    for (int d = b; d < 256; d += G) {
      const uint32_t tot = block_exclusive_scan(hist + (size_t)d * G, G);
      if (threadIdx.x == 0) dtot[d] = tot;
    }
    grid.sync();
    // global base of digit d = totals of smaller digits + this chunk's prefix
    dpre[threadIdx.x] = dtot[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int d = 0; d < 256; ++d) {
        const uint32_t c = dpre[d];
        dpre[d] = run;
        run += c;
      }
    }
    __syncthreads();
    base[threadIdx.x] = dpre[threadIdx.x] + hist[(size_t)threadIdx.x * G + b];
    __syncthreads();
    rs_scatter_block(N, P.rs_chunk, shift, B.key[a], B.idx[a], B.key[a ^ 1], B.idx[a ^ 1], base, b);
    grid.sync();
  }
  for (int s = first; s < N; s += stride) permute_one(P, B, s, cur, hc);
  grid.sync();
  for (int s = first; s < N; s += stride) cells_one(P, B, s);
  if (b == 0 && threadIdx.x == 0) {
    B.ctl[CTL_CUR] = cur ^ 1;
    B.ctl[CTL_HCUR] = hc ^ 1;
  }
}

// ---------------------------------------------------------------------------
// download / debug gathers (internal order -> original ids)
// ---------------------------------------------------------------------------
template <class H>
__global__ __launch_bounds__(256) void k_download(SimParams P, Bufs<H> B, double* x, double* v,
                                                  double* w, int* act) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.N) return;
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  const int id = B.pid[hc][s];
  const double4 p = B.pos[cur][s], vv = B.vel[cur][s], ww = B.angv[cur][s];
  x[id * 3] = p.x; x[id * 3 + 1] = p.y; x[id * 3 + 2] = p.z;
  v[id * 3] = vv.x; v[id * 3 + 1] = vv.y; v[id * 3 + 2] = vv.z;
  w[id * 3] = ww.x; w[id * 3 + 1] = ww.y; w[id * 3 + 2] = ww.z;
  act[id] = p.w < 0. ? 0 : 1;
}

template <class H>
__global__ __launch_bounds__(256) void k_debug(SimParams P, Bufs<H> B, DebugDev D) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.N) return;
  const int N = P.N;
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  const int id = B.pid[hc][s];
  const double4 p = B.pos[cur][s];
  const int cx = (int)floor((p.x - P.bmin[0]) * P.invd[0]) + 1;
  const int cy = (int)floor((p.y - P.bmin[1]) * P.invd[1]) + 1;
  const int cz = (int)floor((p.z - P.bmin[2]) * P.invd[2]) + 1;
  D.cellx[id * 3] = cx; D.cellx[id * 3 + 1] = cy; D.cellx[id * 3 + 2] = cz;
  D.cell_id[id] = (P.size[1] * cz + cy) * P.size[0] + cx;
  D.morton[id] = morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
  const int nn = B.nnb[s];
  D.nnb[id] = nn;
  for (int q = 0; q < nn; ++q) D.nb[(size_t)id * P.nb_cap + q] = B.nbr[(size_t)q * N + s].y;
  const int nw = B.nwl[s];
  D.nwall[id] = nw;
  for (int q = 0; q < nw; ++q) D.wall[(size_t)id * P.wall_cap + q] = B.wl[(size_t)q * N + s];
  const int nh = B.nh[hc][s];
  D.nhist[id] = nh;
  for (int q = 0; q < nh; ++q) {
    D.hist_key[(size_t)id * P.hist_cap + q] = B.hkey[hc][(size_t)q * N + s];
    const auto d = B.hd[hc][(size_t)q * N + s];
    D.hist_delta[((size_t)id * P.hist_cap + q) * 3 + 0] = (double)d.x;
    D.hist_delta[((size_t)id * P.hist_cap + q) * 3 + 1] = (double)d.y;
    D.hist_delta[((size_t)id * P.hist_cap + q) * 3 + 2] = (double)d.z;
  }
  const int nw2 = B.nw[hc][s];
  D.nwhist[id] = nw2;
  for (int q = 0; q < nw2; ++q) {
    D.whist_key[(size_t)id * P.whist_cap + q] = B.wkey[hc][(size_t)q * N + s];
    const auto d = B.wd[hc][(size_t)q * N + s];
    D.whist_delta[((size_t)id * P.whist_cap + q) * 3 + 0] = (double)d.x;
    D.whist_delta[((size_t)id * P.whist_cap + q) * 3 + 1] = (double)d.y;
    D.whist_delta[((size_t)id * P.whist_cap + q) * 3 + 2] = (double)d.z;
  }
}

// per-particle list/contact counts for statistics
template <class H>
__global__ __launch_bounds__(256) void k_counts(SimParams P, Bufs<H> B, unsigned long long* acc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  unsigned long long a = 0, b = 0, c = 0, d = 0;
  if (s < P.N && B.pos[cur][s].w >= 0.) {
    a = 1;
    b = (unsigned long long)B.nnb[s];
    c = (unsigned long long)B.nwl[s];
    d = (unsigned long long)(B.nh[hc][s] + B.nw[hc][s]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
    c += __shfl_down(c, off);
    d += __shfl_down(d, off);
  }
  if (__lane_id() == 0 && (a | b | c | d)) {
    atomicAdd(acc + 0, a);
    atomicAdd(acc + 1, b);
    atomicAdd(acc + 2, c);
    atomicAdd(acc + 3, d);
  }
}

// ---------------------------------------------------------------------------
// probes used by the parity tests
// ---------------------------------------------------------------------------
__global__ void k_tri_probe(const TriGeo* tri, int np, const double* pts, const double* rsq,
                            double* out, int* hit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const TriHit h = tri_distance(tri, pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2], rsq[i]);
  out[i * 4] = h.n.x; out[i * 4 + 1] = h.n.y; out[i * 4 + 2] = h.n.z; out[i * 4 + 3] = h.dist;
  hit[i] = h.hit;
}

// plain streaming copy (calibrates the HBM byte counters and the achievable
// bandwidth): 32 B per lane, grid-stride
__global__ __launch_bounds__(256) void k_copy(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

void launch_copy(const double4* a, double4* b, size_t n, hipStream_t st) {
  k_copy<<<4096, 256, 0, st>>>(a, b, n);
}

__global__ void k_fp64_probe(int n, const double* a, const double* b, double* sq, double* dv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sq[i] = sqrt(a[i]);
  dv[i] = a[i] / b[i];
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int blocks_for(int n, int t) { return (n + t - 1) / t; }

// LDS-held history keys when the capacities fit (pairs <= 16, walls <= 8).
constexpr int PCAP_REG = 16;
constexpr int WCAP_REG = 8;

template <class T, class H>
void launch_step(const SimParams& P, const Bufs<H>& B, hipStream_t st) {
  Rec<T> none{};
  const int g = blocks_for(P.N, 256);
  if (P.hist_cap <= PCAP_REG && P.whist_cap <= WCAP_REG)
    k_step<T, H, false, HistLds<H, PCAP_REG>, HistLds<H, WCAP_REG>><<<g, 256, 0, st>>>(P, B, none);
  else
    k_step<T, H, false, HistScan<H>, HistScan<H>><<<g, 256, 0, st>>>(P, B, none);
}

template <class T, class H>
void launch_step_record(const SimParams& P, const Bufs<H>& B, const RecDev& r, hipStream_t st) {
  Rec<T> rec{r.cap, r.count, r.key, r.kind, r.fn, r.ft, r.delta};
  k_step<T, H, true, HistScan<H>, HistScan<H>><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, rec);
}

void launch_radix_sort(int n, int blocks, int chunk, uint32_t* k0, int* i0, uint32_t* k1, int* i1,
                       uint32_t* hist, const int* ctl, bool gated, hipStream_t st) {
  uint32_t* kk[2] = {k0, k1};
  int* ii[2] = {i0, i1};
  for (int pass = 0; pass < 4; ++pass) {
    const int a = pass & 1, b = a ^ 1;
    if (gated) {
      k_rs_count<true><<<blocks, RS_THREADS, 0, st>>>(n, chunk, 8 * pass, kk[a], hist, ctl);
      k_rs_scan<true><<<1, RS_THREADS, 0, st>>>(256 * blocks, hist, ctl);
      k_rs_scatter<true><<<blocks, RS_THREADS, 0, st>>>(n, chunk, 8 * pass, kk[a], ii[a], kk[b], ii[b], hist, ctl);
    } else {
      k_rs_count<false><<<blocks, RS_THREADS, 0, st>>>(n, chunk, 8 * pass, kk[a], hist, ctl);
      k_rs_scan<false><<<1, RS_THREADS, 0, st>>>(256 * blocks, hist, ctl);
      k_rs_scatter<false><<<blocks, RS_THREADS, 0, st>>>(n, chunk, 8 * pass, kk[a], ii[a], kk[b], ii[b], hist, ctl);
    }
  }
}

template <class H>
int coop_grid_limit() {
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (!prop.cooperativeLaunch) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rebuild_coop<H>, 256, 0) != hipSuccess) return 0;
  return per_cu * prop.multiProcessorCount;
}

template <class H>
hipError_t launch_rebuild(const SimParams& P, const Bufs<H>& B, bool coop, hipStream_t st) {
  if (coop) {
    SimParams p = P;
    Bufs<H> b = B;
    void* args[] = {&p, &b};
    const hipError_t e =
        hipLaunchCooperativeKernel((const void*)k_rebuild_coop<H>, dim3(P.rs_blocks), dim3(256), args, 0, st);
    if (e != hipSuccess) return e;
    k_lists<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B);
    return hipGetLastError();
  }
  const int g = blocks_for(P.N, 256);
  k_keys<H><<<g, 256, 0, st>>>(P, B);
  launch_radix_sort(P.N, P.rs_blocks, P.rs_chunk, B.key[0], B.idx[0], B.key[1], B.idx[1], B.rs_hist,
                    B.ctl, true, st);
  k_permute<H><<<g, 256, 0, st>>>(P, B);
  k_cells<H><<<g, 256, 0, st>>>(P, B);
  k_lists<H><<<g, 256, 0, st>>>(P, B);
  return hipGetLastError();
}

template <class H>
void launch_download(const SimParams& P, const Bufs<H>& B, double* x, double* v, double* w, int* a,
                     hipStream_t st) {
  k_download<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, x, v, w, a);
}

template <class H>
void launch_debug(const SimParams& P, const Bufs<H>& B, const DebugDev& D, hipStream_t st) {
  k_debug<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, D);
}

template <class H>
void launch_counts(const SimParams& P, const Bufs<H>& B, unsigned long long* acc, hipStream_t st) {
  k_counts<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, acc);
}

void launch_tri_probe(const TriGeo* tri, int np, const double* pts, const double* rsq, double* out,
                      int* hit, hipStream_t st) {
  k_tri_probe<<<blocks_for(np, 256), 256, 0, st>>>(tri, np, pts, rsq, out, hit);
}

void launch_fp64_probe(int n, const double* a, const double* b, double* sq, double* dv, hipStream_t st) {
  k_fp64_probe<<<blocks_for(n, 256), 256, 0, st>>>(n, a, b, sq, dv);
}

// explicit instantiations: FP64 (T=H=double) and mixed (T=H=float)
template void launch_step<double, double>(const SimParams&, const Bufs<double>&, hipStream_t);
template void launch_step<float, float>(const SimParams&, const Bufs<float>&, hipStream_t);
template void launch_step_record<double, double>(const SimParams&, const Bufs<double>&, const RecDev&, hipStream_t);
template void launch_step_record<float, float>(const SimParams&, const Bufs<float>&, const RecDev&, hipStream_t);
template hipError_t launch_rebuild<double>(const SimParams&, const Bufs<double>&, bool, hipStream_t);
template hipError_t launch_rebuild<float>(const SimParams&, const Bufs<float>&, bool, hipStream_t);
template int coop_grid_limit<double>();
template int coop_grid_limit<float>();
template void launch_download<double>(const SimParams&, const Bufs<double>&, double*, double*, double*, int*, hipStream_t);
template void launch_download<float>(const SimParams&, const Bufs<float>&, double*, double*, double*, int*, hipStream_t);
template void launch_debug<double>(const SimParams&, const Bufs<double>&, const DebugDev&, hipStream_t);
template void launch_debug<float>(const SimParams&, const Bufs<float>&, const DebugDev&, hipStream_t);
template void launch_counts<double>(const SimParams&, const Bufs<double>&, unsigned long long*, hipStream_t);
template void launch_counts<float>(const SimParams&, const Bufs<float>&, unsigned long long*, hipStream_t);

}  // namespace gdem
