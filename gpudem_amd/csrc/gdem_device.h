// Device-side building blocks shared by the single-scene and the batched
// kernels: vector helpers, streaming loads, the sphere-triangle distance,
// contact-history rows, the per-particle contact step and the list builders.
//
// FP64 arithmetic follows the reference operation for operation (the build
// uses -ffp-contract=off): pair contact d_particle_collision_verlet +
// d_calc_normal_force + d_calc_tangential_force (device_dem.cu:916-981,
// 793-847, 622-740), wall contact d_wall_collision_verlet +
// d_calc_normal_force_wall + d_calc_tangential_force_wall + d_dist_triangle
// (device_dem.cu:288-499, 13-205), history update d_update_history(_wall)
// (device_dem.cu:743-790), integration k_integrate (device_dem.cu:1065-1091),
// dk_checkOoB (:1116-1125) and k_shouldRefreshNeighborList (:1095-1112).
#pragma once
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
__device__ __forceinline__ int2 ld_nt(const int2* p) {
  typedef int v2 __attribute__((ext_vector_type(2)));
  const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return make_int2(v[0], v[1]);
}
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v[0], v[1], v[2], v[3]);
}
template <class H>
__device__ __forceinline__ Del3<H> ld_nt(const Del3<H>* p) {
  Del3<H> r;
  r.x = __builtin_nontemporal_load(&p->x);
  r.y = __builtin_nontemporal_load(&p->y);
  r.z = __builtin_nontemporal_load(&p->z);
  return r;
}
template <class H>
__device__ __forceinline__ void st_nt(Del3<H> a, Del3<H>* p) {
  __builtin_nontemporal_store(a.x, &p->x);
  __builtin_nontemporal_store(a.y, &p->y);
  __builtin_nontemporal_store(a.z, &p->z);
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
  __syncthreads();  // every thread of the workgroup is past its reads
  if (threadIdx.x != 0) return false;
  const int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return prev == (int)gridDim.x - 1;
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

template <class G>
__device__ __forceinline__ TriHit tri_distance_g(const G& tg, double px, double py, double pz, double Rsq) {
  TriHit res;
  res.n = V3<double>{0., 0., 0.};
  res.dist = 1.e10;
  res.hit = -1;
  const V3<double> nt{tg.n[0], tg.n[1], tg.n[2]};
  double dist = 0.;
  dist += nt.x * px;
  dist += nt.y * py;
  dist += nt.z * pz;
  dist += tg.d;
  double distsq = dist * dist;
  if (distsq > Rsq) return res;
  const V3<double> pos{px, py, pz};
  const V3<double> P0{tg.p0[0], tg.p0[1], tg.p0[2]};
  const V3<double> ap = vsub(pos, P0);
  const double d00 = tg.d00, d01 = tg.d01, d11 = tg.d11, den = tg.denom;
  const double d20 = ap.x * tg.e01[0] + ap.y * tg.e01[1] + ap.z * tg.e01[2];
  const double d21 = ap.x * tg.e02[0] + ap.y * tg.e02[1] + ap.z * tg.e02[2];
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
    hit = tg.v0;
  } else if (u <= 0. && w <= 0.0) {
    q = V3<double>{tg.p1[0], tg.p1[1], tg.p1[2]};
    hit = tg.v1;
  } else if (u <= 0. && v <= 0.0) {
    q = V3<double>{tg.p2[0], tg.p2[1], tg.p2[2]};
    hit = tg.v2;
  } else if (u <= 0.0) {
    const V3<double> P1{tg.p1[0], tg.p1[1], tg.p1[2]};
    const V3<double> E{tg.e12[0], tg.e12[1], tg.e12[2]};
    double t = dot(vsub(pos, P1), E) * tg.d22inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P1, scal(t, E));
    hit = tg.e1;
  } else if (v <= 0.0) {
    const V3<double> E{tg.e02[0], tg.e02[1], tg.e02[2]};
    double t = d21 * tg.d11inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P0, scal(t, E));
    hit = tg.e2;
  } else {
    const V3<double> E{tg.e01[0], tg.e01[1], tg.e01[2]};
    double t = d20 * tg.d00inv;
    t = fmax(0., fmin(1.0, t));
    q = vadd(P0, scal(t, E));
    hit = tg.e0;
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

// Reads the triangle's fields as the tests need them (the plane test first).
__device__ __forceinline__ TriHit tri_distance(const TriGeo* __restrict__ tg, double px, double py,
                                               double pz, double Rsq) {
  return tri_distance_g(*tg, px, py, pz, Rsq);
}

// The same test in two memory round trips, for the contact pass.  Reading the
// fields as tri_distance_g needs them costs up to four dependent trips
// (plane; barycentric terms; region data; feature id).  Here the first trip
// (tri_first) requests the plane, the first vertex, both edges from it and
// the barycentric terms together; a point outside the face region then
// requests its region's data in one second trip: the base point, edge, inverse squared length and feature
// id, at offsets selected per lane.  The arithmetic is tri_distance_g's,
// expression for expression: the edge parameter dot(pos - P, E) * inv equals
// its d20 * d00inv and d21 * d11inv for the edges from the first vertex.
// The first trip's fields.
struct TriFirst {
  V3<double> nt, P0, E01, E02;
  double pd, d00, d01, d11, den;
};
__device__ __forceinline__ TriFirst tri_first(const TriGeo* __restrict__ g) {
  TriFirst f;
  f.nt = V3<double>{g->n[0], g->n[1], g->n[2]};
  f.pd = g->d;
  f.P0 = V3<double>{g->p0[0], g->p0[1], g->p0[2]};
  f.E01 = V3<double>{g->e01[0], g->e01[1], g->e01[2]};
  f.E02 = V3<double>{g->e02[0], g->e02[1], g->e02[2]};
  f.d00 = g->d00, f.d01 = g->d01, f.d11 = g->d11, f.den = g->denom;
  return f;
}
// The test given the first trip's fields (tri_first, issued by the caller
// before work that does not need them).
__device__ __forceinline__ TriHit tri_distance_2trip(const TriFirst& f, const TriGeo* __restrict__ g, double px,
                                                     double py, double pz, double Rsq) {
  TriHit res;
  res.n = V3<double>{0., 0., 0.};
  res.dist = 1.e10;
  res.hit = -1;
  const V3<double> nt = f.nt, P0 = f.P0, E01 = f.E01, E02 = f.E02;
  const double pd = f.pd, d00 = f.d00, d01 = f.d01, d11 = f.d11, den = f.den;
  double dist = 0.;
  dist += nt.x * px;
  dist += nt.y * py;
  dist += nt.z * pz;
  dist += pd;
  double distsq = dist * dist;
  if (distsq > Rsq) return res;
  const V3<double> pos{px, py, pz};
  const V3<double> ap = vsub(pos, P0);
  const double d20 = ap.x * E01.x + ap.y * E01.y + ap.z * E01.z;
  const double d21 = ap.x * E02.x + ap.y * E02.y + ap.z * E02.z;
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
  // region: base point, edge, inverse squared edge length, feature id
  int op, oe, oi, oh;
  bool edge = true;
  if (v <= 0. && w <= 0.0) {
    op = offsetof(TriGeo, p0), oh = offsetof(TriGeo, v0), edge = false;
  } else if (u <= 0. && w <= 0.0) {
    op = offsetof(TriGeo, p1), oh = offsetof(TriGeo, v1), edge = false;
  } else if (u <= 0. && v <= 0.0) {
    op = offsetof(TriGeo, p2), oh = offsetof(TriGeo, v2), edge = false;
  } else if (u <= 0.0) {
    op = offsetof(TriGeo, p1), oh = offsetof(TriGeo, e1);
  } else if (v <= 0.0) {
    op = offsetof(TriGeo, p0), oh = offsetof(TriGeo, e2);
  } else {
    op = offsetof(TriGeo, p0), oh = offsetof(TriGeo, e0);
  }
  if (u <= 0.0) {
    oe = offsetof(TriGeo, e12), oi = offsetof(TriGeo, d22inv);
  } else if (v <= 0.0) {
    oe = offsetof(TriGeo, e02), oi = offsetof(TriGeo, d11inv);
  } else {
    oe = offsetof(TriGeo, e01), oi = offsetof(TriGeo, d00inv);
  }
  const char* b = reinterpret_cast<const char*>(g);
  const double* Pp = reinterpret_cast<const double*>(b + op);
  const double* Ep = reinterpret_cast<const double*>(b + oe);
  const V3<double> P{Pp[0], Pp[1], Pp[2]};
  const V3<double> E{Ep[0], Ep[1], Ep[2]};
  const double inv = *reinterpret_cast<const double*>(b + oi);
  const int hit = *reinterpret_cast<const int*>(b + oh);
  // (evaluated for a vertex too, and discarded, so that every load of the
  // trip is issued before the first wait)
  double t = dot(vsub(pos, P), E) * inv;
  t = fmax(0., fmin(1.0, t));
  const V3<double> qe = vadd(P, scal(t, E));
  const V3<double> q = edge ? qe : P;
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
// scene's capacity): a wave's slot row is 64 consecutive words in memory and
// in LDS, so the keys are loaded straight into LDS (global_load_lds) without
// occupying registers, while the screening pass runs; rows_ready() waits for
// them.  The per-contact search then costs no memory round trip.  HistScan
// searches the keys in memory and serves capacities beyond the LDS budget.
template <class H, int CAP>
struct HistLds {
  int* gkey;
  Del3<H>* gd;
  int s, N, cap;
  int n;           // live slots
  uint64_t marks;  // slots in contact this step
  int* k0;         // thread 0's column; thread t's column is k0 + t, slot q at [q * 256]
  // ok: this lane's particle is active.  The wave loads as many slot rows as
  // its longest row (never more than the capacity); in each, only the lanes
  // whose own row reaches it fetch their key (issue).  Slots past a lane's
  // own row hold zeros and are never searched.
  // prepare returns the number of slot rows the wave loads; issue loads them
  // (split so that the wave shuffles of several rows come before the first
  // LDS load: the compiler makes any LDS access after such a load wait for it)
  __device__ __forceinline__ int prepare(int nh, bool ok) {
    n = nh;
    marks = 0;
    int wmax = ok ? nh : 0;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));
    return min(__builtin_amdgcn_readfirstlane(wmax), min(cap, CAP));
  }
  // Buffer loads: the [cap][N] key array is the buffer, each lane's slot is
  // its one 32-bit vector offset and each row's offset is scalar, so the rows
  // cost no vector address registers (the host selects this row type only
  // when cap * N * 4 bytes fit the 32-bit buffer range).  A lane past its own
  // row length gives an offset beyond the buffer: the load then fetches
  // nothing and writes 0 to the (never searched) LDS word.
  __device__ __forceinline__ void issue(int wmax, bool ok) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)gkey, (short)0, (int)((uint32_t)cap * (uint32_t)N * 4u), 0x00020000);
    const uint32_t lane_off = (uint32_t)s * 4u;
    const int mine = ok ? n : 0;
    int* dst = k0 + (threadIdx.x & ~63);  // this wave's part of slot row 0
    for (int q = 0; q < wmax; ++q)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(dst + q * 256), 4,
                                               q < mine ? lane_off : 0xfffffff0u, (uint32_t)q * (uint32_t)N * 4u, 0,
                                               0);
  }
  __device__ __forceinline__ void load(int nh, bool ok) { issue(prepare(nh, ok), ok); }
  // slot of `key` among the first n_o slots of the particle held by thread
  // col (slot index so)
  // (keys of a row are distinct; the slots are read four at a time, so a
  // search costs one LDS round trip per four slots, not one per slot)
  __device__ __forceinline__ int find_in(int col, int so, int n_o, int key) const {
    const int* c = k0 + col;
    for (int q0 = 0; q0 < n_o; q0 += 4) {
      int hit = -1;
#pragma unroll
      for (int u = 3; u >= 0; --u) {
        const int q = q0 + u;
        const int kq = c[(q < CAP ? q : CAP - 1) * 256];
        if (q < n_o && kq == key) hit = q;
      }
      if (hit >= 0) return hit;
    }
    return -1;
  }
};

// Whether HistLds can serve a scene: the capacities fit its LDS columns and
// every key row offset fits the 32-bit range of its buffer loads.
__host__ __device__ inline bool hist_lds_fits(int N, int hist_cap, int whist_cap, int pcap_lds, int wcap_lds) {
  const unsigned long long cap = (unsigned long long)(hist_cap > whist_cap ? hist_cap : whist_cap);
  return hist_cap <= pcap_lds && whist_cap <= wcap_lds && cap * (unsigned long long)N * 4ull < 0xfffffff0ull;
}

template <class H>
struct HistScan {
  int* gkey;
  Del3<H>* gd;
  int s, N, cap;
  int n;
  uint64_t marks;
  __device__ __forceinline__ void load(int nh, bool) {
    n = nh;
    marks = 0;
  }
  __device__ __forceinline__ int find_in(int, int so, int n_o, int key) const {
    for (int q = 0; q < n_o; ++q)
      if (gkey[(size_t)q * N + so] == key) return q;
    return -1;
  }
};

// Waits until the rows' keys are in LDS (the wave's own LDS loads have
// landed: s_waitcnt vmcnt(0)) before any lane searches another lane's column.
template <class Row>
__device__ __forceinline__ void rows_ready(Row&) {}
template <class H, int CAP>
__device__ __forceinline__ void rows_ready(HistLds<H, CAP>&) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt, lgkmcnt not waited for
  __builtin_amdgcn_wave_barrier();
}

// Tangential spring with Coulomb cap and history (d_calc_tangential_force,
// device_dem.cu:622-740; the wall twin :403-499 is identical with the wall
// history).  The slot of `key` is found or appended (d_calc_tangential_force
// :630-653); T is the force type and FP64 keeps the reference's operations.
// The spring-dashpot arithmetic of d_calc_tangential_force
// (device_dem.cu:654-737): the displacement `old` advanced by vt dt and
// projected onto the tangent plane, the Coulomb-capped spring force, and the
// displacement to remember (returned in dn).
// Reciprocal and square root of the contact force arithmetic: the
// reference's correctly rounded operations in FP64, the hardware's one-ulp
// instructions (v_rcp_f32, v_sqrt_f32) in the mixed form, whose forces carry
// a relative error of order 1e-7 either way.
template <class T>
__device__ __forceinline__ T frcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return T(1) / x;
}
template <class T>
__device__ __forceinline__ T fsqrt(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrtf(x);
  else return sqrt(x);
}

template <class T>
__device__ __forceinline__ V3<T> tangential_force(V3<T> old, V3<T> fn, V3<T> vt, V3<T> n, T eta, T k, T mu, T dt,
                                                  V3<T>& dn_out) {
  V3<T> dn{old.x + vt.x * dt, old.y + vt.y * dt, old.z + vt.z * dt};
  const T dd = dot(dn, n);
  dn = vsub(dn, scal(dd, n));
  V3<T> ft = scal(-k, dn);
  const T ftsq = dot(ft, ft);
  const T fnsq = dot(fn, fn);
  if (ftsq > (mu * mu) * fnsq) {
    if (ftsq != T(0)) {
      const V3<T> t = scal(frcp(fsqrt(ftsq) + T(1e-15)), ft);
      const T fnn = fsqrt(fnsq);
      ft = scal(mu * fnn, t);
      if constexpr (sizeof(T) == 4) dn = scal(-frcp(k), ft);
      else dn = scal(T(-1) / k, ft);
    } else {
      ft = V3<T>{T(0), T(0), T(0)};
    }
  } else {
    ft.x -= eta * vt.x;
    ft.y -= eta * vt.y;
    ft.z -= eta * vt.z;
  }
  dn_out = dn;
  return ft;
}

template <class H, class T>
__device__ __forceinline__ void store_delta(Del3<H>* p, V3<T> dn) {
  Del3<H> o;
  o.x = (H)dn.x;
  o.y = (H)dn.y;
  o.z = (H)dn.z;
  st_nt(o, p);
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
  double* acc;   // [n*3] or null: linear acceleration of the step
  double* anga;  // [n*3] or null: angular acceleration of the step
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

// PRow/WRow: history row types for pairs and walls.  Pair candidates are
// screened first (indices and positions fetched screen_batch() at a time so
// their latencies overlap, FP64 overlap test), then the touching ones are
// evaluated in list order.
// Occupancy bound of the step kernels (waves per SIMD), per precision: 4 for
// the mixed form (at most 128 VGPRs, which its layout fits without spills,
// and 40 KB of LDS per workgroup), 3 for the FP64 form, whose wider
// registers would spill at 4.
#ifndef GDEM_FP64_WAVES
#define GDEM_FP64_WAVES 3
#endif
#ifndef GDEM_MIXED_WAVES
#define GDEM_MIXED_WAVES 4
#endif
template <class T>
constexpr int step_waves() {
  return sizeof(T) == 8 ? GDEM_FP64_WAVES : GDEM_MIXED_WAVES;
}

// The FP64 form stages the workgroup's velocity records in LDS with the
// positions, so a touching partner inside the tile is read from LDS rather
// than gathered from memory (where its line has usually left the L2 since the
// workgroup read it as an own record); it gathers spins from memory, as its
// double records leave no LDS room for both at three workgroups per CU.  The
// mixed form runs at four workgroups per CU instead (40 KB of LDS each, the
// positions, history keys and stash), which leaves no room for a velocity
// tile: it gathers partners' velocities and spins from memory.  The fourth
// wave per SIMD is worth more than the tile (mixed force kernel on the chute:
// 112.1 us at three waves with a float velocity and spin tile, 108.3 us at
// four without).
template <class T>
constexpr bool v_tile() {
  return sizeof(T) == 8;
}

// Touching pairs of a particle whose list entries (partner slots) the
// screening pass keeps in LDS for the force pass (later ones are read again
// from the list): 8, so that a workgroup takes 48 KB of LDS in the FP64 form
// (three fit a CU) and 40 KB in the mixed form (four fit).
template <class T>
constexpr int hit_stash() {
  return 8;
}

// The hit stash is used only in the pair pass and the vertex/edge features
// only in the wall pass, so they share their LDS: a thread's ve_max<T>()
// feature words are its own stash words (ve_slot), so no thread's features
// overlap another thread's stash, whatever pass the other waves of the
// workgroup are in.
template <class T>
constexpr int ve_max() {
  return hit_stash<T>();
}

template <class T>
struct VwTile {};
template <>
struct VwTile<double> {
  double4 tv[256];  // velocity records of the tile's slots
};

template <class T, int PC, int WC>
struct StepLds : VwTile<T> {
  int pkey[PC > 0 ? PC : 1][256];
  int wkey[WC > 0 ? WC : 1][256];
  int hit[hit_stash<T>()][256];
};

// Stages a thread's own velocity in the FP64 form's tile; t is its slot's
// place in the tile.
template <class T, class Lds>
__device__ __forceinline__ void stage_vw(Lds& lds, int t, const double4& v) {
  if constexpr (v_tile<T>()) lds.tv[t] = v;
}

// feature word u (< ve_max<T>()) of thread tx
template <class Lds>
__device__ __forceinline__ int* ve_slot(Lds& lds, int tx, int u) {
  return &lds.hit[u][tx];
}

template <class Row>
__device__ __forceinline__ void bind_lds(Row&, int*, int) {}
template <class H, int CAP>
__device__ __forceinline__ void bind_lds(HistLds<H, CAP>& r, int* col0, int) {
  r.k0 = col0;
}

// Starts loading the pair and wall history rows of a particle (rows_ready()
// before the keys are searched).
template <class A, class B>
__device__ __forceinline__ void load_rows(A& a, int na, B& b, int nb, bool ok) {
  a.load(na, ok);
  b.load(nb, ok);
}
template <class H, int CA, int CB>
__device__ __forceinline__ void load_rows(HistLds<H, CA>& a, int na, HistLds<H, CB>& b, int nb, bool ok) {
  const int wa = a.prepare(na, ok), wb = b.prepare(nb, ok);
  a.issue(wa, ok);
  b.issue(wb, ok);
}

template <class Row>
struct RowCap {
  static constexpr int value = 0;
};
template <class H, int CAP>
struct RowCap<HistLds<H, CAP>> {
  static constexpr int value = CAP;
};

// Everything the contact step of one particle reads and writes: the current
// and next state buffers, material, histories and candidate lists (all indexed
// by the particle's slot s, lists and histories slot-major with stride N), and
// the scene constants.
template <class H>
struct StepArgs {
  using H4 = typename Vec4<H>::type;
  const double4* pin;
  const double4* vin;
  const double4* win;
  double4* pout;
  double4* vout;
  double4* wout;
  const H4* mat;
  int* hkey;
  Del3<H>* hd;
  int* nh;
  int* wkey;
  Del3<H>* wd;
  int* nw;
  const double4* ref;
  const int* pid;  // original ids (read by the recorder only)
  const int* nbr;
  const int* nnb;
  const int* wl;
  const int* nwl;
  const TriGeo* tri;
  int N;
  int nb_cap, hist_cap, whist_cap;
  double dt, mu;
  double g[3], bmin[3], bmax[3];
  double threshSq;
  double qh[3], qeps;  // packed reference position: step per axis, error bound (refq_set)
};

// Pair-list entries screened per batch, and (PRE) fetched with the own
// records: 8 in the FP64 form, 4 in the mixed form, whose register budget at
// four waves per SIMD holds four candidate positions in flight, not eight.
template <class H>
constexpr int screen_batch() {
  return sizeof(H) == 4 ? 4 : 8;
}
template <class H>
constexpr int pre_rows() {
  return 2 * screen_batch<H>();
}

// A particle's own records: state, material, list and history lengths, and
// (PRE) the first pre_rows() pair-list entries (two screening batches).  All
// loads are independent, so they share one memory round trip; lanes without
// a particle load nothing.  The entries are fetched before the list length is
// known: rows past the length (always inside the [nb_cap][N] array) hold
// stale entries, which the screening pass masks.  This saves the round trips
// that wait for the length and for the second batch's entries, at the price
// of reading pre_rows() 4-byte rows whatever the length: a gain for dense
// packings (lists of about eight and longer), a loss for dilute ones.
template <class H>
struct Own {
  double4 p, v, w;  // w.w: the packed reference position (refq_pack)
  typename Vec4<H>::type mt;
  int nnb, nwl, nh0, nw0;
  int nb0[pre_rows<H>()];
};

template <bool PRE, class H>
__device__ __forceinline__ Own<H> load_own(const StepArgs<H>& A, int s, bool valid) {
  Own<H> o;
  o.p = make_double4(0., 0., 0., -1.);
  o.v = make_double4(0., 0., 0., 0.);
  o.w = o.v;
  o.mt = {};
  o.nnb = o.nwl = o.nh0 = o.nw0 = 0;
#pragma unroll
  for (int u = 0; u < pre_rows<H>(); ++u) o.nb0[u] = -1;
  if (valid) {
    o.p = A.pin[s];
    o.v = A.vin[s];
    o.w = A.win[s];
    o.mt = ld_nt(A.mat + s);
    o.nnb = ld_nt(A.nnb + s);
    o.nwl = ld_nt(A.nwl + s);
    o.nh0 = ld_nt(A.nh + s);
    o.nw0 = ld_nt(A.nw + s);
    // (rows clamped rather than skipped: a branch here would leave the
    // compiler unable to count the loads in flight, and it would then wait
    // for all of them before the next dependent access)
    if constexpr (PRE) {
#pragma unroll
      for (int u = 0; u < pre_rows<H>(); ++u)
        o.nb0[u] = A.nbr[(size_t)(u < A.nb_cap ? u : A.nb_cap - 1) * A.N + s];
    }
  }
  return o;
}

// The position at the last list build travels packed in the unused fourth
// word of the particle's spin record, which the step reads and writes anyway:
// 21 bits per axis, a fixed-point offset from the box corner bmin in steps of
// h = (bmax - bmin) / (2^21 - 1).  A position outside the box (never that of
// an active particle) is stored as REFQ_EXACT instead.  The exact position
// stays in its own array for the rare test that needs it.
constexpr int REFQ_BITS = 21;
constexpr unsigned long long REFQ_MASK = (1ull << REFQ_BITS) - 1;
constexpr double REFQ_LEVELS = 2097151.0;  // 2^21 - 1
constexpr unsigned long long REFQ_EXACT = ~0ull;

// Step per axis and the bound on |ref - decoded ref|: each component is
// rounded to the nearest step, so it is off by at most h/2, plus the
// rounding of the double arithmetic that forms and decodes it (a few ulps
// of |bmin| + |bmax|).
__device__ __forceinline__ void refq_set(const double lo[3], const double hi[3], double h[3], double& eps) {
  double hh = 0., big = 0.;
  for (int a = 0; a < 3; ++a) {
    h[a] = (hi[a] - lo[a]) * (1.0 / REFQ_LEVELS);
    hh += h[a] * h[a];
    big += fabs(lo[a]) + fabs(hi[a]);
  }
  eps = 0.5 * (1.0 + 1e-8) * sqrt(hh) + 4e-15 * big;
}

__device__ __forceinline__ double refq_pack(const double4& p, const double lo[3], const double hi[3],
                                            const double h[3]) {
  const double c[3] = {p.x, p.y, p.z};
  unsigned long long u = 0;
  for (int a = 0; a < 3; ++a) {
    if (!(c[a] >= lo[a] && c[a] <= hi[a])) return __longlong_as_double((long long)REFQ_EXACT);
    const double f = fmin(fmax(rint((c[a] - lo[a]) / h[a]), 0.), REFQ_LEVELS);
    u |= (unsigned long long)f << (REFQ_BITS * a);
  }
  return __longlong_as_double((long long)u);
}

// The reference's refresh test (main.cu, cpu_dem.cu:1529-1552):
// |x - ref|^2 > thresh in FP64, with ref the position at the last list
// build.  It is first decided on the packed position q: |ref - q| <= eps, so
// the distance from q is within eps of the exact distance.  Outside a band of
// twice that width around the threshold distance (plus a margin for the
// double rounding of the distance itself) the two tests agree; in it
// (rarely, a fraction of a micrometre for a box of half a metre) the exact
// position is read and the reference's expression is evaluated as is.
template <class H>
__device__ __forceinline__ bool refresh_test(const double4& po, double tag, const double4* ref,
                                             const StepArgs<H>& A) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(tag);
  if (u != REFQ_EXACT) {
    const double qx = fma((double)(unsigned)(u & REFQ_MASK), A.qh[0], A.bmin[0]);
    const double qy = fma((double)(unsigned)((u >> REFQ_BITS) & REFQ_MASK), A.qh[1], A.bmin[1]);
    const double qz = fma((double)(unsigned)((u >> (2 * REFQ_BITS)) & REFQ_MASK), A.qh[2], A.bmin[2]);
    const double ex = po.x - qx, ey = po.y - qy, ez = po.z - qz;
    const double d = sqrt(ex * ex + ey * ey + ez * ez);
    const double t = sqrt(A.threshSq);
    const double m = A.qeps + 1e-12 * (d + t) + 1e-38;
    if (d > t + m) return true;
    if (d < t - m) return false;
  }
  const double4 rf = ld_nt(ref);
  const double fx = po.x - rf.x, fy = po.y - rf.y, fz = po.z - rf.z;
  return (fx * fx + fy * fy + fz * fz) > A.threshSq;
}

// ---- wave-level helpers (64 lanes) ------------------------------------------
__device__ __forceinline__ int wave_incl_sum(int v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

// The same for 0 <= v < 128, bit-sliced over ballots and lane-mask counts:
// VALU and scalar work only, where the shuffle form waits on six dependent
// cross-lane (LDS-unit) permutes.
__device__ __forceinline__ int wave_incl_sum7(int v) {
  int sum = 0;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    const int bit = (v >> b) & 1;
    const uint64_t m = __ballot(bit);
    const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    sum += (below + bit) << b;
  }
  return sum;
}

// Maximum over the wave of 0 <= v < 128 (wave-uniform), bit by bit from the
// top with ballots.
__device__ __forceinline__ int wave_max7(int v) {
  int m = 0;
#pragma unroll
  for (int b = 6; b >= 0; --b) {
    const int c = m | (1 << b);
    if (__ballot(v >= c) != 0) m = c;
  }
  return m;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v);
}

// Lane owning work item g: the first lane whose inclusive count exceeds g
// (incl is nondecreasing over the lanes; the result is 0..63 for any g, 63
// when no lane's count exceeds g).  A 4-way search: three levels of three
// independent shuffles, instead of six dependent ones.
__device__ __forceinline__ int owner_lane(int incl, int g) {
  int lo = 0;
#pragma unroll
  for (int span = 16; span > 0; span >>= 2) {
    const int v1 = __shfl(incl, lo + span - 1);
    const int v2 = __shfl(incl, lo + 2 * span - 1);
    const int v3 = __shfl(incl, lo + 3 * span - 1);
    lo += ((v1 <= g) + (v2 <= g) + (v3 <= g)) * span;
  }
  return lo;
}

// Position of the k-th (from 0) set bit of m.
__device__ __forceinline__ int nth_bit(uint64_t m, int k) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
    const uint64_t low = m & ((1ull << w) - 1);
    const int c = __popcll(low);
    if (k >= c) {
      k -= c;
      m >>= w;
      pos += w;
    } else {
      m = low;
    }
  }
  return pos;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  const unsigned lo = __shfl((unsigned)x, src), hi = __shfl((unsigned)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// lanes [lo, hi) as a mask
__device__ __forceinline__ uint64_t lane_range(int lo, int hi) {
  if (hi <= lo) return 0;
  const uint64_t upto = hi >= 64 ? ~0ull : ((1ull << hi) - 1);
  const uint64_t below = lo <= 0 ? 0ull : ((1ull << lo) - 1);
  return upto & ~below;
}

// Contact forces, history update, integration, out-of-bounds and refresh test
// of the particle in slot s, one thread per particle; every lane of the wave
// calls it (valid: the lane holds a particle).  PRow/WRow are the history row
// types for pairs and walls; PRE: own holds the first batch of list entries
// (load_own<true>), REC: every contact is also written to rec.
//
// Pair candidates are screened per particle (indices and positions fetched
// eight at a time, FP64 overlap test).  The wave's touching pairs are then
// spread over all 64 lanes, 64 at a time: each lane evaluates one pair of
// some particle (gathers, history, contact arithmetic), and every particle
// adds its pairs' forces in list order, so the sums are the ones a thread
// walking its own list forms.  Wall candidates are distance-tested the same
// way, then each particle handles its wall contacts in list order (vertex and
// edge deduplication needs that order).
//
// own holds the particle's records (load_own, issued by the caller before it
// stages own.p in tile_p[s - tile0]); tile_p holds the position records of
// slots [tile0, tile0 + 256), this workgroup's particles, so neighbours inside
// the tile (most of them, in Morton order) are read from LDS instead of
// memory.  The lanes of a wave hold consecutive slots (lane i: the wave's
// lowest slot + i).
template <class T, class H, bool REC, bool PRE, class PRow, class WRow, class Lds>
__device__ __forceinline__ void step_particle(const StepArgs<H>& A, int s, bool valid, const Own<H>& own,
                                              const double4* tile_p, int tile0, Lds& lds, const Rec<T>& rec,
                                              uint32_t& err, bool& refresh) {
  using H4 = typename Vec4<H>::type;
  constexpr int HS = hit_stash<T>();
  constexpr int SB = screen_batch<H>(), PR = pre_rows<H>();
  constexpr int VEM = ve_max<T>();
  constexpr bool VT = v_tile<T>();
  const int tx = threadIdx.x;
  const int lane = __lane_id();
  const int wtx = tx - lane;  // thread index of lane 0 of this wave
  // lane 0's slot: the lowest of the wave, so a particle's whenever any lane
  // of the wave holds one; lane i holds slot s_lane0 + i
  const int s_lane0 = __shfl(s, 0);
  const int N = A.N;
  const double4* __restrict__ pin = A.pin;
  const double4* __restrict__ vin = A.vin;
  const double4* __restrict__ win = A.win;
  // The LDS read is marked non-temporal (no effect on an LDS access) so the
  // compiler cannot fold it and the memory read into one flat load of a
  // selected address: flat loads go through the vector-memory path even for
  // LDS addresses, and the memory read keeps its normal cache policy.
  auto pos_of = [&](int j) -> double4 {
    const unsigned t = (unsigned)(j - tile0);
    double4 r = ld_nt(tile_p + (t & 255u));
    if (t >= 256u) r = pin[j];
    return r;
  };
  const double4 p = own.p, v = own.v, w = own.w;
  const bool act = valid && p.w >= 0.;
  if (valid && !act) {
    st_nt(p, A.pout + s);
    st_nt(v, A.vout + s);
    st_nt(w, A.wout + s);
  }
  const H4 mt = own.mt;
  const int nnb = act ? own.nnb : 0, nwl = act ? own.nwl : 0;
  const int nh0 = act ? own.nh0 : 0, nw0 = act ? own.nw0 : 0;
  if constexpr (PRE) {
    // The prefetched list entries arrive with the records above and are
    // waited for here (the empty asm uses them): left to the compiler, the
    // wait for the last entry would come after the wall-list and candidate
    // position loads below, and would wait for those too.
#pragma unroll
    for (int u = 0; u < PR; ++u) asm volatile("" ::"v"(own.nb0[u]));
  }
  const T k = (T)mt.x, etac = (T)mt.y, sqrtm = (T)mt.z;
  const double invmoi = (double)mt.w;
  const double ri = p.w, invm = v.w;
  const T riT = (T)ri;
  const T mu = (T)A.mu, dtT = (T)A.dt;
  const V3<T> vi{(T)v.x, (T)v.y, (T)v.z};
  const V3<T> wi{(T)w.x, (T)w.y, (T)w.z};
  V3<T> f{T(0), T(0), T(0)}, mom{T(0), T(0), T(0)};
  int nrec = 0;
  int self_id = 0;
  if constexpr (REC) self_id = A.pid[s];

  // The wave's wall candidates are spread over its lanes, 64 at a time; the
  // first 64 are fetched now so the load overlaps the pair pass.
  const int w_incl = wave_incl_sum7(nwl);
  const int w_total = __builtin_amdgcn_readfirstlane(__shfl(w_incl, 63));
  int w_t0 = 0;
  {
    // shuffles run with every lane active
    const int o = owner_lane(w_incl, lane);
    const int o_first = __shfl(w_incl - nwl, o), so = s_lane0 + o;
    if (lane < w_total) w_t0 = ld_nt(A.wl + (size_t)(lane - o_first) * N + so);
  }

  // (PRE) The first batch of pair candidates came with the particle's own
  // records; their positions are requested now, before the history keys
  // below, because any LDS access after those loads (these tile reads
  // included) waits for them.
  int ev0[SB];
  double4 cv0[SB];
  if constexpr (PRE) {
#pragma unroll
    for (int u = 0; u < SB; ++u) ev0[u] = u < nnb ? own.nb0[u] : -1;
#pragma unroll
    for (int u = 0; u < SB; ++u) cv0[u] = ev0[u] >= 0 ? pos_of(ev0[u]) : make_double4(0., 0., 0., -1.);
  }

  // history rows of pairs and walls
  PRow prow;
  bind_lds(prow, &lds.pkey[0][0], tx);
  prow.gkey = A.hkey;
  prow.gd = A.hd;
  prow.s = s;
  prow.N = N;
  prow.cap = A.hist_cap;
  WRow wrow;
  bind_lds(wrow, &lds.wkey[0][0], tx);
  wrow.gkey = A.wkey;
  wrow.gd = A.wd;
  wrow.s = s;
  wrow.N = N;
  wrow.cap = A.whist_cap;
  // (after the wave shuffles and tile reads above: an LDS access after the
  // loads below waits for them, and the first one is the screening pass's,
  // which waits for its own candidate loads anyway)
  load_rows(prow, nh0, wrow, nw0, act);
  // ---- particle pairs -------------------------------------------------------
  {
    PRow& row = prow;
    // screening: which candidates overlap (FP64 test of the reference)
    // The first HS touching entries go to this thread's LDS column.
    uint64_t hits = 0;
    int nhit = 0;
    int* stash = &lds.hit[0][tx];
    auto test = [&](const int q0, const int* ev, const double4* cv) {
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        const double4 pj = cv[u];
        if (pj.w < 0.) continue;
        const double dx = p.x - pj.x, dy = p.y - pj.y, dz = p.z - pj.z;
        const double distsq = dx * dx + dy * dy + dz * dz;
        const double R = ri + pj.w;
        if (distsq < R * R) {
          hits |= (uint64_t)1 << (q0 + u);
          if (nhit < HS) stash[nhit * 256] = ev[u];
          ++nhit;
        }
      }
    };
    int q_start = 0;
    if constexpr (PRE) {
      // (outside the loop, so that no wait for a later batch's loads precedes
      // it; likewise the second batch, whose entries are in registers)
      test(0, ev0, cv0);
      int ev1[SB];
      double4 cv1[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) ev1[u] = SB + u < nnb ? own.nb0[SB + u] : -1;
#pragma unroll
      for (int u = 0; u < SB; ++u) cv1[u] = ev1[u] >= 0 ? pos_of(ev1[u]) : make_double4(0., 0., 0., -1.);
      test(SB, ev1, cv1);
      q_start = PR;
    }
    for (int q0 = q_start; q0 < nnb; q0 += SB) {
      int ev[SB];
      double4 cv[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) ev[u] = (q0 + u < nnb) ? A.nbr[(size_t)(q0 + u) * N + s] : -1;
#pragma unroll
      for (int u = 0; u < SB; ++u) cv[u] = ev[u] >= 0 ? pos_of(ev[u]) : make_double4(0., 0., 0., -1.);
      test(q0, ev, cv);
    }
    // the entries are read from other lanes' columns below
    __builtin_amdgcn_wave_barrier();
    // the keys are searched in other lanes' columns from here on
    rows_ready(prow);
    rows_ready(wrow);
    // the touching pairs, spread over the lanes
    const int cnt = __popcll(hits);
    const int incl = wave_incl_sum7(cnt);
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63));
    const int first = incl - cnt;
    int added = 0;  // new history slots this particle's pairs have asked for
    uint64_t marks = 0;
    // Pairs are handled 64 per round, one per lane.  Per pair: its list
    // entry and history slot (LDS), then the partner's position, velocity
    // and spin and the history displacement (memory).  The memory reads are
    // issued by every lane at valid addresses (a lane without a pair reads
    // its owner's records and slot row 0) and the arithmetic runs on every
    // lane, so no read sits in a branch: after such a branch the compiler
    // cannot count the reads in flight and waits for all of them, the
    // previous round's history stores included.
    struct PairIn {
      int e;        // partner slot (also the history key)
      int slot;     // history slot, -1 for a new contact
      int so;       // owner's slot
      int o, o_first, o_n;  // owner's lane, its first pair's index, its history length
      bool in_t;    // the partner is in the workgroup's tile
      double4 pj, vj, wj;
      Del3<H> d;
    };
    // The partner's records are requested as soon as the list entry is
    // known, before the history key search (dependent LDS reads), which only
    // the displacement read needs.  A partner inside the tile has its
    // position (and, in the FP64 form, its velocity record) read from LDS in
    // pair_finish; those memory reads of its lane go to the tile's first slot
    // instead, one line the whole wave shares, so that the reads stay
    // unconditional.
    auto pair_fetch = [&](const int base, PairIn& q) {
      const int g = base + lane;
      const bool has = g < total;
      const int o = owner_lane(incl, g);
      const int o_first = __shfl(first, o);
      // A lane without a pair uses the slot of lane 0, which holds a particle
      // whenever the wave has pairs.
      const int so = has ? s_lane0 + o : s_lane0;
      const int o_n = __shfl(nh0, o);
      const int r = g - o_first;  // rank among the owner's touching pairs
      // (the LDS read is marked non-temporal so it stays an LDS instruction)
      int e = ld_nt(&lds.hit[r < HS ? r : HS - 1][wtx + o]);
      // Pairs past the stash: the entry is read from the list.  (The branch
      // on the ballot is uniform, so the shuffle inside runs on every lane:
      // in a divergent branch it would read lanes that are off in it.)
      if (__ballot(has && r >= HS) != 0) {
        const uint64_t o_hits = shfl64(hits, o);
        if (has && r >= HS) {
          e = A.nbr[(size_t)nth_bit(o_hits, r) * N + so];
          // wait for it here, inside the rare branch (see above)
          asm volatile("" ::"v"(e));
        }
      }
      q.so = so;
      q.o = o;
      q.o_first = o_first;
      q.o_n = o_n;
      q.e = has ? e : so;
      q.in_t = (unsigned)(q.e - tile0) < 256u;
      const int me = q.in_t ? tile0 : q.e;
      q.pj = pin[me];
      q.vj = vin[VT ? me : q.e];
      q.wj = win[q.e];
      // (keeps the compiler from sinking the loads above below the search)
      asm volatile("" ::: "memory");
      q.slot = has ? row.find_in(wtx + o, so, o_n, e) : -1;
      q.d = ld_nt(row.gd + (size_t)(q.slot >= 0 ? q.slot : 0) * N + q.so);
    };
    auto pair_finish = [&](const int base, const PairIn& q) {
      const int g = base + lane;
      const bool has = g < total;
      const int o = q.o;
      const int o_first = q.o_first;
      const int so = q.so;
      const int o_n = q.o_n;
      const int o_added = __shfl(added, o);
      // the owner's position record, from the workgroup's tile
      const double4 op = ld_nt(tile_p + (so - tile0));
      const double opx = op.x, opy = op.y, opz = op.z, ori = op.w;
      const T oinvm = __shfl((T)invm, o);
      const T ok = __shfl(k, o), oetac = __shfl(etac, o), oriT = __shfl(riT, o);
      const V3<T> ovi{__shfl(vi.x, o), __shfl(vi.y, o), __shfl(vi.z, o)};
      const V3<T> owi{__shfl(wi.x, o), __shfl(wi.y, o), __shfl(wi.z, o)};
      const int e = q.e;
      const int slot = q.slot;
      double4 pj = q.pj;
      V3<T> vj{(T)q.vj.x, (T)q.vj.y, (T)q.vj.z}, wj{(T)q.wj.x, (T)q.wj.y, (T)q.wj.z};
      T invmj = (T)q.vj.w;
      {
        const int te = (e - tile0) & 255;
        const double4 pl = ld_nt(tile_p + te);
        if constexpr (VT) {
          const double4 vl = ld_nt(lds.tv + te);
          if (q.in_t) {
            vj = V3<T>{vl.x, vl.y, vl.z};
            invmj = vl.w;
          }
        }
        if (q.in_t) pj = pl;
      }
      // new contacts take their particle's next free slots in list order
      const uint64_t fresh = __ballot(has && slot < 0);
      const uint64_t before_me = lane_range(max(o_first - base, 0), lane);
      int wslot = -1;
      if (has) {
        wslot = slot;
        if (slot < 0) {
          const int ns = o_n + o_added + __popcll(fresh & before_me);
          if (ns < row.cap) {
            wslot = ns;
            st_nt(e, row.gkey + (size_t)ns * N + so);
          } else {
            err |= GDEM_ERR_HIST_OVERFLOW_BIT;
          }
        }
      }
      // The arithmetic runs on every lane (a lane without a pair works on
      // its owner's own records and its results are never read), so the
      // memory reads it uses stay unconditional; only the store is guarded.
      const double dx = opx - pj.x, dy = opy - pj.y, dz = opz - pj.z;
      const double distsq = dx * dx + dy * dy + dz * dz;
      const double R = ori + pj.w;
      const double dist = sqrt(distsq);
      const double delMagD = R - dist;
      V3<T> n;
      if constexpr (sizeof(T) == 8) {
        n = V3<T>{dx / dist, dy / dist, dz / dist};
      } else {
        const T inv = frcp((T)dist);
        n = V3<T>{(T)dx * inv, (T)dy * inv, (T)dz * inv};
      }
      const T delMag = (T)delMagD;
      const V3<T> del = scal(-delMag, n);
      const V3<T> vr{ovi.x - vj.x, ovi.y - vj.y, ovi.z - vj.z};
      const T vrn = dot(vr, n);
      const V3<T> vn = scal(vrn, n);
      const T meff = frcp(oinvm + invmj);
      const T eta = oetac * fsqrt(meff);
      V3<T> fn;
      fn.x = -ok * del.x - eta * vn.x;
      fn.y = -ok * del.y - eta * vn.y;
      fn.z = -ok * del.z - eta * vn.z;
      const T rj = (T)pj.w;
      V3<T> vrot{oriT * owi.x + rj * wj.x, oriT * owi.y + rj * wj.y, oriT * owi.z + rj * wj.z};
      vrot = cross(vrot, n);
      const V3<T> vt = vadd(vsub(vr, vn), vrot);
      const V3<T> old = slot >= 0 ? V3<T>{(T)q.d.x, (T)q.d.y, (T)q.d.z} : V3<T>{T(0), T(0), T(0)};
      V3<T> dn;
      const V3<T> ft = tangential_force<T>(old, fn, vt, n, eta, ok, mu, dtT, dn);
      if (has && wslot >= 0) store_delta<H, T>(row.gd + (size_t)wslot * N + so, dn);
      const V3<T> mm = scal(oriT, cross(n, ft));
      // every particle adds its pairs of this round in list order
      const int my_lo = max(first - base, 0), my_hi = min(first + cnt - base, 64);
      const int my_n = max(my_hi - my_lo, 0);
      const int rounds = wave_max7(my_n);
      // The mixed form adds each pair's fn + ft as one vector (six values
      // fetched per pair instead of nine); FP64 adds fn, then ft, as the
      // reference's accumulation does.
      constexpr bool FOLD = sizeof(T) == 4 && !REC;
      const V3<T> fa = FOLD ? vadd(fn, ft) : fn;
      for (int t = 0; t < rounds; ++t) {
        const bool mine = t < my_n;
        const int src = mine ? my_lo + t : lane;
        const T a0 = __shfl(fa.x, src), a1 = __shfl(fa.y, src), a2 = __shfl(fa.z, src);
        T b0 = T(0), b1 = T(0), b2 = T(0);
        if constexpr (!FOLD) {
          b0 = __shfl(ft.x, src), b1 = __shfl(ft.y, src), b2 = __shfl(ft.z, src);
        }
        const T c0 = __shfl(mm.x, src), c1 = __shfl(mm.y, src), c2 = __shfl(mm.z, src);
        const int ws = __shfl(wslot, src);
        int rkey = 0;
        double rdel = 0.;
        if constexpr (REC) {
          rkey = __shfl(e, src);
          rdel = __shfl(delMagD, src);
        }
        if (mine) {
          f.x += a0;
          f.y += a1;
          f.z += a2;
          if constexpr (!FOLD) {
            f.x += b0;
            f.y += b1;
            f.z += b2;
          }
          mom.x += c0;
          mom.y += c1;
          mom.z += c2;
          if (ws >= 0) marks |= (uint64_t)1 << ws;
          if constexpr (REC) {
            if (nrec < rec.cap) {
              const size_t r = (size_t)self_id * rec.cap + nrec;
              rec.key[r] = A.pid[rkey];  // the partner's original id
              rec.kind[r] = 0;
              rec.fn[r * 3] = a0; rec.fn[r * 3 + 1] = a1; rec.fn[r * 3 + 2] = a2;
              rec.ft[r * 3] = b0; rec.ft[r * 3 + 1] = b1; rec.ft[r * 3 + 2] = b2;
              rec.delta[r] = rdel;
              ++nrec;
            }
          }
        }
      }
      added += __popcll(fresh & lane_range(my_lo, my_hi));
    };
    for (int base = 0; base < total; base += 64) {
      PairIn q;
      pair_fetch(base, q);
      pair_finish(base, q);
    }
    row.n = min(nh0 + added, row.cap);
    row.marks = marks;
    if (act) st_nt(truncate_history(row.n, row.marks), A.nh + s);
  }

  // ---- walls ----------------------------------------------------------------
  // The wave's wall candidates are spread over the lanes, 64 at a time, like
  // the touching pairs: each lane tests its candidate's distance and, for a
  // contact, evaluates the contact force.  A vertex or edge contact whose
  // feature an earlier contact of the same particle (in list order) already
  // touched is skipped (d_particle_collision_verlet's hitAt bookkeeping);
  // new history slots go to the kept contacts in list order; each particle
  // then adds its kept contacts in list order.
  {
    WRow& row = wrow;
    // (every lane's pair pass is done before the wall pass writes the union)
    __builtin_amdgcn_wave_barrier();
    // features of this particle's kept vertex/edge contacts of earlier rounds:
    // *ve_slot(lds, tx, u), u < min(nve, VEM)
    int nve = 0;
    int added = 0;  // new history slots this particle's contacts have taken
    uint64_t marks = 0;
    const T eta = etac * sqrtm;
    const int incl = w_incl, total = w_total, first = incl - nwl;
    // One round: candidates base + lane; t_in is this lane's candidate.
    auto wall_round = [&](const int base, const int t_in) {
      const int g = base + lane;
      const bool has = g < total;
      const int o = owner_lane(incl, g);
      const int o_first = __shfl(first, o);
      const int so = s_lane0 + o;
      const double4 op = ld_nt(tile_p + (so - tile0));  // the owner's position record
      const double opx = op.x, opy = op.y, opz = op.z, ori = op.w;
      const double orsq = ori * ori;
      const int o_n = __shfl(nw0, o), o_added = __shfl(added, o), o_nve = __shfl(nve, o);
      int t = 0, slot = -1, hit = -1;
      double dist = 1.e10;
      V3<T> nT{T(0), T(0), T(0)}, old{T(0), T(0), T(0)};
      if (has) {
        // candidate, its history slot and displacement, and its distance
        t = t_in;
        // The triangle's first fields are requested before the history key
        // search (dependent LDS reads); the empty asm keeps the compiler
        // from sinking the loads below it.  The displacement (slot row 0,
        // always allocated, stands in when there is none, so the load needs
        // no branch) follows the search.
        const TriFirst tf = tri_first(A.tri + t);
        asm volatile("" ::: "memory");
        slot = row.find_in(wtx + o, so, o_n, t);
        const Del3<H> d = ld_nt(row.gd + (size_t)(slot >= 0 ? slot : 0) * N + so);
        const TriHit th = tri_distance_2trip(tf, A.tri + t, opx, opy, opz, orsq);
        if (slot >= 0) old = V3<T>{(T)d.x, (T)d.y, (T)d.z};
        dist = th.dist;
        hit = th.hit;
        nT = V3<T>{(T)th.n.x, (T)th.n.y, (T)th.n.z};
      }
      const bool con = has && dist < ori;
      // this particle's candidates of the round are lanes [my_lo, my_hi);
      // the candidate's owner's are [c_lo, ...)
      const int my_lo = max(first - base, 0), my_hi = min(first + nwl - base, 64);
      const int my_n = max(my_hi - my_lo, 0);
      const int span = wave_max7(my_n);
      const int c_lo = max(o_first - base, 0);
      bool dup = false;
      if (con && hit != -1) {
        for (int z = 0; z < min(o_nve, VEM); ++z)
          if (*ve_slot(lds, wtx + o, z) == hit) dup = true;
      }
      for (int kk = 1; kk < span; ++kk) {
        const int src = max(lane - kk, 0);
        const int c_src = __shfl((int)con, src), h_src = __shfl(hit, src);
        if (lane - kk >= c_lo && con && hit != -1 && c_src && h_src == hit) dup = true;
      }
      const bool keep = con && !dup;
      // new contacts take their particle's next free slots in list order
      const uint64_t fresh = __ballot(keep && slot < 0);
      int wslot = slot;
      if (keep && slot < 0) {
        const int ns = o_n + o_added + __popcll(fresh & lane_range(c_lo, lane));
        if (ns < row.cap) {
          wslot = ns;
          st_nt(t, row.gkey + (size_t)ns * N + so);
        } else {
          err |= GDEM_ERR_WHIST_OVERFLOW_BIT;
        }
      }
      const T ok = __shfl(k, o), oeta = __shfl(eta, o), oriT = __shfl(riT, o);
      const V3<T> ovi{__shfl(vi.x, o), __shfl(vi.y, o), __shfl(vi.z, o)};
      const V3<T> owi{__shfl(wi.x, o), __shfl(wi.y, o), __shfl(wi.z, o)};
      V3<T> fn{T(0), T(0), T(0)}, ft{T(0), T(0), T(0)}, mm{T(0), T(0), T(0)};
      double delMagD = 0.;
      if (keep) {
        delMagD = ori - dist;
        const T delMag = (T)delMagD;
        const V3<T> n = nT;
        const V3<T> del = scal(-delMag, n);
        const T vrn = dot(ovi, n);
        const V3<T> vn = scal(vrn, n);
        fn.x = -ok * del.x - oeta * vn.x;
        fn.y = -ok * del.y - oeta * vn.y;
        fn.z = -ok * del.z - oeta * vn.z;
        V3<T> vrot{oriT * owi.x, oriT * owi.y, oriT * owi.z};
        vrot = cross(vrot, n);
        const V3<T> vt = vadd(vsub(ovi, vn), vrot);
        V3<T> dn;
        ft = tangential_force<T>(old, fn, vt, n, oeta, ok, mu, dtT, dn);
        if (wslot >= 0) store_delta<H, T>(row.gd + (size_t)wslot * N + so, dn);
        mm = scal(oriT, cross(n, ft));
      }
      // every particle adds its kept contacts of this round in list order
      // (fn + ft folded in the mixed form, as for the pairs)
      constexpr bool FOLD = sizeof(T) == 4 && !REC;
      const V3<T> fa = FOLD ? vadd(fn, ft) : fn;
      for (int u = 0; u < span; ++u) {
        const bool mine = u < my_n;
        const int src = mine ? my_lo + u : lane;
        const int kp = __shfl((int)keep, src);
        const T a0 = __shfl(fa.x, src), a1 = __shfl(fa.y, src), a2 = __shfl(fa.z, src);
        T b0 = T(0), b1 = T(0), b2 = T(0);
        if constexpr (!FOLD) {
          b0 = __shfl(ft.x, src), b1 = __shfl(ft.y, src), b2 = __shfl(ft.z, src);
        }
        const T c0 = __shfl(mm.x, src), c1 = __shfl(mm.y, src), c2 = __shfl(mm.z, src);
        const int ws = __shfl(wslot, src), hh = __shfl(hit, src);
        int rkey = 0;
        double rdel = 0.;
        if constexpr (REC) {
          rkey = __shfl(t, src);
          rdel = __shfl(delMagD, src);
        }
        if (mine && kp) {
          f.x += a0;
          f.y += a1;
          f.z += a2;
          if constexpr (!FOLD) {
            f.x += b0;
            f.y += b1;
            f.z += b2;
          }
          mom.x += c0;
          mom.y += c1;
          mom.z += c2;
          if (ws >= 0) marks |= (uint64_t)1 << ws;
          if (hh != -1) {
            if (nve < VEM) *ve_slot(lds, tx, nve) = hh;
            else err |= GDEM_ERR_VE_OVERFLOW_BIT;
            ++nve;
          }
          if constexpr (REC) {
            if (nrec < rec.cap) {
              const size_t r = (size_t)self_id * rec.cap + nrec;
              rec.key[r] = rkey;
              rec.kind[r] = 1;
              rec.fn[r * 3] = a0; rec.fn[r * 3 + 1] = a1; rec.fn[r * 3 + 2] = a2;
              rec.ft[r * 3] = b0; rec.ft[r * 3 + 1] = b1; rec.ft[r * 3 + 2] = b2;
              rec.delta[r] = rdel;
              ++nrec;
            }
          }
        }
      }
      added += __popcll(fresh & lane_range(my_lo, my_hi));
    };
    // The first round is separate so that its candidates are the ones
    // fetched at the start: a candidate loaded in the round itself makes the
    // compiler wait for every earlier memory access, the pair pass's history
    // stores included, before the key search.  Waves with more than 64 wall
    // candidates fetch the later ones round by round.
    if (total > 0) wall_round(0, w_t0);
    for (int base = 64; base < total; base += 64) {
      const int g = base + lane;
      const int o = owner_lane(incl, g);
      const int o_first = __shfl(first, o), so = s_lane0 + o;
      const int tn = g < total ? ld_nt(A.wl + (size_t)(g - o_first) * N + so) : 0;
      wall_round(base, tn);
    }
    row.n = min(nw0 + added, row.cap);
    row.marks = marks;
    if (act) st_nt(truncate_history(row.n, row.marks), A.nw + s);
  }
  if constexpr (REC) {
    if (act) rec.count[self_id] = nrec;
  }
  if (!act) return;

  // ---- integration (k_integrate), out-of-bounds, refresh check --------------
  // The particle's own records are read again here (the position from the
  // tile, the rest from memory, where they are still cached) rather than held
  // in registers through the contact passes, which would spill at either
  // form's occupancy bound; the empty asm keeps the compiler from reusing the
  // values loaded at the start.
  asm volatile("" ::: "memory");
  const double4 p1 = ld_nt(tile_p + (s - tile0));
  const double4 v1 = vin[s];
  const double4 w1 = win[s];
  const double dt = A.dt;
  const double ax = (double)f.x * invm + A.g[0];
  const double ay = (double)f.y * invm + A.g[1];
  const double az = (double)f.z * invm + A.g[2];
  double4 vo, po, wo;
  vo.x = v1.x + ax * dt;
  vo.y = v1.y + ay * dt;
  vo.z = v1.z + az * dt;
  vo.w = v1.w;
  po.x = p1.x + vo.x * dt;
  po.y = p1.y + vo.y * dt;
  po.z = p1.z + vo.z * dt;
  const double bx = (double)mom.x * invmoi;
  const double by = (double)mom.y * invmoi;
  const double bz = (double)mom.z * invmoi;
  if constexpr (REC) {
    if (rec.acc) {
      rec.acc[(size_t)self_id * 3] = ax;
      rec.acc[(size_t)self_id * 3 + 1] = ay;
      rec.acc[(size_t)self_id * 3 + 2] = az;
    }
    if (rec.anga) {
      rec.anga[(size_t)self_id * 3] = bx;
      rec.anga[(size_t)self_id * 3 + 1] = by;
      rec.anga[(size_t)self_id * 3 + 2] = bz;
    }
  }
  wo.x = w1.x + bx * dt;
  wo.y = w1.y + by * dt;
  wo.z = w1.z + bz * dt;
  wo.w = w1.w;
  const bool out = po.x < A.bmin[0] || po.x > A.bmax[0] || po.y < A.bmin[1] || po.y > A.bmax[1] ||
                   po.z < A.bmin[2] || po.z > A.bmax[2];
  po.w = out ? -p1.w : p1.w;
  if (!out) refresh = refresh_test(po, w1.w, A.ref + s, A);
  if (!(isfinite(po.x) && isfinite(po.y) && isfinite(po.z) && isfinite(vo.x) && isfinite(vo.y) &&
        isfinite(vo.z)))
    err |= GDEM_ERR_NAN_BIT;
  st_nt(po, A.pout + s);
  st_nt(vo, A.vout + s);
  st_nt(wo, A.wout + s);
}

// What the candidate-list builders read and write.  Cell ids are global:
// the grid of this scene occupies [cell_base, cell_base + size.x*size.y*size.z).
struct ListArgs {
  const double4* pos;
  const int* pid;
  const int* cell_of;
  const int2* cell_tab;
  int* nbr;
  int* nnb;
  double4* ref;
  double4* angv;  // spin records of the same state buffer as pos (fourth word: packed ref)
  int* wl;
  int* nwl;
  const TriGeo* tri;
  const BvhNode* bvh;
  const double* tbox;
  const int* wc_off;
  const int* wc_tri;
  unsigned int* err;
  int N;
  int size[3];
  int cell_base;
  double skin;
  double qlo[3], qhi[3], qh[3];  // box and step of the packed reference position
  int nb_cap, wall_cap;
  int ntri, wall_table;
};

// Verlet pair list (k_update_neighborlist_endsort, BoundingBox.cu:1106-1162):
// all active j != i in the 27 surrounding cells with d.d < (r_i+r_j+skin)^2,
// ordered by original id.  Also records the reference position.
constexpr int NB_MAX = 64;

// Candidates are met slab by slab: each z-slab of 3 x 3 cells has its nine
// cell ranges loaded at once, then its particles are tested PL_BATCH at a
// time with their positions and ids loaded together (a few memory round trips
// per slab instead of one per cell and per particle).  visit(q, id) is called
// for every candidate in range.
constexpr int PL_BATCH = 16;

template <class Visit>
__device__ __forceinline__ void pair_candidates(const ListArgs& L, int s, const double4& p, Visit&& visit) {
  const double4* __restrict__ pos = L.pos;
  const int* __restrict__ pid = L.pid;
  const int c = L.cell_of[s] - L.cell_base;
  const int sx = L.size[0], sy = L.size[1], sz = L.size[2];
  const int cx = c % sx;
  const int cy = (c / sx) % sy;
  const int cz = c / (sx * sy);
  for (int dz = -1; dz <= 1; ++dz) {
    const int Z = cz + dz;
    if (Z < 0 || Z >= sz) continue;
    int st[9], cnt[9];
    int total = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int Y = cy + k / 3 - 1, X = cx + k % 3 - 1;
      int2 ce = make_int2(-1, -1);
      if (Y >= 0 && Y < sy && X >= 0 && X < sx) ce = L.cell_tab[L.cell_base + (sy * Z + Y) * sx + X];
      st[k] = ce.x;
      cnt[k] = ce.x < 0 ? 0 : ce.y - ce.x;
    }
    // The slab's t-th particle is slot t + off(k) for the last cell k whose
    // first particle is number pre[k] <= t (off(k) = its first slot minus
    // pre[k]); off is formed as a telescoping sum of the differences dlt, so
    // no cell is indexed at run time (empty cells cancel out).
    int pre[9], dlt[9];
    {
      int prev = 0;
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int off = st[k] - total;
        pre[k] = total;
        dlt[k] = off - prev;
        prev = off;
        total += cnt[k];
      }
    }
    for (int t0 = 0; t0 < total; t0 += PL_BATCH) {
      int qv[PL_BATCH];
#pragma unroll
      for (int u = 0; u < PL_BATCH; ++u) {
        const int t = t0 + u;
        int q = t;
#pragma unroll
        for (int k = 0; k < 9; ++k) q += t >= pre[k] ? dlt[k] : 0;
        qv[u] = (t < total && q != s) ? q : -1;
      }
      double4 pj[PL_BATCH];
      int idj[PL_BATCH];
#pragma unroll
      for (int u = 0; u < PL_BATCH; ++u) {
        pj[u] = qv[u] >= 0 ? pos[qv[u]] : make_double4(0., 0., 0., -1.);
        idj[u] = qv[u] >= 0 ? pid[qv[u]] : 0;
      }
#pragma unroll
      for (int u = 0; u < PL_BATCH; ++u) {
        if (qv[u] < 0) continue;
        const double ddx = p.x - pj[u].x, ddy = p.y - pj[u].y, ddz = p.z - pj[u].z;
        const double distsq = ddx * ddx + ddy * ddy + ddz * ddz;
        const double R2 = p.w + pj[u].w + L.skin;
        if (distsq < R2 * R2) visit(qv[u], idj[u]);
      }
    }
  }
}

// Lists up to PL_LDS_CAP long are gathered in the thread's LDS column
// (col[k * COLS], COLS the workgroup's thread count), one store per accepted
// candidate, and sorted there once at the end; longer capacities use private
// arrays.
constexpr int PL_BLOCK = 128;   // threads per workgroup of the list kernels
constexpr int PL_LDS_CAP = 32;

template <int COLS>
struct PairLdsT {
  int2 e[PL_LDS_CAP][COLS];
};
using PairLds = PairLdsT<PL_BLOCK>;

template <int COLS>
__device__ __forceinline__ int pairlist_lds(const ListArgs& L, int s, int2* col) {
  const int N = L.N;
  const double4 p = L.pos[s];
  const int cap = L.nb_cap;
  int n = 0;
  bool ovf = false;
  pair_candidates(L, s, p, [&](int q, int id) {
    if (n >= cap) {
      ovf = true;
      return;
    }
    col[n * COLS] = make_int2(q, id);
    ++n;
  });
  // ids are distinct, so the order is total
  for (int i = 1; i < n; ++i) {
    const int2 x = col[i * COLS];
    int j = i - 1;
    for (; j >= 0; --j) {
      const int2 y = col[j * COLS];
      if (y.y < x.y) break;
      col[(j + 1) * COLS] = y;
    }
    col[(j + 1) * COLS] = x;
  }
  for (int k = 0; k < n; ++k) L.nbr[(size_t)k * N + s] = col[k * COLS].x;
  L.nnb[s] = n;
  if (ovf) atomicOr(L.err, GDEM_ERR_NB_OVERFLOW_BIT);
  return n;
}

__device__ __forceinline__ int pairlist_arrays(const ListArgs& L, int s) {
  const int N = L.N;
  const double4 p = L.pos[s];
  const int cap = L.nb_cap < NB_MAX ? L.nb_cap : NB_MAX;
  int cand[NB_MAX];
  int cpid[NB_MAX];
  int n = 0;
  bool ovf = false;
  pair_candidates(L, s, p, [&](int q, int id) {
    if (n >= cap) {
      ovf = true;
      return;
    }
    int at = n;
    while (at > 0 && cpid[at - 1] > id) {
      cpid[at] = cpid[at - 1];
      cand[at] = cand[at - 1];
      --at;
    }
    cpid[at] = id;
    cand[at] = q;
    ++n;
  });
  for (int q = 0; q < n; ++q) L.nbr[(size_t)q * N + s] = cand[q];
  L.nnb[s] = n;
  if (ovf) atomicOr(L.err, GDEM_ERR_NB_OVERFLOW_BIT);
  return n;
}

// lds: the workgroup's list columns (one per thread of a COLS-thread
// workgroup).  LONG: capacities above PL_LDS_CAP, gathered in private arrays
// (a separate kernel instantiation, so the usual one carries no scratch).
// Returns the list length.
template <int COLS, bool LONG>
__device__ __forceinline__ int pairlist_one(const ListArgs& L, int s, bool active, PairLdsT<COLS>& lds) {
  {
    const double4 p = L.pos[s];
    L.ref[s] = p;
    L.angv[s].w = refq_pack(p, L.qlo, L.qhi, L.qh);
  }
  if (!active) {
    L.nnb[s] = 0;
    return 0;
  }
  if constexpr (LONG) return pairlist_arrays(L, s);
  else return pairlist_lds<COLS>(L, s, &lds.e[0][threadIdx.x]);
}

// Wall candidates by BVH traversal (k_update_neighborlist_wall,
// BVH.cu:493-540): leaf box test against (r+skin)^2, exact test
// dist < r + skin, ascending triangle ids.
constexpr int WL_MAX = 48;

// BVH: traversal (its stack and sort buffer are private arrays, so it is a
// kernel instantiation of its own); otherwise the per-cell table, which every
// scene with triangles and no BVH request has.
template <bool BVH>
__device__ __forceinline__ void walllist_one(const ListArgs& L, int s, bool active) {
  const int N = L.N;
  if (!active || L.ntri == 0) {
    L.nwl[s] = 0;
    return;
  }
  const double4 p = L.pos[s];
  const double R = p.w + L.skin;
  const double Rsq = R * R;
  const int cap = L.wall_cap < WL_MAX ? L.wall_cap : WL_MAX;
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
  if constexpr (!BVH) {
    // static per-cell table: ascending triangle ids already, so candidates
    // go straight to their slots
    // (entries and their boxes are loaded WL_BATCH at a time)
    int n = 0;
    bool ovf = false;
    constexpr int WL_BATCH = 4;
    const int c = L.cell_of[s];
    const int b0 = L.wc_off[c], b1 = L.wc_off[c + 1];
    for (int q0 = b0; q0 < b1; q0 += WL_BATCH) {
      int tv[WL_BATCH];
#pragma unroll
      for (int u = 0; u < WL_BATCH; ++u) tv[u] = q0 + u < b1 ? L.wc_tri[q0 + u] : -1;
      bool near[WL_BATCH];
#pragma unroll
      for (int u = 0; u < WL_BATCH; ++u) {
        near[u] = false;
        if (tv[u] >= 0) {
          const double* bx = L.tbox + (size_t)tv[u] * 6;
          near[u] = consider(tv[u], bx, bx + 3);
        }
      }
      for (int u = 0; u < WL_BATCH; ++u) {
        if (!near[u]) continue;
        const int t = tv[u];
        const TriHit h = tri_distance(L.tri + t, p.x, p.y, p.z, Rsq);
        if (h.dist < p.w + L.skin) {
          if (n >= cap) {
            ovf = true;
            continue;
          }
          L.wl[(size_t)n * N + s] = t;
          ++n;
        }
      }
    }
    L.nwl[s] = n;
    if (ovf) atomicOr(L.err, GDEM_ERR_WALL_OVERFLOW_BIT);
  } else {
    int list[WL_MAX];
    int n = 0;
    bool ovf = false;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const BvhNode* nd = L.bvh + stack[--sp];
      const int nchild = (nd->child[0] == nd->child[1]) ? 1 : 2;
      for (int c = 0; c < nchild; ++c) {
        if (!consider(-1, nd->lo[c], nd->hi[c])) continue;
        const int ch = nd->child[c];
        if (ch >= 0) {
          if (sp < 64) stack[sp++] = ch;
          else ovf = true;
        } else {
          const int t = ~ch;
          const TriHit h = tri_distance(L.tri + t, p.x, p.y, p.z, Rsq);
          if (h.dist < p.w + L.skin) {
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
    for (int q = 0; q < n; ++q) L.wl[(size_t)q * N + s] = list[q];
    L.nwl[s] = n;
    if (ovf) atomicOr(L.err, GDEM_ERR_WALL_OVERFLOW_BIT);
  }
}

}  // namespace gdem
