// Device-side data layout of the gdem engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gdem {

// One triangle, indexed by its original (file-order) id; 256 B so a wave's
// gathers stay line-aligned.  Fields are the precompute of
// TriangleMesh.cu:536-677: vertices in ascending merged-vertex-id order,
// edge vectors, barycentric dot products and their inverses, plane.
struct alignas(256) TriGeo {
  double n[3];
  double d;  // plane offset
  double p0[3], p1[3], p2[3];
  double e01[3], e02[3], e12[3];
  double d00, d01, d11, denom, d00inv, d11inv, d22inv;
  int v0, v1, v2;  // merged vertex ids (hitAt of a vertex contact)
  int e0, e1, e2;  // edge ids + nShift (hitAt of an edge contact)
};
static_assert(sizeof(TriGeo) == 256, "TriGeo must be 256 B");

// Bounding-volume node with both children's boxes inline: one 128 B load per
// traversal step.  child >= 0 is an internal node, child < 0 a leaf holding
// triangle ~child.
struct alignas(128) BvhNode {
  double lo[2][3];
  double hi[2][3];
  int child[2];
  int pad[6];
};
static_assert(sizeof(BvhNode) == 128, "BvhNode must be 128 B");

// Per-scene constants passed by value to every kernel.
struct SimParams {
  double dt, mu;
  double g[3];
  double bmin[3], bmax[3];
  double invd[3];
  double skin, threshSq;
  int size[3];
  int use_morton;  // sort by morton3D of the cell (all axes <= 1024 cells) or by cell id
  int N;
  int nb_cap, wall_cap, hist_cap, whist_cap;
  int ntri, nnode;
  int rs_blocks, rs_chunk;  // radix sort grid and per-block chunk
  int list_blocks;          // grid of the pair-list kernel (co-resident workgroups, grid-stride)
  int gbar_spin;            // poll limit of the rebuild kernel's grid barriers
  int wall_table;           // 1: wall candidates from the static per-cell table, 0: BVH traversal
  int list_prefetch;        // 1: the step fetches the first 16 pair-list entries with the own records
};

// Control words shared by all kernels of a scene (device memory).
enum Ctl : int {
  CTL_CUR = 0,       // which state buffer is current (flips every step)
  CTL_HCUR = 1,      // which material/history buffer is current (flips at re-sort)
  CTL_REFRESH = 2,   // set by the step kernel when a list rebuild is due
  CTL_ERR = 3,       // GDEM_ERR_* bits
  CTL_DONE_STEP = 4, // last-block counters
  CTL_DONE_PERM = 5,
  CTL_NACT = 6,      // active particles after the last re-sort
  CTL_DONE_WALL = 7,
  CTL_REBUILDS = 8,
  CTL_NBSUM = 10,    // words 10-11: the sum of the pair-list lengths built by the last rebuild
                     // (unsigned 64-bit; the control block is 8-byte aligned)
  CTL_WORDS = 16
};

// 4-wide vector of the history/material precision.
template <class H>
struct Vec4;
template <>
struct Vec4<double> {
  using type = double4;
};
template <>
struct Vec4<float> {
  using type = float4;
};

// A tangential history displacement, packed (12 B in the mixed form, 24 B in
// FP64): a wave's slot row is read and written as one contiguous run.
template <class H>
struct Del3 {
  H x, y, z;
};

// All device buffers of a scene.  Passed by value; which half of a [2]
// array is current is read from the control words at kernel start.
template <class H>
struct Bufs {
  using H4 = typename Vec4<H>::type;
  double4* pos[2];   // x, y, z, r   (r < 0 marks an inactive particle)
  double4* vel[2];   // vx, vy, vz, 1/m
  double4* angv[2];  // wx, wy, wz, packed position at the last list build
  H4* mat[2];        // k, etaconst, sqrt(m), 1/I (FP64, or FP32 in the mixed form)
  int* pid[2];       // original particle id
  int* hkey[2];      // pair history keys (partner slot), slot-major [hist_cap][N]
  Del3<H>* hd[2];    // pair history deltas, slot-major [hist_cap][N]
  int* nh[2];
  int* wkey[2];      // wall history keys [whist_cap][N]
  Del3<H>* wd[2];    // [whist_cap][N]
  int* nw[2];
  double4* ref;      // positions at the last list build (exact; the packed form is angv's w)
  int* nbr;          // pair candidates (partner slot), slot-major [nb_cap][N], ascending original id
  int* nnb;
  int* inv;          // new slot of each old slot at a rebuild (history keys are remapped through it)
  int* wl;           // wall candidates (triangle id) [wall_cap][N]
  int* nwl;
  uint32_t* key[2];  // radix sort ping-pong
  int* idx[2];
  int* cell_tmp;     // cell id per particle before the permutation
  int* cell_of;      // cell id per particle after it
  int2* cell_tab;    // [ncell] (first, end) of each occupied cell, (-1, -1) otherwise
  uint32_t* rs_hist; // radix counts [RS_DIGITS * rs_blocks] + digit totals [RS_DIGITS]
  uint32_t* gbar;    // grid barrier words of the rebuild kernel [GBAR_WORDS], zeroed at creation
  const TriGeo* tri;
  const BvhNode* bvh;
  const double* tbox;   // per-triangle box [ntri][6]: lo xyz, hi xyz
  const int* wc_off;    // per-cell wall table [ncell + 1]
  const int* wc_tri;
  int* ctl;
};

// ---------------------------------------------------------------------------
// batched scenes
// ---------------------------------------------------------------------------
// Many independent scenes share one set of arrays.  Scene s owns a contiguous
// particle segment [seg0, seg0 + nseg) whose length is a multiple of 64, so
// every wave belongs to exactly one scene; a contiguous range of global cell
// ids [cell_base, cell_base + ncell] (the extra entry ends the scene's wall
// table); and a contiguous range of triangles.  Constants are written once at
// admission.
struct BScene {
  double dt, mu;
  double g[3];
  double bmin[3], bmax[3];
  double invd[3];
  double skin, threshSq;
  int size[3];
  int cell_base, ncell;
  int seg0, nseg;
  int n;  // real particles (the rest of the segment is inactive padding)
  int ntri, wall_table;
  int start_step, end_step;  // the scene steps while start_step <= STEP < end_step
};

// Mutable per-scene words.  The current state buffer of a scene is
// (STEP & 1) ^ q and its current material/history buffer is q: a rebuild
// permutes the scene into the other buffers and flips q, leaving every other
// scene untouched.
// A scene's cell table is scanned by SCAN_CHUNKS workgroups, one per
// contiguous chunk of its cells.
constexpr int SCAN_CHUNKS = 8;

struct SlotCtl {
  int q;        // buffer parity
  int q2;       // parity after the rebuild in progress
  int fl[2];    // refresh requested by the step kernel at STEP parity 0 / 1
  int inact;    // inactive particles counted by the binning pass
  int nact;     // active particles after the last rebuild
  unsigned int err;
  int rebuilds;
  int chunk[SCAN_CHUNKS];  // active particles binned into each chunk of the cells (zero outside a rebuild)
};

enum BCtl : int {
  BC_STEP = 0,  // batch steps completed
  BC_NDUE0 = 1,  // entries of the due-scene list of STEP parity 0 / 1
  BC_NDUE1 = 2,
  BC_DONE = 3,  // last-block counter of the step kernel
  BC_WORDS = 8
};

template <class H>
struct BBufs {
  using H4 = typename Vec4<H>::type;
  double4* pos[2];
  double4* vel[2];
  double4* angv[2];
  H4* mat[2];
  int* pid[2];  // id within the scene, -1 for padding
  int* hkey[2];
  Del3<H>* hd[2];
  int* nh[2];
  int* wkey[2];
  Del3<H>* wd[2];
  int* nw[2];
  double4* ref;
  int* nbr;
  int* nnb;
  int* wl;
  int* nwl;
  int* bin_cell;   // global cell id of each slot at binning (-1 inactive)
  int* bin_rank;   // rank within its cell (or among the inactive)
  int* cell_of;    // global cell id after the permutation
  int* ccount;     // per-cell counters, zero outside a rebuild
  int2* cell_tab;  // per-cell (first, end) slots
  const TriGeo* tri;
  const double* tbox;
  const int* wc_off;  // per global cell (plus one end entry per scene)
  const int* wc_tri;  // global triangle ids
  const int* wave_scene;  // scene of every wave of 64 slots, -1 if free
  const BScene* scenes;
  SlotCtl* sctl;
  // Scenes that requested a rebuild, per STEP parity: due[p * nslots + j],
  // j < ctl[BC_NDUE0 + p], each (scene, first wave, waves, 0).  A scene is
  // listed at most once per parity (its flag fl[p] guards the append);
  // eviction scrubs a scene's entries to (-1, 0, 0, 0), and admission reuses
  // a scrubbed entry before it appends, so a list never holds more than
  // nslots entries.
  int4* due;
  int* ctl;
  int cap;  // slot stride of the slot-major arrays
  int nb_cap, wall_cap, hist_cap, whist_cap;
  int nslots;
};

}  // namespace gdem
