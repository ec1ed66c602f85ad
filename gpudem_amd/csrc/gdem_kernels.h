// Launch interface between the host engine and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_types.h"

#define GDEM_ERR_NB_OVERFLOW_BIT 1u
#define GDEM_ERR_WALL_OVERFLOW_BIT 2u
#define GDEM_ERR_HIST_OVERFLOW_BIT 4u
#define GDEM_ERR_WHIST_OVERFLOW_BIT 8u
#define GDEM_ERR_VE_OVERFLOW_BIT 16u
#define GDEM_ERR_NAN_BIT 32u
#define GDEM_ERR_BARRIER_BIT 64u

namespace gdem {

// Stable LSD radix sort of 32-bit keys: three passes of 11-bit digits; the
// sorted keys and their source indices end in buffer RS_FINAL of the
// ping-pong pair.  The count table holds RS_DIGITS entries per workgroup plus
// RS_DIGITS digit totals.
constexpr int RS_BITS = 11;
constexpr int RS_DIGITS = 1 << RS_BITS;
constexpr int RS_PASSES = 3;
constexpr int RS_FINAL = RS_PASSES & 1;
// Two-level grid barrier of the rebuild kernel: BAR_SUB arrival
// counters, a top counter, a generation word and an abort word, each on its
// own 128 B line.
constexpr int BAR_SUB = 32;
constexpr int GBAR_ABORT = (BAR_SUB + 2) * 32;
constexpr int GBAR_WORDS = (BAR_SUB + 3) * 32;
constexpr int GBAR_SPIN = 1 << 20;  // default poll limit of a barrier wait (about a second)

struct RecDev {
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

// device copies of a gdem_history_view (original particle order)
struct HistIn {
  const int* nhist;
  const int* hist_key;
  const double* hist_delta;
  const int* nwhist;
  const int* whist_key;
  const double* whist_delta;
};

struct DebugDev {
  int* cellx;
  int* cell_id;
  uint32_t* morton;
  int* nb;
  int* nnb;
  int* wall;
  int* nwall;
  int* hist_key;
  double* hist_delta;
  int* nhist;
  int* whist_key;
  double* whist_delta;
  int* nwhist;
};

// Allocation-free download: the staging array (gdem_noalloc.hip) holds up to
// NOALLOC_MAX_N particles.
constexpr int NOALLOC_MAX_N = 1 << 18;
struct NoallocStage {
  double* x;  // [NOALLOC_MAX_N * 3]
  double* v;
  double* w;
  int* act;   // [NOALLOC_MAX_N]
};
// Resolves the staging array's address on `device`, which must be current.
hipError_t noalloc_init(int device);
// The staging pointers of `device` (null members before noalloc_init).
NoallocStage noalloc_stage(int device);

// per-slot view of the last rebuild (gdem_order_view)
struct OrderDev {
  uint32_t* key;
  int* order;
  int* cell;
  int* first;
  int* end;
  int* nsorted;
};

template <class T, class H>
void launch_step(const SimParams& P, const Bufs<H>& B, hipStream_t st);
template <class T, class H>
void launch_step_record(const SimParams& P, const Bufs<H>& B, const RecDev& r, hipStream_t st);
// coop: one grid-barrier kernel (grid = P.rs_blocks, co-resident); otherwise
// the chain of per-phase kernels.  Both can be captured into a hipGraph.
template <class H>
hipError_t launch_rebuild(const SimParams& P, const Bufs<H>& B, bool coop, hipStream_t st);
// Largest co-resident grid of the grid-barrier rebuild kernel (0 if the
// occupancy query fails).
template <class H>
int coop_grid_limit();
// Grid of the pair-list kernel for n slots: no more workgroups than can be
// resident at once (the kernel strides over the slots).
template <class H>
int list_grid(int n);
template <class H>
void launch_download(const SimParams& P, const Bufs<H>& B, double* x, double* v, double* w, int* a,
                     hipStream_t st);
template <class H>
void launch_debug(const SimParams& P, const Bufs<H>& B, const DebugDev& D, hipStream_t st);
template <class H>
void launch_debug_order(const SimParams& P, const Bufs<H>& B, const OrderDev& D, hipStream_t st);
template <class H>
void launch_counts(const SimParams& P, const Bufs<H>& B, unsigned long long* acc, hipStream_t st);
template <class H>
void launch_set_history(const SimParams& P, const Bufs<H>& B, const HistIn& in, hipStream_t st);
// The active particles' positions and radii in original id order, scaled by
// lf and converted to float32 (cnt: frame_blocks(N) + 1 ints of workspace;
// cnt[frame_blocks(N)] receives the active count).  Uses inv.
template <class H>
void launch_frame(const SimParams& P, const Bufs<H>& B, int* cnt, double lf, float* xyz, float* r, hipStream_t st);
int frame_blocks(int n);

void launch_radix_sort(int n, int blocks, int chunk, uint32_t* k0, int* i0, uint32_t* k1, int* i1, int* inv_out,
                       uint32_t* hist, const int* ctl, bool gated, hipStream_t st);
void launch_tri_probe(const TriGeo* tri, int np, const double* pts, const double* rsq, double* out,
                      int* hit, hipStream_t st);
void launch_copy(const double4* a, double4* b, size_t n, hipStream_t st);
void launch_fp64_probe(int n, const double* a, const double* b, double* sq, double* dv, hipStream_t st);

}  // namespace gdem
