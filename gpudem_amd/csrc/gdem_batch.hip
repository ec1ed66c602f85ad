// gfx950 kernels of the batched engine: many independent scenes stepped by
// one launch per phase.
//
// Every scene owns a 64-aligned segment of slots, so each wave serves exactly
// one scene and reads that scene's constants with scalar loads.  A step is
//   kb_bin -> kb_scan -> kb_permute -> kb_lists   (only the scenes on the
//                                                  due list, whose refresh
//                                                  flag is set, do any work)
//   kb_step                                       (forces + integration of
//                                                  every running scene)
// and is identical for every step, so a run of steps is captured once into a
// hipGraph and replayed.  A rebuild re-bins one scene by cell with a counting
// sort (atomic per-cell ranks, one workgroup scanning the scene's cells, a
// scatter into the other buffer half), so it never touches other scenes.  The
// contact step itself is the same device code as the single-scene kernel, and
// the lists order contacts by particle id, so a scene's trajectory in a batch
// is bit-identical to the same scene stepped alone.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gdem_batch.h"
#include "gdem_device.h"

namespace gdem {

// The step kernel running while STEP == u raises refresh flags at parity u&1;
// the rebuild that follows it runs while STEP == u+1 and so reads parity
// (STEP+1)&1.
__device__ __forceinline__ int rebuild_parity(int step) { return (step + 1) & 1; }

// The batch step counter and the length of the due list that the rebuild
// in progress reads, in one load of the first four control words.
static_assert(BC_STEP == 0 && BC_NDUE0 == 1 && BC_NDUE1 == 2, "control word layout");
struct BatchClock {
  int step, ndue;
};
template <class H>
__device__ __forceinline__ BatchClock batch_clock(const BBufs<H>& B) {
  const int4 c = *reinterpret_cast<const int4*>(B.ctl);
  const int step = __builtin_amdgcn_readfirstlane(c.x);
  return {step, __builtin_amdgcn_readfirstlane(rebuild_parity(step) ? c.z : c.y)};
}

// The rebuild kernels run on a fixed grid and visit the waves of the due
// scenes only.  The due scenes' waves, concatenated in list order, are the
// work items; item g goes to wave g mod (waves of the grid).  Each workgroup
// takes the list DUE_CHUNK entries at a time, with their first waves, wave
// counts and prefix sums in LDS, and finds an item's entry by binary search
// (a scrubbed entry has no waves).  f(sid, slot) runs with the whole wave
// active and must not synchronise the workgroup.
constexpr int DUE_CHUNK = 256;
struct DueLds {
  int sid[DUE_CHUNK];
  int w0[DUE_CHUNK];
  int pre[DUE_CHUNK + 1];  // waves of the chunk's entries before entry j
};

template <class H, class F>
__device__ __forceinline__ void for_due_waves(const BBufs<H>& B, const BatchClock& clk, DueLds& d, F&& f) {
  constexpr int PER = DUE_CHUNK / 64;
  const int4* list = B.due + (size_t)rebuild_parity(clk.step) * B.nslots;
  const int nd = clk.ndue;
  const int tid = threadIdx.x, lane = tid & 63;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (tid >> 6);
  const int GW = gridDim.x * (blockDim.x >> 6);
  int before = 0;  // items of the earlier chunks
  for (int j0 = 0; j0 < nd; j0 += DUE_CHUNK) {
    if (tid < 64) {
      int w[PER], sum = 0;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int j = j0 + tid * PER + u;
        const int4 e = j < nd ? list[j] : make_int4(-1, 0, 0, 0);
        d.sid[tid * PER + u] = e.x;
        d.w0[tid * PER + u] = e.y;
        w[u] = e.z;
        sum += e.z;
      }
      int run = wave_incl_sum(sum) - sum;
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        d.pre[tid * PER + u] = run;
        run += w[u];
      }
      if (tid == 63) d.pre[DUE_CHUNK] = run;
    }
    __syncthreads();
    const int tot = d.pre[DUE_CHUNK];
    for (int g = before + ((gw - before) % GW + GW) % GW; g < before + tot; g += GW) {
      const int local = g - before;
      int lo = 0, hi = DUE_CHUNK;  // d.pre[lo] <= local < d.pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d.pre[mid] <= local) lo = mid;
        else hi = mid;
      }
      const int sid = __builtin_amdgcn_readfirstlane(d.sid[lo]);
      const int w0 = __builtin_amdgcn_readfirstlane(d.w0[lo]);
      f(sid, (w0 + local - d.pre[lo]) * 64 + lane);
    }
    before += tot;
    __syncthreads();  // every wave is done with d before the next chunk
  }
}

// Grids of the rebuild kernels (they walk the due waves, so a fixed grid
// serves any number of due scenes).
constexpr int RB_GRID = 1024;
constexpr int RB_SCAN_GRID = 512;

// Cell of every slot of a due scene (d_calcCellId, BoundingBox.cu:1089-1104)
// and its rank among the particles of that cell.  Inactive particles, and any
// position outside the grid, are ranked among the scene's inactive ones.
template <class H>
__global__ __launch_bounds__(256) void kb_bin(BBufs<H> B) {
  const BatchClock clk = batch_clock(B);
  if (clk.ndue == 0) return;
  const int step = clk.step;
  __shared__ DueLds due;
  for_due_waves(B, clk, due, [&](const int sid, const int i) {
    const BScene& S = B.scenes[sid];
    SlotCtl& C = B.sctl[sid];
    const int cur = (step & 1) ^ __builtin_amdgcn_readfirstlane(C.q);
    const double4 p = B.pos[cur][i];
    bool in = p.w >= 0.;
    int c = -1;
    if (in) {
      const double fx = floor((p.x - S.bmin[0]) * S.invd[0]) + 1;
      const double fy = floor((p.y - S.bmin[1]) * S.invd[1]) + 1;
      const double fz = floor((p.z - S.bmin[2]) * S.invd[2]) + 1;
      in = fx >= 0. && fx < (double)S.size[0] && fy >= 0. && fy < (double)S.size[1] && fz >= 0. &&
           fz < (double)S.size[2];
      if (in) c = S.cell_base + (S.size[1] * (int)fz + (int)fy) * S.size[0] + (int)fx;
    }
    if (in) {
      B.bin_cell[i] = c;
      B.bin_rank[i] = atomicAdd(&B.ccount[c], 1);
    } else {
      B.bin_cell[i] = -1;
      B.bin_rank[i] = atomicAdd(&C.inact, 1);
    }
    // this wave's particles per chunk of the scene's cells (each workgroup of
    // the scan starts from the particles of the chunks before its own)
    const int cc = (S.ncell + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
    const int ch = in ? (c - S.cell_base) / cc : -1;
#pragma unroll
    for (int k = 0; k < SCAN_CHUNKS; ++k) {
      const int m = __popcll(__ballot(ch == k));
      if (m && __lane_id() == 0) atomicAdd(&C.chunk[k], m);
    }
  });
}

// SCAN_CHUNKS workgroup tasks per due scene, each a contiguous chunk of its
// cells: exclusive scan of the cell counts into the (first, end) slot table,
// starting from the particles the binning pass counted in the chunks before
// (no task waits for another), clearing the counters for the next rebuild.
// The last chunk's task records the scene's active count.  The tasks of
// the due scenes are spread over a fixed grid.
constexpr int SCAN_PER = 16;

template <class H>
__global__ __launch_bounds__(256) void kb_scan(BBufs<H> B) {
  const BatchClock clk = batch_clock(B);
  const int nd = clk.ndue;
  const int4* list = B.due + (size_t)rebuild_parity(clk.step) * B.nslots;
  // per round: each thread sums SCAN_PER counts, each wave scans its 64 sums
  // with lane shuffles, and the four wave totals are combined through LDS
  // (two workgroup barriers per round)
  __shared__ int part[4];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  for (int task = blockIdx.x; task < nd * SCAN_CHUNKS; task += gridDim.x) {
    const int4 e = list[task / SCAN_CHUNKS];
    if (__builtin_amdgcn_readfirstlane(e.z) == 0) continue;  // scrubbed: the scene was evicted
    const int sid = __builtin_amdgcn_readfirstlane(e.x);
    const int part_k = task % SCAN_CHUNKS;
    SlotCtl& C = B.sctl[sid];
    const BScene& S = B.scenes[sid];
    const int base = S.cell_base, n = S.ncell;
    const int cc = (n + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
    const int lo = part_k * cc, hi = min(n, lo + cc);
    int carry = 0;
    for (int k = 0; k < part_k; ++k) carry += C.chunk[k];
    // the next round's counts are requested before this round's stores, so
    // their latency overlaps the scan
    int v[SCAN_PER];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      const int x = lo + tid * SCAN_PER + k;
      v[k] = x < hi ? B.ccount[base + x] : 0;
    }
    for (int c0 = lo; c0 < hi; c0 += 256 * SCAN_PER) {
      const int beg = c0 + tid * SCAN_PER;
      const int nbeg = beg + 256 * SCAN_PER;
      int vn[SCAN_PER];
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k) vn[k] = (nbeg + k < hi) ? B.ccount[base + nbeg + k] : 0;
      int sum = 0;
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k) sum += v[k];
      const int incl = wave_incl_sum(sum);
      if (lane == 63) part[wave] = incl;
      __syncthreads();
      int before = 0, total = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int t = part[w];
        if (w < wave) before += t;
        total += t;
      }
      int run = S.seg0 + carry + before + incl - sum;
      __syncthreads();  // every wave has read part[] before the next round writes it
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k) {
        if (beg + k < hi) {
          B.cell_tab[base + beg + k] = make_int2(run, run + v[k]);
          if (v[k]) B.ccount[base + beg + k] = 0;
          run += v[k];
        }
      }
      carry += total;
#pragma unroll
      for (int k = 0; k < SCAN_PER; ++k) v[k] = vn[k];
    }
    if (tid == 0 && part_k == SCAN_CHUNKS - 1) {
      C.nact = carry;
      C.inact = 0;
      C.q2 = C.q ^ 1;
    }
  }
}

// Scatter every slot of a due scene to its cell-ordered position in the other
// buffer half: state, material, ids and contact histories (d_swap_ps,
// BoundingBox.cu:928-1005, here including history rows, stiffness and ids).
template <class H>
__global__ __launch_bounds__(256) void kb_permute(BBufs<H> B) {
  const BatchClock clk = batch_clock(B);
  if (clk.ndue == 0) return;
  const int step = clk.step;
  __shared__ DueLds due;
  for_due_waves(B, clk, due, [&](const int sid, const int i) {
    const BScene& S = B.scenes[sid];
    SlotCtl& C = B.sctl[sid];
    const int q = __builtin_amdgcn_readfirstlane(C.q);
    const int cur = (step & 1) ^ q;
    // new slot of slot j of this scene
    auto dest = [&](int j) {
      const int c = B.bin_cell[j];
      const int r = B.bin_rank[j];
      return c >= 0 ? B.cell_tab[c].x + r : S.seg0 + C.nact + r;
    };
    const int c = B.bin_cell[i];
    const int d = dest(i);
    if (d < S.seg0 || d >= S.seg0 + S.nseg) {
      atomicOr(&C.err, GDEM_ERR_NB_OVERFLOW_BIT);  // inconsistent counts: never write outside the segment
      return;
    }
    const size_t N = (size_t)B.cap;
    B.pos[cur ^ 1][d] = B.pos[cur][i];
    B.vel[cur ^ 1][d] = B.vel[cur][i];
    B.angv[cur ^ 1][d] = B.angv[cur][i];
    B.mat[q ^ 1][d] = B.mat[q][i];
    B.pid[q ^ 1][d] = B.pid[q][i];
    B.cell_of[d] = c;
    const int nh = B.nh[q][i];
    B.nh[q ^ 1][d] = nh;
    for (int k = 0; k < nh; ++k) {
      const int j = B.hkey[q][k * N + i];  // the partner's old slot, in this scene
      B.hkey[q ^ 1][k * N + d] = (j >= S.seg0 && j < S.seg0 + S.nseg) ? dest(j) : -1;
      B.hd[q ^ 1][k * N + d] = B.hd[q][k * N + i];
    }
    const int nw = B.nw[q][i];
    B.nw[q ^ 1][d] = nw;
    for (int k = 0; k < nw; ++k) {
      B.wkey[q ^ 1][k * N + d] = B.wkey[q][k * N + i];
      B.wd[q ^ 1][k * N + d] = B.wd[q][k * N + i];
    }
  });
}

// Pair and wall candidate lists of the permuted scene (the single-scene list
// builders with this scene's grid), then the scene's buffer parity flips.
template <class H, bool LONG>
__global__ __launch_bounds__(PL_BLOCK) void kb_lists(BBufs<H> B) {
  __shared__ PairLds lds;
  const BatchClock clk = batch_clock(B);
  if (clk.ndue == 0) return;
  const int step = clk.step;
  __shared__ DueLds due;
  for_due_waves(B, clk, due, [&](const int sid, const int i) {
    const BScene& S = B.scenes[sid];
    SlotCtl& C = B.sctl[sid];
    const int qn = __builtin_amdgcn_readfirstlane(C.q2);
    const int cur = (step & 1) ^ qn;
    ListArgs L;
    L.pos = B.pos[cur];
    L.pid = B.pid[qn];
    L.cell_of = B.cell_of;
    L.cell_tab = B.cell_tab;
    L.nbr = B.nbr;
    L.nnb = B.nnb;
    L.ref = B.ref;
    L.angv = B.angv[cur];
    L.wl = B.wl;
    L.nwl = B.nwl;
    L.tri = B.tri;
    L.bvh = nullptr;
    L.tbox = B.tbox;
    L.wc_off = B.wc_off;
    L.wc_tri = B.wc_tri;
    L.err = &C.err;
    L.N = B.cap;
    for (int a = 0; a < 3; ++a) L.size[a] = S.size[a];
    L.cell_base = S.cell_base;
    L.skin = S.skin;
    for (int a = 0; a < 3; ++a) {
      L.qlo[a] = S.bmin[a];
      L.qhi[a] = S.bmax[a];
    }
    double qeps;
    refq_set(L.qlo, L.qhi, L.qh, qeps);
    L.nb_cap = B.nb_cap;
    L.wall_cap = B.wall_cap;
    L.ntri = S.ntri;
    L.wall_table = 1;
    const bool active = (i - S.seg0) < C.nact;
    if (i == S.seg0)
      for (int k = 0; k < SCAN_CHUNKS; ++k) C.chunk[k] = 0;  // read by this rebuild's scan only
    pairlist_one<PL_BLOCK, LONG>(L, i, active, lds);
    walllist_one<false>(L, i, active);  // batch scenes always have the per-cell table
    // every later reader of q runs in a later kernel; this kernel reads q2 only
    if (__lane_id() == 0) C.q = qn;
    if (i == S.seg0) C.rebuilds += 1;
  });
}

// Admission of a scene: raises its refresh flag of parity par and lists it,
// in a scrubbed entry if there is one, else at the end.  One wave, in stream
// order after every earlier batch step.
template <class H>
__global__ __launch_bounds__(64) void kb_admit(BBufs<H> B, int sid, int par) {
  const int lane = threadIdx.x;
  int4* list = B.due + (size_t)par * B.nslots;
  const int n = B.ctl[BC_NDUE0 + par];
  int at = n;  // first entry of this scene slot or scrubbed
  for (int j = lane; j < n; j += 64) {
    const int x = list[j].x;
    if (x == sid || x < 0) at = min(at, j);
  }
  for (int off = 32; off > 0; off >>= 1) at = min(at, __shfl_xor(at, off));
  if (lane == 0) {
    const BScene& S = B.scenes[sid];
    list[at] = make_int4(sid, S.seg0 >> 6, S.nseg >> 6, 0);
    if (at == n) B.ctl[BC_NDUE0 + par] = n + 1;
    B.sctl[sid].fl[par] = 1;
  }
}

// Eviction: scrubs the scene's entries from both due lists, so that a
// rebuild it requested at its last step never runs.
template <class H>
__global__ __launch_bounds__(64) void kb_evict(BBufs<H> B, int sid) {
  const int lane = threadIdx.x;
  for (int par = 0; par < 2; ++par) {
    int4* list = B.due + (size_t)par * B.nslots;
    const int n = B.ctl[BC_NDUE0 + par];
    for (int j = lane; j < n; j += 64)
      if (list[j].x == sid) list[j] = make_int4(-1, 0, 0, 0);
  }
}

// Fused contact step of every running scene.
template <class T, class H, class PRow, class WRow>
__global__ __launch_bounds__(256, step_waves<T>()) void kb_step(BBufs<H> B) {
  __shared__ StepLds<T, RowCap<PRow>::value, RowCap<WRow>::value> lds;
  const int step = __builtin_amdgcn_readfirstlane(B.ctl[BC_STEP]);
  const int s = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int sid = __builtin_amdgcn_readfirstlane(B.wave_scene[__builtin_amdgcn_readfirstlane(s >> 6)]);
  uint32_t err = 0;
  bool refresh = false;
  // this workgroup's positions, for in-tile neighbour reads; a neighbour
  // belongs to its particle's scene, whose waves stage from the same buffer
  __shared__ double4 tile_p[256];
  bool running = false;
  StepArgs<H> A{};
  if (sid >= 0) {
    const BScene& S = B.scenes[sid];
    running = step >= S.start_step && step < S.end_step;
    const int q = __builtin_amdgcn_readfirstlane(B.sctl[sid].q);
    const int cur = (step & 1) ^ q;
    A.pin = B.pos[cur];
    A.vin = B.vel[cur];
    A.win = B.angv[cur];
    A.pout = B.pos[cur ^ 1];
    A.vout = B.vel[cur ^ 1];
    A.wout = B.angv[cur ^ 1];
    A.mat = B.mat[q];
    A.hkey = B.hkey[q];
    A.hd = B.hd[q];
    A.nh = B.nh[q];
    A.wkey = B.wkey[q];
    A.wd = B.wd[q];
    A.nw = B.nw[q];
    A.ref = B.ref;
    A.pid = B.pid[q];
    A.nbr = B.nbr;
    A.nnb = B.nnb;
    A.wl = B.wl;
    A.nwl = B.nwl;
    A.tri = B.tri;
    A.N = B.cap;
    A.nb_cap = B.nb_cap;
    A.hist_cap = B.hist_cap;
    A.whist_cap = B.whist_cap;
    A.dt = S.dt;
    A.mu = S.mu;
    for (int a = 0; a < 3; ++a) {
      A.g[a] = S.g[a];
      A.bmin[a] = S.bmin[a];
      A.bmax[a] = S.bmax[a];
    }
    A.threshSq = S.threshSq;
    refq_set(A.bmin, A.bmax, A.qh, A.qeps);
  }
  const Own<H> own = load_own<false>(A, s, running);
  tile_p[threadIdx.x] = own.p;
  stage_vw<T>(lds, threadIdx.x, own.v);
  __syncthreads();
  if (sid >= 0) {
    const BScene& S = B.scenes[sid];
    SlotCtl& C = B.sctl[sid];
    if (running) {
      Rec<T> none{};
      step_particle<T, H, false, false, PRow, WRow>(A, s, true, own, tile_p, s - (int)threadIdx.x, lds, none, err,
                                             refresh);
      // a refresh at the scene's last step would only move its final state
      if (__ballot(refresh) != 0 && __lane_id() == 0 && step + 1 < S.end_step) {
        // the first wave of the scene to raise the flag lists the scene
        if (atomicExch(&C.fl[step & 1], 1) == 0) {
          const int k = atomicAdd(&B.ctl[BC_NDUE0 + (step & 1)], 1);
          B.due[(size_t)(step & 1) * B.nslots + k] = make_int4(sid, S.seg0 >> 6, S.nseg >> 6, 0);
        }
      }
      if (err) atomicOr(&C.err, err);
    }
    // the flag the preceding rebuild consumed
    if ((s >> 6) == (S.seg0 >> 6) && __lane_id() == 0) C.fl[rebuild_parity(step)] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) B.ctl[BC_NDUE0 + rebuild_parity(step)] = 0;
  if (last_block(&B.ctl[BC_DONE])) {
    B.ctl[BC_DONE] = 0;
    atomicExch(&B.ctl[BC_STEP], step + 1);
  }
}

// State of one scene in its particles' own order (ids within the scene).
template <class H>
__global__ __launch_bounds__(256) void kb_download(BBufs<H> B, int sid, double* x, double* v, double* w,
                                                   int* act) {
  const BScene& S = B.scenes[sid];
  const int i = S.seg0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.seg0 + S.nseg) return;
  const int step = B.ctl[BC_STEP];
  const int t = step < S.end_step ? step : S.end_step;
  const int q = B.sctl[sid].q;
  const int cur = (t & 1) ^ q;
  const int id = B.pid[q][i];
  if (id < 0 || id >= S.n) return;
  const double4 p = B.pos[cur][i], vv = B.vel[cur][i], ww = B.angv[cur][i];
  x[id * 3] = p.x; x[id * 3 + 1] = p.y; x[id * 3 + 2] = p.z;
  v[id * 3] = vv.x; v[id * 3 + 1] = vv.y; v[id * 3 + 2] = vv.z;
  w[id * 3] = ww.x; w[id * 3 + 1] = ww.y; w[id * 3 + 2] = ww.z;
  act[id] = p.w < 0. ? 0 : 1;
}

// Per-scene list/contact sums for statistics: active, pair candidates, wall
// candidates, live history slots.
template <class H>
__global__ __launch_bounds__(256) void kb_counts(BBufs<H> B, int sid, unsigned long long* acc) {
  const BScene& S = B.scenes[sid];
  const int i = S.seg0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int step = B.ctl[BC_STEP];
  const int t = step < S.end_step ? step : S.end_step;
  const int q = B.sctl[sid].q;
  const int cur = (t & 1) ^ q;
  unsigned long long a = 0, b = 0, c = 0, d = 0;
  if (i < S.seg0 + S.nseg && B.pos[cur][i].w >= 0.) {
    a = 1;
    b = (unsigned long long)B.nnb[i];
    c = (unsigned long long)B.nwl[i];
    d = (unsigned long long)(B.nh[q][i] + B.nw[q][i]);
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
// launchers
// ---------------------------------------------------------------------------
constexpr int BPCAP_LDS = 16;
constexpr int BWCAP_LDS = 8;

template <class T, class H>
void launch_batch_step(const BBufs<H>& B, int grid, hipStream_t st) {
  if (grid <= 0) return;
  const int rg = std::min(grid, RB_GRID), sg = std::min(B.nslots * SCAN_CHUNKS, RB_SCAN_GRID);
  kb_bin<H><<<rg, 256, 0, st>>>(B);
  kb_scan<H><<<sg, 256, 0, st>>>(B);
  kb_permute<H><<<rg, 256, 0, st>>>(B);
  if (B.nb_cap <= PL_LDS_CAP) kb_lists<H, false><<<rg * (256 / PL_BLOCK), PL_BLOCK, 0, st>>>(B);
  else kb_lists<H, true><<<rg * (256 / PL_BLOCK), PL_BLOCK, 0, st>>>(B);
  if (hist_lds_fits(B.cap, B.hist_cap, B.whist_cap, BPCAP_LDS, BWCAP_LDS))
    kb_step<T, H, HistLds<H, BPCAP_LDS>, HistLds<H, BWCAP_LDS>><<<grid, 256, 0, st>>>(B);
  else
    kb_step<T, H, HistScan<H>, HistScan<H>><<<grid, 256, 0, st>>>(B);
}

template <class H>
void launch_batch_admit(const BBufs<H>& B, int sid, int par, hipStream_t st) {
  kb_admit<H><<<1, 64, 0, st>>>(B, sid, par);
}

template <class H>
void launch_batch_evict(const BBufs<H>& B, int sid, hipStream_t st) {
  kb_evict<H><<<1, 64, 0, st>>>(B, sid);
}

template <class H>
void launch_batch_download(const BBufs<H>& B, int sid, int nseg, double* x, double* v, double* w, int* act,
                           hipStream_t st) {
  kb_download<H><<<(nseg + 255) / 256, 256, 0, st>>>(B, sid, x, v, w, act);
}

template <class H>
void launch_batch_counts(const BBufs<H>& B, int sid, int nseg, unsigned long long* acc, hipStream_t st) {
  kb_counts<H><<<(nseg + 255) / 256, 256, 0, st>>>(B, sid, acc);
}

template void launch_batch_step<double, double>(const BBufs<double>&, int, hipStream_t);
template void launch_batch_step<float, float>(const BBufs<float>&, int, hipStream_t);
template void launch_batch_admit<double>(const BBufs<double>&, int, int, hipStream_t);
template void launch_batch_admit<float>(const BBufs<float>&, int, int, hipStream_t);
template void launch_batch_evict<double>(const BBufs<double>&, int, hipStream_t);
template void launch_batch_evict<float>(const BBufs<float>&, int, hipStream_t);
template void launch_batch_download<double>(const BBufs<double>&, int, int, double*, double*, double*, int*,
                                            hipStream_t);
template void launch_batch_download<float>(const BBufs<float>&, int, int, double*, double*, double*, int*,
                                           hipStream_t);
template void launch_batch_counts<double>(const BBufs<double>&, int, int, unsigned long long*, hipStream_t);
template void launch_batch_counts<float>(const BBufs<float>&, int, int, unsigned long long*, hipStream_t);

}  // namespace gdem
