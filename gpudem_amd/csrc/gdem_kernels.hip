// gfx950 kernels of the single-scene gdem contact engine.
//
// One thread owns one particle (internal Morton order).  Particle state is
// double-buffered (read buffer `cur`, write buffer `cur^1`), so forces and
// integration run in one fused pass and no force/torque arrays or per-step
// memsets exist.  Candidate lists and contact histories are slot-major
// ([slot][N]) so lane i of a wave reads consecutive words.  A list rebuild is
// decided on the device (refresh flag) and every rebuild kernel exits at once
// when the flag is clear, so a step needs no host synchronisation.
#include <hip/hip_runtime.h>

#include "gdem_device.h"
#include "gdem_kernels.h"

namespace gdem {

// ---------------------------------------------------------------------------
// fused contact forces + integration for one step
// ---------------------------------------------------------------------------
template <class T, class H, bool REC, bool PRE, class PRow, class WRow>
__global__ __launch_bounds__(256, step_waves<T>()) void k_step(SimParams P, Bufs<H> B, Rec<T> rec) {
  __shared__ StepLds<T, RowCap<PRow>::value, RowCap<WRow>::value> lds;
  // thread tt of the workgroup holds slot tt of the tile
  const int tile0 = xcd_tile(blockIdx.x, gridDim.x) * 256;
  const int tt = threadIdx.x;
  const int s = tile0 + tt;
  const int cur = __builtin_amdgcn_readfirstlane(B.ctl[CTL_CUR]);
  const int hc = __builtin_amdgcn_readfirstlane(B.ctl[CTL_HCUR]);
  StepArgs<H> A;
  A.pin = B.pos[cur];
  A.vin = B.vel[cur];
  A.win = B.angv[cur];
  A.pout = B.pos[cur ^ 1];
  A.vout = B.vel[cur ^ 1];
  A.wout = B.angv[cur ^ 1];
  A.mat = B.mat[hc];
  A.hkey = B.hkey[hc];
  A.hd = B.hd[hc];
  A.nh = B.nh[hc];
  A.wkey = B.wkey[hc];
  A.wd = B.wd[hc];
  A.nw = B.nw[hc];
  A.ref = B.ref;
  A.pid = B.pid[hc];
  A.nbr = B.nbr;
  A.nnb = B.nnb;
  A.wl = B.wl;
  A.nwl = B.nwl;
  A.tri = B.tri;
  A.N = P.N;
  A.nb_cap = P.nb_cap;
  A.hist_cap = P.hist_cap;
  A.whist_cap = P.whist_cap;
  A.dt = P.dt;
  A.mu = P.mu;
  for (int a = 0; a < 3; ++a) {
    A.g[a] = P.g[a];
    A.bmin[a] = P.bmin[a];
    A.bmax[a] = P.bmax[a];
  }
  A.threshSq = P.threshSq;
  refq_set(A.bmin, A.bmax, A.qh, A.qeps);
  // this workgroup's positions, for in-tile neighbour reads
  __shared__ double4 tile_p[256];
  const bool valid = s < P.N;
  const Own<H> own = load_own<PRE>(A, s, valid);
  tile_p[tt] = own.p;
  stage_vw<T>(lds, tt, own.v);
  __syncthreads();
  uint32_t err = 0;
  bool refresh = false;
  step_particle<T, H, REC, PRE, PRow, WRow>(A, s, valid, own, tile_p, tile0, lds, rec, err, refresh);
  if (__ballot(refresh) != 0 && __lane_id() == 0) B.ctl[CTL_REFRESH] = 1;
  if (err) atomicOr((unsigned int*)&B.ctl[CTL_ERR], err);
  if (last_block(&B.ctl[CTL_DONE_STEP])) {
    B.ctl[CTL_DONE_STEP] = 0;
    atomicExch(&B.ctl[CTL_CUR], cur ^ 1);
  }
}

// ---------------------------------------------------------------------------
// list rebuild (each phase is a per-element device function used both by the
// launch chain and by the single grid-barrier rebuild kernel)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rebuild_due(const int* ctl) {
  return __builtin_amdgcn_readfirstlane(ctl[CTL_REFRESH]) != 0;
}

// Cell coordinates (d_calcCellId, BoundingBox.cu:1089-1104) and sort keys.
// Also clears the table entry of the cell this particle occupied at the last
// rebuild, so the new table holds exactly the occupied cells.
template <class H>
__device__ __forceinline__ uint32_t keys_one(const SimParams& P, const Bufs<H>& B, int s, int cur) {
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
  return key;
}

// ---- stable LSD radix sort, 11-bit digits ---------------------------------
// Workgroup b owns the contiguous chunk [b*chunk, (b+1)*chunk); per-(digit,
// chunk) counts are scanned digit-major so equal digits keep chunk order, and
// inside a chunk keys are ranked in arrival order with wave ballots.
constexpr int RS_THREADS = 256;
constexpr int RS_PER = RS_DIGITS / RS_THREADS;  // digits per thread
constexpr int RS_WAVES = RS_THREADS / 64;

__device__ __forceinline__ uint32_t rs_digit(uint32_t k, int pass) {
  return (k >> (RS_BITS * pass)) & (uint32_t)(RS_DIGITS - 1);
}

// cnt (LDS, RS_DIGITS words) -> hist[d * nb + b]
__device__ __forceinline__ void rs_clear(uint32_t* cnt) {
#pragma unroll
  for (int q = 0; q < RS_PER; ++q) cnt[q * RS_THREADS + threadIdx.x] = 0;
  __syncthreads();
}
__device__ __forceinline__ void rs_store_counts(const uint32_t* cnt, uint32_t* __restrict__ hist, int b, int nb) {
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RS_PER; ++q) {
    const int d = q * RS_THREADS + threadIdx.x;
    hist[(size_t)d * nb + b] = cnt[d];
  }
  __syncthreads();
}

__device__ __forceinline__ void rs_count_block(int n, int chunk, int pass, const uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ hist, int b, int nb, uint32_t* cnt) {
  rs_clear(cnt);
  const int beg = b * chunk;
  const int end = min(n, beg + chunk);
  for (int i = beg + threadIdx.x; i < end; i += RS_THREADS) atomicAdd(&cnt[rs_digit(keys[i], pass)], 1u);
  rs_store_counts(cnt, hist, b, nb);
}

// Exclusive scan, in place, of each digit's counts over the nb chunks
// (hist[d * nb + j]), with the digit totals to dtot[d].  Workgroup b of gs
// takes digits b, b + gs, ...: one thread per digit when nb is small, one
// wave per digit otherwise (no workgroup barriers either way).
__device__ __forceinline__ void rs_scan_digits(uint32_t* __restrict__ hist, uint32_t* __restrict__ dtot, int nb,
                                               int b, int gs) {
  if (nb <= 64) {
    for (int d = b * RS_THREADS + threadIdx.x; d < RS_DIGITS; d += gs * RS_THREADS) {
      uint32_t* h = hist + (size_t)d * nb;
      uint32_t run = 0;
      for (int j = 0; j < nb; ++j) {
        const uint32_t c = h[j];
        h[j] = run;
        run += c;
      }
      dtot[d] = run;
    }
  } else {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = b * RS_WAVES + wave; d < RS_DIGITS; d += gs * RS_WAVES) {
      uint32_t* h = hist + (size_t)d * nb;
      int run = 0;
      for (int j0 = 0; j0 < nb; j0 += 64) {
        const int j = j0 + lane;
        const int c = j < nb ? (int)h[j] : 0;
        const int incl = wave_incl_sum(c);
        if (j < nb) h[j] = (uint32_t)(run + incl - c);
        run += __shfl(incl, 63);
      }
      if (lane == 0) dtot[d] = (uint32_t)run;
    }
  }
}

// Scatter bases of workgroup b: base[d] (LDS) = keys of smaller digits in the
// whole array + keys of digit d in earlier chunks (the scanned hist entry).
__device__ __forceinline__ void rs_bases(const uint32_t* __restrict__ hist, const uint32_t* __restrict__ dtot,
                                         int b, int nb, uint32_t* base) {
  __shared__ uint32_t part[RS_THREADS];
  uint32_t v[RS_PER];
  uint32_t sum = 0;
#pragma unroll
  for (int q = 0; q < RS_PER; ++q) {
    v[q] = dtot[threadIdx.x * RS_PER + q];
    sum += v[q];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < RS_THREADS; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
#pragma unroll
  for (int q = 0; q < RS_PER; ++q) {
    const int d = threadIdx.x * RS_PER + q;
    base[d] = run + hist[(size_t)d * nb + b];
    run += v[q];
  }
  __syncthreads();
}

// Stable scatter of workgroup b's chunk from the bases in base (LDS).  Each
// round ranks 256 keys: within a wave by ballots over the digit bits, across
// waves through per-wave digit counts in wcnt (LDS, RS_WAVES x RS_DIGITS
// 16-bit words: round tag << 8 | count), so the counts are cleared only when
// the 8-bit tag wraps.
__device__ __forceinline__ void rs_clear_wcnt(uint16_t* wcnt) {
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RS_WAVES * RS_PER; ++q) wcnt[q * RS_THREADS + threadIdx.x] = 0;
  __syncthreads();
}

// inv (the last pass of a rebuild's sort, else null): inv[source index] =
// its sorted position, the inverse of the permutation in iout
__device__ __forceinline__ void rs_scatter_block(int n, int chunk, int pass, const uint32_t* __restrict__ kin,
                                                 const int* __restrict__ iin, uint32_t* __restrict__ kout,
                                                 int* __restrict__ iout, int* __restrict__ inv, uint32_t* base,
                                                 uint16_t* wcnt, int b) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int beg = b * chunk;
  const int end = min(n, beg + chunk);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t round = 255;
  for (int r0 = beg; r0 < end; r0 += RS_THREADS) {
    // tags run 1..255; cleared words (tag 0) never match
    if (++round == 256) {
      rs_clear_wcnt(wcnt);
      round = 1;
    }
    const int i = r0 + tid;
    const bool valid = i < end;
    const uint32_t k = valid ? kin[i] : 0u;
    const int idx = valid ? iin[i] : 0;
    const uint32_t d = rs_digit(k, pass);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < RS_BITS; ++bit) {
      const uint64_t bal = __ballot((d >> bit) & 1u);
      match &= ((d >> bit) & 1u) ? bal : ~bal;
    }
    const int rank = __popcll(match & lt);
    const uint32_t cw = (uint32_t)__popcll(match);
    if (valid && rank == 0) wcnt[wave * RS_DIGITS + d] = (uint16_t)((round << 8) | cw);
    __syncthreads();
    uint32_t before = 0;
    bool last_wave = true;  // no later wave holds digit d this round
    uint32_t bd = 0;
    if (valid) {
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) {
        const uint32_t e = wcnt[w * RS_DIGITS + d];  // cw <= 64 fits the low byte
        const bool live = (e >> 8) == round;
        if (w < wave && live) before += e & 0xffu;
        if (w > wave && live) last_wave = false;
      }
      bd = base[d];
    }
    __syncthreads();  // every base read before the updates below
    if (valid) {
      const uint32_t pos = bd + before + (uint32_t)rank;
      if (pos < (uint32_t)n) {  // always, unless a grid barrier failed
        kout[pos] = k;
        iout[pos] = idx;
        if (inv && (unsigned)idx < (unsigned)n) inv[idx] = (int)pos;
      }
      if (rank == 0 && last_wave) base[d] = bd + before + cw;
    }
    __syncthreads();
  }
}

// Gather every per-particle array into sorted order (d_swap_ps,
// BoundingBox.cu:928-1005, but including history rows, stiffness and ids).
template <class H>
__device__ __forceinline__ void permute_one(const SimParams& P, const Bufs<H>& B, int s, int cur, int hc) {
  const int N = P.N;
  const int src = B.idx[RS_FINAL][s];
  if ((unsigned)src >= (unsigned)N) return;  // always in range, unless a grid barrier failed
  B.pos[cur ^ 1][s] = B.pos[cur][src];
  B.vel[cur ^ 1][s] = B.vel[cur][src];
  B.angv[cur ^ 1][s] = B.angv[cur][src];
  B.mat[hc ^ 1][s] = B.mat[hc][src];
  B.pid[hc ^ 1][s] = B.pid[hc][src];
  B.cell_of[s] = B.cell_tmp[src];
  const int nh = B.nh[hc][src];
  B.nh[hc ^ 1][s] = nh;
  for (int q = 0; q < nh; ++q) {
    const int k = B.hkey[hc][(size_t)q * N + src];  // the partner's old slot
    B.hkey[hc ^ 1][(size_t)q * N + s] = (unsigned)k < (unsigned)N ? B.inv[k] : -1;
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
  const uint32_t* key = B.key[RS_FINAL];
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

// ---- the launch chain (one kernel per phase; used under hipGraph capture) ----
template <class H>
__global__ __launch_bounds__(256) void k_keys(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < P.N) keys_one(P, B, s, B.ctl[CTL_CUR]);
  if (s == 0) *reinterpret_cast<unsigned long long*>(&B.ctl[CTL_NBSUM]) = 0ull;
}

template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_count(int n, int chunk, int pass,
                                                         const uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ hist, const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  __shared__ uint32_t cnt[RS_DIGITS];
  rs_count_block(n, chunk, pass, keys, hist, blockIdx.x, gridDim.x, cnt);
}

// The digits' counts over the nb chunks, scanned; the totals go to
// hist[RS_DIGITS * nb + d].
constexpr int RS_SCAN_BLOCKS = RS_DIGITS / RS_WAVES;

template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan(int nb, uint32_t* __restrict__ hist, const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  rs_scan_digits(hist, hist + (size_t)RS_DIGITS * nb, nb, blockIdx.x, gridDim.x);
}

template <bool GATED>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(int n, int chunk, int pass,
                                                           const uint32_t* __restrict__ kin,
                                                           const int* __restrict__ iin,
                                                           uint32_t* __restrict__ kout,
                                                           int* __restrict__ iout,
                                                           int* __restrict__ inv,
                                                           const uint32_t* __restrict__ hist,
                                                           const int* ctl) {
  if (GATED && !rebuild_due(ctl)) return;
  __shared__ uint32_t base[RS_DIGITS];
  __shared__ uint16_t wcnt[RS_WAVES * RS_DIGITS];
  const int nb = gridDim.x;
  rs_bases(hist, hist + (size_t)RS_DIGITS * nb, blockIdx.x, nb, base);
  rs_scatter_block(n, chunk, pass, kin, iin, kout, iout, inv, base, wcnt, blockIdx.x);
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

// Arguments of the list builders for state buffer cur and material/id
// buffer hc.
template <class H>
__device__ __forceinline__ ListArgs list_args(const SimParams& P, const Bufs<H>& B, int cur, int hc) {
  ListArgs L;
  L.pos = B.pos[cur];
  L.pid = B.pid[hc];
  L.cell_of = B.cell_of;
  L.cell_tab = B.cell_tab;
  L.nbr = B.nbr;
  L.nnb = B.nnb;
  L.ref = B.ref;
  L.angv = B.angv[cur];
  L.wl = B.wl;
  L.nwl = B.nwl;
  L.tri = B.tri;
  L.bvh = B.bvh;
  L.tbox = B.tbox;
  L.wc_off = B.wc_off;
  L.wc_tri = B.wc_tri;
  L.err = (unsigned int*)&B.ctl[CTL_ERR];
  L.N = P.N;
  for (int a = 0; a < 3; ++a) L.size[a] = P.size[a];
  L.cell_base = 0;
  L.skin = P.skin;
  for (int a = 0; a < 3; ++a) {
    L.qlo[a] = P.bmin[a];
    L.qhi[a] = P.bmax[a];
  }
  double qeps;
  refq_set(L.qlo, L.qhi, L.qh, qeps);
  L.nb_cap = P.nb_cap;
  L.wall_cap = P.wall_cap;
  L.ntri = P.ntri;
  L.wall_table = P.wall_table;
  return L;
}

// Cell table, and the wall list of each slot (it needs only the slot's own
// permuted record and the static triangle tables).
template <class H, bool BVH>
__global__ __launch_bounds__(256) void k_cells(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < P.N) {
    cells_one(P, B, s);
    const ListArgs L = list_args(P, B, B.ctl[CTL_CUR], B.ctl[CTL_HCUR]);
    walllist_one<BVH>(L, s, B.key[RS_FINAL][s] != 0xffffffffu);
  }
}

// Adds a thread's list lengths to the workgroup's sum in LDS (zeroed before
// a barrier that precedes this call) and the workgroup's sum to the scene's
// control word; every thread of the workgroup calls it.
__device__ __forceinline__ void add_list_lengths(int* ctl, int* wg_sum, int mine) {
  if (mine) atomicAdd(wg_sum, mine);
  __syncthreads();
  if (threadIdx.x == 0 && *wg_sum)
    atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[CTL_NBSUM]), (unsigned long long)*wg_sum);
}

// Pair lists of the sorted particles, at full occupancy (both the chain and
// the grid-barrier rebuild end here).  The last workgroup clears the refresh
// flag, which ends the rebuild.
template <class H, bool LONG>
__global__ __launch_bounds__(PL_BLOCK) void k_lists(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  // a cancelled grid barrier left the previous lists and the refresh flag in
  // place for the host to rebuild through the launch chain
  if (__builtin_amdgcn_readfirstlane(B.gbar[GBAR_ABORT]) != 0) return;
  __shared__ PairLds lds;
  __shared__ int nbsum;
  if (threadIdx.x == 0) nbsum = 0;
  __syncthreads();
  const ListArgs L = list_args(P, B, B.ctl[CTL_CUR], B.ctl[CTL_HCUR]);
  const int nact = B.ctl[CTL_NACT];
  int mine = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < P.N; s += gridDim.x * blockDim.x)
    mine += pairlist_one<PL_BLOCK, LONG>(L, s, s < nact, lds);
  add_list_lengths(B.ctl, &nbsum, mine);
  if (last_block(&B.ctl[CTL_DONE_WALL])) {
    B.ctl[CTL_DONE_WALL] = 0;
    B.ctl[CTL_REBUILDS] += 1;
    atomicExch(&B.ctl[CTL_REFRESH], 0);
  }
}

// ---- the global rebuild phases in one grid-barrier kernel ------------------
// Two-level generation barrier over a co-resident grid: workgroups arrive on
// BAR_SUB counters (b % BAR_SUB), the last arrival of each arrives on the top
// counter, and the last of those resets it and advances the generation word,
// which the others poll.  A single counter for every workgroup serialises the
// arrivals (about 0.1 us each on MI355X).  Each workgroup releases its writes
// before arriving and acquires after the generation moves; the counters are
// reset before the generation word is released, so the next barrier finds
// them at zero.  The wait is bounded (spin polls, about a second by default).
// A workgroup whose wait runs out tries to cancel the barrier by setting the
// generation word's top bit with a compare-and-swap from the value it
// arrived with; the last arrival advances the word with a compare-and-swap
// from the same value.  Exactly one of the two succeeds, so either every
// workgroup passes the barrier or none does: a cancelled barrier sets
// GDEM_ERR_BARRIER and the sticky abort word, every workgroup returns false,
// and no phase after it runs.  The state, the lists and the cell of every
// slot then remain those of the last complete rebuild (the phases before the
// barrier write only sort buffers and clear cell-table entries that the next
// rebuild clears again), the refresh flag stays set, and the host switches
// the scene to the launch chain (gdem_scene_step / gdem_scene_status).
constexpr uint32_t GBAR_CANCEL = 0x80000000u;

__device__ __forceinline__ bool grid_barrier(uint32_t* gbar, int* ctl, int spin) {
  __shared__ int go;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int G = gridDim.x, b = blockIdx.x;
    const int nsub = G < BAR_SUB ? G : BAR_SUB;
    const int sub = b % nsub;
    const int members = (G - sub + nsub - 1) / nsub;
    uint32_t* cnt = gbar + sub * 32;
    uint32_t* top = gbar + BAR_SUB * 32;
    uint32_t* gen = gbar + (BAR_SUB + 1) * 32;
    const uint32_t g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = (g0 & GBAR_CANCEL) == 0;  // a barrier of this launch was cancelled
    if (ok) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const uint32_t prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == (uint32_t)members - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t t = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (t == (uint32_t)nsub - 1) {
          __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          uint32_t expect = g0;
          __hip_atomic_compare_exchange_strong(gen, &expect, (g0 + 1u) & ~GBAR_CANCEL, __ATOMIC_RELEASE,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      for (int it = 0;; ++it) {
        const uint32_t gv = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gv != g0) {
          ok = (gv & GBAR_CANCEL) == 0;
          break;
        }
        if (it >= spin) {
          uint32_t expect = g0;
          if (__hip_atomic_compare_exchange_strong(gen, &expect, g0 | GBAR_CANCEL, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicOr((unsigned int*)&ctl[CTL_ERR], GDEM_ERR_BARRIER_BIT);
            __hip_atomic_store(gbar + GBAR_ABORT, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
          } else {
            ok = (expect & GBAR_CANCEL) == 0;  // completed (or cancelled by another) meanwhile
          }
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    go = ok;
  }
  __syncthreads();
  return go != 0;
}

// The whole rebuild in one kernel launched every step with a co-resident
// grid, which returns at once (before any barrier) when the refresh flag is
// clear: keys (counted for the first radix pass as they are made), the three
// radix passes, the permutation, the cell table and each slot's wall list,
// then -- after one more barrier, once every slot's permuted record and cell
// entry exist -- the pair lists.  One launch per step instead of two halves
// the cost of the steps that rebuild nothing.  Every workgroup reads the
// control words before the first barrier; the last workgroup to finish
// flips the buffers, counts the rebuild and clears the refresh flag.
template <class H, bool LONG, bool BVH>
__global__ __launch_bounds__(256) void k_rebuild_grid(SimParams P, Bufs<H> B) {
  if (!rebuild_due(B.ctl)) return;
  if (__builtin_amdgcn_readfirstlane(B.gbar[GBAR_ABORT]) != 0) return;  // an earlier barrier timed out
  const int cur = B.ctl[CTL_CUR];
  const int hc = B.ctl[CTL_HCUR];
  const int N = P.N;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int stride = G * blockDim.x;
  const int first = b * blockDim.x + threadIdx.x;
  uint32_t* hist = B.rs_hist;
  uint32_t* dtot = B.rs_hist + (size_t)RS_DIGITS * G;
  // the sort's counts and bases, and later the pair lists' columns, share LDS
  __shared__ union {
    struct {
      uint32_t dig[RS_DIGITS];  // digit counts, then scatter bases
      uint16_t wcnt[RS_WAVES * RS_DIGITS];
    } rs;
    PairLdsT<256> pl;
  } lds;
  uint32_t* dig = lds.rs.dig;
  uint16_t* wcnt = lds.rs.wcnt;
  __shared__ int nbsum;
  if (threadIdx.x == 0) {
    nbsum = 0;
    if (b == 0) *reinterpret_cast<unsigned long long*>(&B.ctl[CTL_NBSUM]) = 0ull;
  }
  // keys of this workgroup's chunk, counted for the first pass as they are made
  rs_clear(dig);
  {
    const int beg = b * P.rs_chunk;
    const int end = min(N, beg + P.rs_chunk);
    for (int s = beg + threadIdx.x; s < end; s += RS_THREADS) atomicAdd(&dig[rs_digit(keys_one(P, B, s, cur), 0)], 1u);
  }
  rs_store_counts(dig, hist, b, G);
  if (!grid_barrier(B.gbar, B.ctl, P.gbar_spin)) return;
  for (int pass = 0; pass < RS_PASSES; ++pass) {
    const int a = pass & 1;
    if (pass > 0) {
      rs_count_block(N, P.rs_chunk, pass, B.key[a], hist, b, G, dig);
      if (!grid_barrier(B.gbar, B.ctl, P.gbar_spin)) return;
    }
    rs_scan_digits(hist, dtot, G, b, G);
    if (!grid_barrier(B.gbar, B.ctl, P.gbar_spin)) return;
    rs_bases(hist, dtot, b, G, dig);
    rs_scatter_block(N, P.rs_chunk, pass, B.key[a], B.idx[a], B.key[a ^ 1], B.idx[a ^ 1],
                     pass == RS_PASSES - 1 ? B.inv : nullptr, dig, wcnt, b);
    if (!grid_barrier(B.gbar, B.ctl, P.gbar_spin)) return;
  }
  // The cell table needs the sorted keys (complete at the last barrier) and
  // the cell of each slot, which this thread's own permute step wrote; so
  // does the wall list, with the slot's permuted position.
  for (int s = first; s < N; s += stride) permute_one(P, B, s, cur, hc);
  const ListArgs L = list_args(P, B, cur ^ 1, hc ^ 1);
  for (int s = first; s < N; s += stride) {
    cells_one(P, B, s);
    walllist_one<BVH>(L, s, B.key[RS_FINAL][s] != 0xffffffffu);
  }
  // pair lists: every slot's neighbour cells and records are in place
  if (!grid_barrier(B.gbar, B.ctl, P.gbar_spin)) return;
  const int nact = B.ctl[CTL_NACT];
  int mine = 0;
  for (int s = first; s < N; s += stride) mine += pairlist_one<256, LONG>(L, s, s < nact, lds.pl);
  add_list_lengths(B.ctl, &nbsum, mine);
  if (last_block(&B.ctl[CTL_DONE_WALL])) {
    B.ctl[CTL_DONE_WALL] = 0;
    B.ctl[CTL_REBUILDS] += 1;
    B.ctl[CTL_CUR] = cur ^ 1;
    B.ctl[CTL_HCUR] = hc ^ 1;
    atomicExch(&B.ctl[CTL_REFRESH], 0);
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
  for (int q = 0; q < nn; ++q) D.nb[(size_t)id * P.nb_cap + q] = B.pid[hc][B.nbr[(size_t)q * N + s]];
  const int nw = B.nwl[s];
  D.nwall[id] = nw;
  for (int q = 0; q < nw; ++q) D.wall[(size_t)id * P.wall_cap + q] = B.wl[(size_t)q * N + s];
  const int nh = B.nh[hc][s];
  D.nhist[id] = nh;
  for (int q = 0; q < nh; ++q) {
    const int k = B.hkey[hc][(size_t)q * N + s];  // the partner's slot
    D.hist_key[(size_t)id * P.hist_cap + q] = (unsigned)k < (unsigned)N ? B.pid[hc][k] : -1;
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

// What the last rebuild stored, per slot: its sorted key, the original id it
// holds, its cell and that cell's table entry.
template <class H>
__global__ __launch_bounds__(256) void k_debug_order(SimParams P, Bufs<H> B, OrderDev D) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.N) return;
  const int hc = B.ctl[CTL_HCUR];
  D.key[s] = B.key[RS_FINAL][s];
  D.order[s] = B.pid[hc][s];
  const int c = B.cell_of[s];
  D.cell[s] = c;
  const int2 e = c >= 0 ? B.cell_tab[c] : make_int2(-1, -1);
  D.first[s] = e.x;
  D.end[s] = e.y;
  if (s == 0) D.nsorted[0] = B.ctl[CTL_NACT];
}

// inv[original id] = current slot, for k_set_history (inv is free between
// rebuilds)
template <class H>
__global__ __launch_bounds__(256) void k_slot_of_id(SimParams P, Bufs<H> B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.N) return;
  const int id = B.pid[B.ctl[CTL_HCUR]][s];
  if ((unsigned)id < (unsigned)P.N) B.inv[id] = s;
}

// ---- solver_output frames compacted on the device ---------------------------
// A frame holds the active particles in original id order (write_frame_bin,
// solver_output.cu:259-316).  inv maps each id to its slot (k_slot_of_id);
// a workgroup takes FRAME_BLOCK consecutive ids, FRAME_PER per thread.
constexpr int FRAME_PER = 8;
constexpr int FRAME_BLOCK = 256 * FRAME_PER;

// Exclusive scan over the workgroup of one count per thread (4 waves, LDS
// for the wave totals); returns the total in *total.
__device__ __forceinline__ int wg_excl_sum(int v, int* part, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_incl_sum(v);
  if (lane == 63) part[wave] = incl;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int t = part[w];
    if (w < wave) before += t;
    all += t;
  }
  *total = all;
  return before + incl - v;
}

template <class H>
__global__ __launch_bounds__(256) void k_frame_count(SimParams P, Bufs<H> B, int* cnt) {
  __shared__ int part[4];
  const int cur = B.ctl[CTL_CUR];
  const int id0 = blockIdx.x * FRAME_BLOCK + threadIdx.x * FRAME_PER;
  int c = 0;
#pragma unroll
  for (int k = 0; k < FRAME_PER; ++k) {
    const int id = id0 + k;
    if (id < P.N) c += B.pos[cur][B.inv[id]].w >= 0. ? 1 : 0;
  }
  int total = 0;
  (void)wg_excl_sum(c, part, &total);
  if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the nb block counts in place, the sum in
// cnt[nb].
__global__ __launch_bounds__(256) void k_frame_scan(int nb, int* cnt) {
  __shared__ int part[4];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += 256) {
    const int b = b0 + (int)threadIdx.x;
    const int v = b < nb ? cnt[b] : 0;
    int total = 0;
    const int ex = wg_excl_sum(v, part, &total);
    __syncthreads();  // every wave has read part[] before the next round writes it
    if (b < nb) cnt[b] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) cnt[nb] = carry;
}

// The frame's float32 values, each scaled by lf in FP64 before the
// conversion, at the active particles' ranks.
template <class H>
__global__ __launch_bounds__(256) void k_frame_write(SimParams P, Bufs<H> B, const int* off, double lf, float* xyz,
                                                     float* r) {
  __shared__ int part[4];
  const int cur = B.ctl[CTL_CUR];
  const int id0 = blockIdx.x * FRAME_BLOCK + threadIdx.x * FRAME_PER;
  double4 p[FRAME_PER];
  int c = 0;
#pragma unroll
  for (int k = 0; k < FRAME_PER; ++k) {
    const int id = id0 + k;
    p[k] = id < P.N ? B.pos[cur][B.inv[id]] : make_double4(0., 0., 0., -1.);
    c += p[k].w >= 0. ? 1 : 0;
  }
  int total = 0;
  int at = off[blockIdx.x] + wg_excl_sum(c, part, &total);
#pragma unroll
  for (int k = 0; k < FRAME_PER; ++k) {
    if (p[k].w < 0.) continue;
    xyz[(size_t)at * 3 + 0] = (float)(p[k].x * lf);
    xyz[(size_t)at * 3 + 1] = (float)(p[k].y * lf);
    xyz[(size_t)at * 3 + 2] = (float)(p[k].z * lf);
    r[at] = (float)(p[k].w * lf);
    ++at;
  }
}

// Contact histories from original particle order into the current internal
// order (snapshot restore); pair keys become partner slots.  Rows longer than the capacity raise the
// overflow bits and are cut at the capacity.
template <class H>
__global__ __launch_bounds__(256) void k_set_history(SimParams P, Bufs<H> B, HistIn in) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.N) return;
  const int N = P.N;
  const int hc = B.ctl[CTL_HCUR];
  const int id = B.pid[hc][s];
  int nh = in.nhist[id];
  if (nh > P.hist_cap) {
    atomicOr((unsigned int*)&B.ctl[CTL_ERR], GDEM_ERR_HIST_OVERFLOW_BIT);
    nh = P.hist_cap;
  }
  B.nh[hc][s] = nh;
  for (int q = 0; q < nh; ++q) {
    const size_t src = (size_t)id * P.hist_cap + q;
    const int k = in.hist_key[src];  // the partner's original id -> its slot
    B.hkey[hc][(size_t)q * N + s] = (unsigned)k < (unsigned)N ? B.inv[k] : -1;
    Del3<H> d;
    d.x = (H)in.hist_delta[src * 3];
    d.y = (H)in.hist_delta[src * 3 + 1];
    d.z = (H)in.hist_delta[src * 3 + 2];
    B.hd[hc][(size_t)q * N + s] = d;
  }
  int nw = in.nwhist[id];
  if (nw > P.whist_cap) {
    atomicOr((unsigned int*)&B.ctl[CTL_ERR], GDEM_ERR_WHIST_OVERFLOW_BIT);
    nw = P.whist_cap;
  }
  B.nw[hc][s] = nw;
  for (int q = 0; q < nw; ++q) {
    const size_t src = (size_t)id * P.whist_cap + q;
    B.wkey[hc][(size_t)q * N + s] = in.whist_key[src];
    Del3<H> d;
    d.x = (H)in.whist_delta[src * 3];
    d.y = (H)in.whist_delta[src * 3 + 1];
    d.z = (H)in.whist_delta[src * 3 + 2];
    B.wd[hc][(size_t)q * N + s] = d;
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
  const bool lds_rows = hist_lds_fits(P.N, P.hist_cap, P.whist_cap, PCAP_REG, WCAP_REG);
  if (lds_rows && P.list_prefetch)
    k_step<T, H, false, true, HistLds<H, PCAP_REG>, HistLds<H, WCAP_REG>><<<g, 256, 0, st>>>(P, B, none);
  else if (lds_rows)
    k_step<T, H, false, false, HistLds<H, PCAP_REG>, HistLds<H, WCAP_REG>><<<g, 256, 0, st>>>(P, B, none);
  else if (P.list_prefetch)
    k_step<T, H, false, true, HistScan<H>, HistScan<H>><<<g, 256, 0, st>>>(P, B, none);
  else
    k_step<T, H, false, false, HistScan<H>, HistScan<H>><<<g, 256, 0, st>>>(P, B, none);
}

template <class T, class H>
void launch_step_record(const SimParams& P, const Bufs<H>& B, const RecDev& r, hipStream_t st) {
  Rec<T> rec{r.cap, r.count, r.key, r.kind, r.fn, r.ft, r.delta, r.acc, r.anga};
  k_step<T, H, true, true, HistScan<H>, HistScan<H>><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, rec);
}

void launch_radix_sort(int n, int blocks, int chunk, uint32_t* k0, int* i0, uint32_t* k1, int* i1, int* inv_out,
                       uint32_t* hist, const int* ctl, bool gated, hipStream_t st) {
  uint32_t* kk[2] = {k0, k1};
  int* ii[2] = {i0, i1};
  for (int pass = 0; pass < RS_PASSES; ++pass) {
    const int a = pass & 1, b = a ^ 1;
    int* inv = pass == RS_PASSES - 1 ? inv_out : nullptr;
    if (gated) {
      k_rs_count<true><<<blocks, RS_THREADS, 0, st>>>(n, chunk, pass, kk[a], hist, ctl);
      k_rs_scan<true><<<RS_SCAN_BLOCKS, RS_THREADS, 0, st>>>(blocks, hist, ctl);
      k_rs_scatter<true><<<blocks, RS_THREADS, 0, st>>>(n, chunk, pass, kk[a], ii[a], kk[b], ii[b], inv, hist, ctl);
    } else {
      k_rs_count<false><<<blocks, RS_THREADS, 0, st>>>(n, chunk, pass, kk[a], hist, ctl);
      k_rs_scan<false><<<RS_SCAN_BLOCKS, RS_THREADS, 0, st>>>(blocks, hist, ctl);
      k_rs_scatter<false><<<blocks, RS_THREADS, 0, st>>>(n, chunk, pass, kk[a], ii[a], kk[b], ii[b], inv, hist, ctl);
    }
  }
}

template <class H>
int list_grid(int n) {
  const int want = blocks_for(n, PL_BLOCK);
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return want;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lists<H, false>, PL_BLOCK, 0) != hipSuccess || per_cu < 1)
    return want;
  return std::max(1, std::min(want, per_cu * prop.multiProcessorCount));
}

template <class H>
int coop_grid_limit() {
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rebuild_grid<H, false, false>, 256, 0) != hipSuccess) return 0;
  return per_cu * prop.multiProcessorCount;
}

template <class H>
hipError_t launch_rebuild(const SimParams& P, const Bufs<H>& B, bool coop, hipStream_t st) {
  if (coop) {
    // A normal launch: hipLaunchCooperativeKernel adds about 11 us of idle
    // time on each side of the kernel, and the grid is sized to be resident.
    const bool lng = P.nb_cap > PL_LDS_CAP, bvh = P.ntri > 0 && !P.wall_table;
    if (!lng && !bvh) k_rebuild_grid<H, false, false><<<P.rs_blocks, 256, 0, st>>>(P, B);
    else if (!lng) k_rebuild_grid<H, false, true><<<P.rs_blocks, 256, 0, st>>>(P, B);
    else if (!bvh) k_rebuild_grid<H, true, false><<<P.rs_blocks, 256, 0, st>>>(P, B);
    else k_rebuild_grid<H, true, true><<<P.rs_blocks, 256, 0, st>>>(P, B);
    return hipGetLastError();
  }
  const int g = blocks_for(P.N, 256);
  k_keys<H><<<g, 256, 0, st>>>(P, B);
  launch_radix_sort(P.N, P.rs_blocks, P.rs_chunk, B.key[0], B.idx[0], B.key[1], B.idx[1], B.inv, B.rs_hist,
                    B.ctl, true, st);
  k_permute<H><<<g, 256, 0, st>>>(P, B);
  if (P.ntri > 0 && !P.wall_table) k_cells<H, true><<<g, 256, 0, st>>>(P, B);
  else k_cells<H, false><<<g, 256, 0, st>>>(P, B);
  if (P.nb_cap <= PL_LDS_CAP) k_lists<H, false><<<P.list_blocks, PL_BLOCK, 0, st>>>(P, B);
  else k_lists<H, true><<<P.list_blocks, PL_BLOCK, 0, st>>>(P, B);
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
void launch_debug_order(const SimParams& P, const Bufs<H>& B, const OrderDev& D, hipStream_t st) {
  k_debug_order<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, D);
}

template <class H>
void launch_counts(const SimParams& P, const Bufs<H>& B, unsigned long long* acc, hipStream_t st) {
  k_counts<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, acc);
}

template <class H>
void launch_frame(const SimParams& P, const Bufs<H>& B, int* cnt, double lf, float* xyz, float* r, hipStream_t st) {
  const int nb = blocks_for(P.N, FRAME_BLOCK);
  k_slot_of_id<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B);
  k_frame_count<H><<<nb, 256, 0, st>>>(P, B, cnt);
  k_frame_scan<<<1, 256, 0, st>>>(nb, cnt);
  k_frame_write<H><<<nb, 256, 0, st>>>(P, B, cnt, lf, xyz, r);
}

int frame_blocks(int n) { return blocks_for(n, FRAME_BLOCK); }

template <class H>
void launch_set_history(const SimParams& P, const Bufs<H>& B, const HistIn& in, hipStream_t st) {
  k_slot_of_id<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B);
  k_set_history<H><<<blocks_for(P.N, 256), 256, 0, st>>>(P, B, in);
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
template int list_grid<double>(int);
template int list_grid<float>(int);
template void launch_download<double>(const SimParams&, const Bufs<double>&, double*, double*, double*, int*, hipStream_t);
template void launch_download<float>(const SimParams&, const Bufs<float>&, double*, double*, double*, int*, hipStream_t);
template void launch_debug<double>(const SimParams&, const Bufs<double>&, const DebugDev&, hipStream_t);
template void launch_debug<float>(const SimParams&, const Bufs<float>&, const DebugDev&, hipStream_t);
template void launch_debug_order<double>(const SimParams&, const Bufs<double>&, const OrderDev&, hipStream_t);
template void launch_debug_order<float>(const SimParams&, const Bufs<float>&, const OrderDev&, hipStream_t);
template void launch_counts<double>(const SimParams&, const Bufs<double>&, unsigned long long*, hipStream_t);
template void launch_counts<float>(const SimParams&, const Bufs<float>&, unsigned long long*, hipStream_t);
template void launch_frame<double>(const SimParams&, const Bufs<double>&, int*, double, float*, float*, hipStream_t);
template void launch_frame<float>(const SimParams&, const Bufs<float>&, int*, double, float*, float*, hipStream_t);
template void launch_set_history<double>(const SimParams&, const Bufs<double>&, const HistIn&, hipStream_t);
template void launch_set_history<float>(const SimParams&, const Bufs<float>&, const HistIn&, hipStream_t);

}  // namespace gdem
