// C ABI of the gdem engine: contexts, scenes, stepping, downloads, parity hooks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_internal.h"
#include "gdem_kernels.h"
#include "gdem_mesh.h"

struct gdem_scene {
  gdem_ctx* ctx = nullptr;
  gdem::SimParams P{};
  int precision = GDEM_FP64;
  int graph_steps = 0;
  bool coop = false;   // rebuild as one grid-barrier kernel (else the launch chain)
  bool chain = false;  // launch chain requested (per-phase profiling)
  int64_t steps = 0;
  int rebuild_base = 0;
  gdem::MeshHost mesh;
  gdem::DevArena arena;
  gdem::TriGeo* d_tri = nullptr;
  gdem::BvhNode* d_bvh = nullptr;
  double* d_tbox = nullptr;
  int* d_wc_off = nullptr;
  int* d_wc_tri = nullptr;
  int* d_ctl = nullptr;
  unsigned long long* d_acc = nullptr;
  gdem::Bufs<double> b64{};
  gdem::Bufs<float> b32{};
  hipGraphExec_t graph = nullptr;
  int timing = 0;
  std::vector<hipEvent_t> ev;  // pairs around force launches
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double last_total_ms = 0, last_force_ms = 0;
  int last_force_launches = 0;
  int chain_blocks = 1;               // radix-sort grid of the launch chain
  uint32_t* h_abort = nullptr;        // pinned copy of the barrier's abort word
  hipEvent_t abort_ev = nullptr;      // recorded after that copy
  bool abort_probe = false;           // a copy is in flight or done
  int* h_lstat = nullptr;             // pinned copy of the control words CTL_NACT .. CTL_NBSUM + 1 (the
                                      // active count, and the 64-bit sum of the last rebuild's list lengths)
  bool lstat_probe = false;           // a copy is in flight or done (abort_ev follows it)

  ~gdem_scene() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (abort_ev) (void)hipEventDestroy(abort_ev);
    if (h_abort) (void)hipHostFree(h_abort);
    if (h_lstat) (void)hipHostFree(h_lstat);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
  }
};

namespace gdem {
namespace {

// control words CTL_NACT .. CTL_NBSUM + 1, copied to the host after each step call
constexpr int LSTAT_WORDS = CTL_NBSUM + 2 - CTL_NACT;

template <class H>
int alloc_bufs(gdem_scene* sc, Bufs<H>& B, size_t ncell) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  auto& A = sc->arena;
  for (int h = 0; h < 2; ++h) {
    HIPCHK(c, A.alloc(&B.pos[h], N));
    HIPCHK(c, A.alloc(&B.vel[h], N));
    HIPCHK(c, A.alloc(&B.angv[h], N));
    HIPCHK(c, A.alloc(&B.mat[h], N));
    HIPCHK(c, A.alloc(&B.pid[h], N));
    HIPCHK(c, A.alloc(&B.hkey[h], N * sc->P.hist_cap));
    HIPCHK(c, A.alloc(&B.hd[h], N * sc->P.hist_cap));
    HIPCHK(c, A.alloc(&B.nh[h], N));
    HIPCHK(c, A.alloc(&B.wkey[h], N * sc->P.whist_cap));
    HIPCHK(c, A.alloc(&B.wd[h], N * sc->P.whist_cap));
    HIPCHK(c, A.alloc(&B.nw[h], N));
    HIPCHK(c, A.alloc(&B.key[h], N));
    HIPCHK(c, A.alloc(&B.idx[h], N));
  }
  HIPCHK(c, A.alloc(&B.ref, N));
  HIPCHK(c, A.alloc(&B.nbr, N * sc->P.nb_cap));
  HIPCHK(c, A.alloc(&B.inv, N));
  HIPCHK(c, A.alloc(&B.nnb, N));
  HIPCHK(c, A.alloc(&B.wl, N * sc->P.wall_cap));
  HIPCHK(c, A.alloc(&B.nwl, N));
  HIPCHK(c, A.alloc(&B.cell_tmp, N));
  HIPCHK(c, A.alloc(&B.cell_of, N));
  HIPCHK(c, A.alloc(&B.cell_tab, ncell));
  HIPCHK(c, A.alloc(&B.rs_hist, (size_t)RS_DIGITS * std::max(sc->P.rs_blocks, sc->chain_blocks) + RS_DIGITS));
  HIPCHK(c, A.alloc(&B.gbar, GBAR_WORDS));
  HIPCHK(c, hipMemsetAsync(B.gbar, 0, GBAR_WORDS * sizeof(uint32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(B.cell_tab, 0xff, ncell * sizeof(int2), c->stream));
  HIPCHK(c, hipMemsetAsync(B.cell_of, 0xff, N * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(B.nh[0], 0, N * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(B.nw[0], 0, N * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(B.nnb, 0, N * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(B.nwl, 0, N * sizeof(int), c->stream));
  B.tri = sc->d_tri;
  B.bvh = sc->d_bvh;
  B.tbox = sc->d_tbox;
  B.wc_off = sc->d_wc_off;
  B.wc_tri = sc->d_wc_tri;
  B.ctl = sc->d_ctl;
  return GDEM_OK;
}

template <class H>
Bufs<H>& bufs(gdem_scene* sc);
template <>
Bufs<double>& bufs<double>(gdem_scene* sc) { return sc->b64; }
template <>
Bufs<float>& bufs<float>(gdem_scene* sc) { return sc->b32; }

template <class H>
int enqueue_steps(gdem_scene* sc, int n) {
  Bufs<H>& B = bufs<H>(sc);
  hipStream_t st = sc->ctx->stream;
  for (int i = 0; i < n; ++i) {
    launch_step<H, H>(sc->P, B, st);
    HIPCHK(sc->ctx, launch_rebuild<H>(sc->P, B, sc->coop, st));
  }
  return GDEM_OK;
}

// A grid barrier of the rebuild kernel was cancelled (GDEM_ERR_BARRIER): the
// kernel stopped before any phase after it, so the state, the lists and the
// cell of every slot are those of the last complete rebuild and the refresh
// flag is still set.  The scene switches to the launch chain, which needs no
// co-resident grid, and rebuilds through it at once.  `wait`: synchronise
// and read the abort word; otherwise use the copy the last gdem_scene_step
// left, if it has landed.
template <class H>
int recover_barrier(gdem_scene* sc, bool wait) {
  if (!sc->coop) return GDEM_OK;
  gdem_ctx* c = sc->ctx;
  hipStream_t st = c->stream;
  Bufs<H>& B = bufs<H>(sc);
  uint32_t ab = 0;
  if (wait) {
    HIPCHK(c, hipStreamSynchronize(st));
    HIPCHK(c, hipMemcpy(&ab, B.gbar + GBAR_ABORT, 4, hipMemcpyDeviceToHost));
  } else {
    if (!sc->abort_probe || hipEventQuery(sc->abort_ev) != hipSuccess) return GDEM_OK;
    ab = *sc->h_abort;
  }
  if (!ab) return GDEM_OK;
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipMemsetAsync(B.gbar, 0, GBAR_WORDS * sizeof(uint32_t), st));
  sc->coop = false;
  sc->abort_probe = false;
  SimParams& P = sc->P;
  P.rs_blocks = sc->chain_blocks;
  P.rs_chunk = ((P.N + P.rs_blocks - 1) / P.rs_blocks + 255) / 256 * 256;
  if (sc->graph) {
    HIPCHK(c, hipGraphExecDestroy(sc->graph));
    sc->graph = nullptr;
  }
  HIPCHK(c, launch_rebuild<H>(P, B, false, st));
  HIPCHK(c, hipStreamSynchronize(st));
  return GDEM_OK;
}

// After the steps of a call, the abort word and the list statistics are
// copied to pinned host memory behind them, so the next call can see a
// cancelled barrier and the current mean list length without waiting.
template <class H>
int probe_barrier(gdem_scene* sc) {
  gdem_ctx* c = sc->ctx;
  if (sc->coop)
    HIPCHK(c, hipMemcpyAsync(sc->h_abort, bufs<H>(sc).gbar + GBAR_ABORT, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(sc->h_lstat, sc->d_ctl + CTL_NACT, LSTAT_WORDS * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipEventRecord(sc->abort_ev, c->stream));
  sc->abort_probe = sc->coop;
  sc->lstat_probe = true;
  return GDEM_OK;
}

// The step fetches its first 16 pair-list entries together with the own
// records (one memory round trip fewer) when the lists are long enough to use
// them: at a mean length of LIST_PREFETCH_MIN or more.  Shorter lists would
// mostly fetch rows past their ends, so they are read on demand.  The choice
// follows the lengths the last rebuild built, as seen by the latest probe; a
// change drops the captured graph, which holds the other kernel.
constexpr int LIST_PREFETCH_MIN = 4;

// the list-length sum held in control words CTL_NBSUM, CTL_NBSUM + 1, from a
// copy of the control words that starts at word `first`
unsigned long long nbsum_of(const int* words, int first) {
  unsigned long long v = 0;
  std::memcpy(&v, words + (CTL_NBSUM - first), sizeof(v));
  return v;
}

void set_prefetch(gdem_scene* sc, unsigned long long nbsum, long long nact) {
  const int want = nact > 0 && nbsum >= (unsigned long long)LIST_PREFETCH_MIN * (unsigned long long)nact ? 1 : 0;
  if (want == sc->P.list_prefetch) return;
  sc->P.list_prefetch = want;
  if (sc->graph) {
    (void)hipGraphExecDestroy(sc->graph);
    sc->graph = nullptr;
  }
}

void choose_prefetch(gdem_scene* sc) {
  if (!sc->lstat_probe || hipEventQuery(sc->abort_ev) != hipSuccess) return;
  set_prefetch(sc, nbsum_of(sc->h_lstat, CTL_NACT), sc->h_lstat[0]);
}

template <class H>
int step_impl(gdem_scene* sc, int nsteps) {
  gdem_ctx* c = sc->ctx;
  hipStream_t st = c->stream;
  Bufs<H>& B = bufs<H>(sc);
  {
    const int rc = recover_barrier<H>(sc, false);
    if (rc) return rc;
  }
  choose_prefetch(sc);
  if (sc->timing > 0) {
    // direct launches; an event pair brackets the force kernel of every
    // `timing`-th step (sampling keeps the host ahead of the device)
    const int stride = sc->timing;
    const int nsamp = (nsteps + stride - 1) / stride;
    for (auto e : sc->ev) (void)hipEventDestroy(e);
    sc->ev.assign((size_t)2 * nsamp, nullptr);
    for (auto& e : sc->ev) HIPCHK(c, hipEventCreate(&e));
    HIPCHK(c, hipEventRecord(sc->t0, st));
    for (int i = 0; i < nsteps; ++i) {
      const bool samp = (i % stride) == 0;
      if (samp) HIPCHK(c, hipEventRecord(sc->ev[2 * (i / stride)], st));
      launch_step<H, H>(sc->P, B, st);
      if (samp) HIPCHK(c, hipEventRecord(sc->ev[2 * (i / stride) + 1], st));
      HIPCHK(c, launch_rebuild<H>(sc->P, B, sc->coop, st));
    }
    HIPCHK(c, hipEventRecord(sc->t1, st));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(sc->t1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, sc->t0, sc->t1));
    sc->last_total_ms = ms;
    double fsum = 0;
    for (int i = 0; i < nsamp; ++i) {
      float m = 0;
      HIPCHK(c, hipEventElapsedTime(&m, sc->ev[2 * i], sc->ev[2 * i + 1]));
      fsum += m;
    }
    sc->last_force_ms = fsum;
    sc->last_force_launches = nsamp;
    sc->steps += nsteps;
    return probe_barrier<H>(sc);
  }
  int left = nsteps;
  if (sc->graph_steps > 0 && nsteps >= sc->graph_steps) {
    if (!sc->graph) {
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int erc = enqueue_steps<H>(sc, sc->graph_steps);
      HIPCHK(c, hipStreamEndCapture(st, &g));
      if (erc) return erc;
      hipError_t e = hipGraphInstantiate(&sc->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(c, e);
    }
    while (left >= sc->graph_steps) {
      HIPCHK(c, hipGraphLaunch(sc->graph, st));
      left -= sc->graph_steps;
    }
  }
  const int erc = enqueue_steps<H>(sc, left);
  if (erc) return erc;
  HIPCHK(c, hipGetLastError());
  sc->steps += nsteps;
  return probe_barrier<H>(sc);
}

template <class H>
int create_impl(gdem_scene* sc, const gdem_particles_desc* pd) {
  gdem_ctx* c = sc->ctx;
  const int N = sc->P.N;
  size_t ncell = (size_t)sc->P.size[0] * sc->P.size[1] * sc->P.size[2];
  Bufs<H>& B = bufs<H>(sc);
  // radix-sort / rebuild grid: one workgroup per chunk of >= 256 keys
  SimParams& P = sc->P;
  const int want = (N + 255) / 256;
  // The rebuild kernel's grid is at most the co-residency the occupancy query
  // reports (two workgroups per CU: the pair lists' LDS columns), and at most
  // 1024 workgroups: every workgroup must be resident at once for its grid
  // barriers.  A normal launch of a grid that fits is resident once the
  // preceding kernel has drained; if other work holds the device for longer
  // than the barrier's bounded wait, the barrier is cancelled and the scene
  // continues on the launch chain.
  const int limit = sc->chain ? 0 : std::min(1024, coop_grid_limit<H>());
  sc->coop = limit > 0;
  sc->chain_blocks = std::max(1, std::min(1024, (N + 2047) / 2048));
  P.rs_blocks = sc->coop ? std::max(1, std::min(limit, want)) : sc->chain_blocks;
  P.rs_chunk = ((N + P.rs_blocks - 1) / P.rs_blocks + 255) / 256 * 256;
  P.list_blocks = list_grid<H>(N);
  // GDEM_GBAR_SPIN (environment) lowers the barrier's poll limit, to exercise
  // its timeout path; unset, the default holds.
  P.gbar_spin = GBAR_SPIN;
  if (const char* e = std::getenv("GDEM_GBAR_SPIN")) P.gbar_spin = std::max(0, std::atoi(e));
  int rc = alloc_bufs<H>(sc, B, ncell);
  if (rc) return rc;
  using H4 = typename Vec4<H>::type;
  std::vector<double4> pos(N), vel(N), angv(N);
  std::vector<H4> mat(N);
  std::vector<int> pid(N);
  for (int i = 0; i < N; ++i) {
    const double r = pd->r[i], m = pd->m[i];
    const bool act = pd->is_active ? pd->is_active[i] == 1 : true;
    pos[i] = make_double4(pd->x[3 * i], pd->x[3 * i + 1], pd->x[3 * i + 2], act ? r : -r);
    const double* v = pd->v ? pd->v + 3 * i : nullptr;
    const double* w = pd->angv ? pd->angv + 3 * i : nullptr;
    // ParticleSystem.cu:294-335: invm = 1/m, sqrtm = sqrt(m), moi = 2/5 m r^2
    vel[i] = make_double4(v ? v[0] : 0., v ? v[1] : 0., v ? v[2] : 0., 1. / m);
    angv[i] = make_double4(w ? w[0] : 0., w ? w[1] : 0., w ? w[2] : 0., 0.);
    const double rsq = r * r;
    const double moi = 2. / 5. * m * rsq;
    const double mrow[4] = {pd->k[i], pd->etaconst[i], sqrt(m), 1. / moi};
    mat[i].x = (H)mrow[0];
    mat[i].y = (H)mrow[1];
    mat[i].z = (H)mrow[2];
    mat[i].w = (H)mrow[3];
    pid[i] = i;
  }
  hipStream_t st = c->stream;
  HIPCHK(c, hipMemcpyAsync(B.pos[0], pos.data(), N * sizeof(double4), hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B.vel[0], vel.data(), N * sizeof(double4), hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B.angv[0], angv.data(), N * sizeof(double4), hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B.mat[0], mat.data(), N * sizeof(H4), hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B.pid[0], pid.data(), N * sizeof(int), hipMemcpyHostToDevice, st));
  int ctl[CTL_WORDS] = {0};
  ctl[CTL_REFRESH] = 1;  // the initial lists are built like main.cu:151,196 do on the host
  HIPCHK(c, hipMemcpyAsync(sc->d_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice, st));
  HIPCHK(c, launch_rebuild<H>(sc->P, B, sc->coop, st));
  HIPCHK(c, hipStreamSynchronize(st));
  rc = recover_barrier<H>(sc, true);
  if (rc) return rc;
  int ctl2[CTL_WORDS];
  HIPCHK(c, hipMemcpy(ctl2, sc->d_ctl, sizeof(ctl2), hipMemcpyDeviceToHost));
  sc->rebuild_base = ctl2[CTL_REBUILDS];
  sc->P.list_prefetch = 1;
  set_prefetch(sc, nbsum_of(ctl2, 0), ctl2[CTL_NACT]);
  return GDEM_OK;
}

template <class H>
int download_impl(gdem_scene* sc, gdem_particles_out* out) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  DevArena tmp;
  double *x, *v, *w;
  int* a;
  HIPCHK(c, tmp.alloc(&x, 3 * N));
  HIPCHK(c, tmp.alloc(&v, 3 * N));
  HIPCHK(c, tmp.alloc(&w, 3 * N));
  HIPCHK(c, tmp.alloc(&a, N));
  launch_download<H>(sc->P, bufs<H>(sc), x, v, w, a, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out->x) HIPCHK(c, hipMemcpy(out->x, x, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->v) HIPCHK(c, hipMemcpy(out->v, v, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->angv) HIPCHK(c, hipMemcpy(out->angv, w, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->is_active) HIPCHK(c, hipMemcpy(out->is_active, a, N * 4, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

template <class H>
int frame_impl(gdem_scene* sc, double lf, float* xyz_out, float* r_out, int* n_active) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  DevArena tmp;
  float *xyz, *r;
  int* cnt;
  const int nb = frame_blocks((int)N);
  HIPCHK(c, tmp.alloc(&xyz, 3 * N));
  HIPCHK(c, tmp.alloc(&r, N));
  HIPCHK(c, tmp.alloc(&cnt, (size_t)nb + 1));
  launch_frame<H>(sc->P, bufs<H>(sc), cnt, lf, xyz, r, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int na = 0;
  HIPCHK(c, hipMemcpy(&na, cnt + nb, sizeof(int), hipMemcpyDeviceToHost));
  if (na > 0) {
    HIPCHK(c, hipMemcpy(xyz_out, xyz, (size_t)na * 3 * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(r_out, r, (size_t)na * sizeof(float), hipMemcpyDeviceToHost));
  }
  *n_active = na;
  return GDEM_OK;
}

template <class H>
int debug_impl(gdem_scene* sc, gdem_debug_view* vw) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  const SimParams& P = sc->P;
  DevArena tmp;
  DebugDev D{};
  HIPCHK(c, tmp.alloc(&D.cellx, 3 * N));
  HIPCHK(c, tmp.alloc(&D.cell_id, N));
  HIPCHK(c, tmp.alloc(&D.morton, N));
  HIPCHK(c, tmp.alloc(&D.nb, N * P.nb_cap));
  HIPCHK(c, tmp.alloc(&D.nnb, N));
  HIPCHK(c, tmp.alloc(&D.wall, N * P.wall_cap));
  HIPCHK(c, tmp.alloc(&D.nwall, N));
  HIPCHK(c, tmp.alloc(&D.hist_key, N * P.hist_cap));
  HIPCHK(c, tmp.alloc(&D.hist_delta, N * P.hist_cap * 3));
  HIPCHK(c, tmp.alloc(&D.nhist, N));
  HIPCHK(c, tmp.alloc(&D.whist_key, N * P.whist_cap));
  HIPCHK(c, tmp.alloc(&D.whist_delta, N * P.whist_cap * 3));
  HIPCHK(c, tmp.alloc(&D.nwhist, N));
  HIPCHK(c, hipMemsetAsync(D.nb, 0xff, N * P.nb_cap * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(D.wall, 0xff, N * P.wall_cap * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(D.hist_key, 0xff, N * P.hist_cap * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(D.whist_key, 0xff, N * P.whist_cap * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(D.hist_delta, 0, N * P.hist_cap * 24, c->stream));
  HIPCHK(c, hipMemsetAsync(D.whist_delta, 0, N * P.whist_cap * 24, c->stream));
  launch_debug<H>(P, bufs<H>(sc), D, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    if (dst) HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GDEM_OK;
  };
  int rc = 0;
  rc |= cp(vw->cellx, D.cellx, 3 * N * 4);
  rc |= cp(vw->cell_id, D.cell_id, N * 4);
  rc |= cp(vw->morton, D.morton, N * 4);
  rc |= cp(vw->nb, D.nb, N * P.nb_cap * 4);
  rc |= cp(vw->nnb, D.nnb, N * 4);
  rc |= cp(vw->wall, D.wall, N * P.wall_cap * 4);
  rc |= cp(vw->nwall, D.nwall, N * 4);
  rc |= cp(vw->hist_key, D.hist_key, N * P.hist_cap * 4);
  rc |= cp(vw->hist_delta, D.hist_delta, N * P.hist_cap * 24);
  rc |= cp(vw->nhist, D.nhist, N * 4);
  rc |= cp(vw->whist_key, D.whist_key, N * P.whist_cap * 4);
  rc |= cp(vw->whist_delta, D.whist_delta, N * P.whist_cap * 24);
  rc |= cp(vw->nwhist, D.nwhist, N * 4);
  return rc ? GDEM_E_HIP : GDEM_OK;
}

template <class H>
int debug_order_impl(gdem_scene* sc, gdem_order_view* vw) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  DevArena tmp;
  OrderDev D{};
  HIPCHK(c, tmp.alloc(&D.key, N));
  HIPCHK(c, tmp.alloc(&D.order, N));
  HIPCHK(c, tmp.alloc(&D.cell, N));
  HIPCHK(c, tmp.alloc(&D.first, N));
  HIPCHK(c, tmp.alloc(&D.end, N));
  HIPCHK(c, tmp.alloc(&D.nsorted, 1));
  launch_debug_order<H>(sc->P, bufs<H>(sc), D, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    if (dst) HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GDEM_OK;
  };
  int rc = 0;
  rc |= cp(vw->key, D.key, N * 4);
  rc |= cp(vw->order, D.order, N * 4);
  rc |= cp(vw->cell, D.cell, N * 4);
  rc |= cp(vw->cell_first, D.first, N * 4);
  rc |= cp(vw->cell_end, D.end, N * 4);
  rc |= cp(vw->n_sorted, D.nsorted, 4);
  return rc ? GDEM_E_HIP : GDEM_OK;
}

template <class H>
int record_impl(gdem_scene* sc, gdem_contact_view* vw) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  const int cap = vw->rec_cap;
  DevArena tmp;
  RecDev R{};
  R.cap = cap;
  HIPCHK(c, tmp.alloc(&R.count, N));
  HIPCHK(c, tmp.alloc(&R.key, N * cap));
  HIPCHK(c, tmp.alloc(&R.kind, N * cap));
  HIPCHK(c, tmp.alloc(&R.fn, N * cap * 3));
  HIPCHK(c, tmp.alloc(&R.ft, N * cap * 3));
  HIPCHK(c, tmp.alloc(&R.delta, N * cap));
  if (vw->accel) {
    HIPCHK(c, tmp.alloc(&R.acc, N * 3));
    HIPCHK(c, hipMemsetAsync(R.acc, 0, N * 24, c->stream));
  }
  if (vw->ang_accel) {
    HIPCHK(c, tmp.alloc(&R.anga, N * 3));
    HIPCHK(c, hipMemsetAsync(R.anga, 0, N * 24, c->stream));
  }
  HIPCHK(c, hipMemsetAsync(R.count, 0, N * 4, c->stream));
  launch_step_record<H, H>(sc->P, bufs<H>(sc), R, c->stream);
  HIPCHK(c, launch_rebuild<H>(sc->P, bufs<H>(sc), sc->coop, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  sc->steps += 1;
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    if (dst) HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GDEM_OK;
  };
  int rc = 0;
  rc |= cp(vw->count, R.count, N * 4);
  rc |= cp(vw->key, R.key, N * cap * 4);
  rc |= cp(vw->kind, R.kind, N * cap * 4);
  rc |= cp(vw->fn, R.fn, N * cap * 24);
  rc |= cp(vw->ft, R.ft, N * cap * 24);
  rc |= cp(vw->delta, R.delta, N * cap * 8);
  if (vw->accel) rc |= cp(vw->accel, R.acc, N * 24);
  if (vw->ang_accel) rc |= cp(vw->ang_accel, R.anga, N * 24);
  return rc ? GDEM_E_HIP : GDEM_OK;
}

template <class H>
int set_history_impl(gdem_scene* sc, const gdem_history_view* vw) {
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  const SimParams& P = sc->P;
  for (size_t i = 0; i < N; ++i)
    if (vw->nhist[i] < 0 || vw->nhist[i] > P.hist_cap || vw->nwhist[i] < 0 || vw->nwhist[i] > P.whist_cap)
      return fail(c, GDEM_E_SLOT_OVERFLOW, "history row longer than the scene's capacity");
  DevArena tmp;
  int *nh, *hk, *nw, *wk;
  double *hd, *wd;
  HIPCHK(c, tmp.alloc(&nh, N));
  HIPCHK(c, tmp.alloc(&hk, N * P.hist_cap));
  HIPCHK(c, tmp.alloc(&hd, N * P.hist_cap * 3));
  HIPCHK(c, tmp.alloc(&nw, N));
  HIPCHK(c, tmp.alloc(&wk, N * P.whist_cap));
  HIPCHK(c, tmp.alloc(&wd, N * P.whist_cap * 3));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(nh, vw->nhist, N * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(hk, vw->hist_key, N * P.hist_cap * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(hd, vw->hist_delta, N * P.hist_cap * 24, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(nw, vw->nwhist, N * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(wk, vw->whist_key, N * P.whist_cap * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(wd, vw->whist_delta, N * P.whist_cap * 24, hipMemcpyHostToDevice));
  HistIn in{nh, hk, hd, nw, wk, wd};
  launch_set_history<H>(P, bufs<H>(sc), in, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GDEM_OK;
}

template <class H>
int stats_impl(gdem_scene* sc, gdem_stats* out) {
  gdem_ctx* c = sc->ctx;
  {
    const int rc = recover_barrier<H>(sc, true);
    if (rc) return rc;
  }
  HIPCHK(c, hipMemsetAsync(sc->d_acc, 0, 4 * sizeof(unsigned long long), c->stream));
  launch_counts<H>(sc->P, bufs<H>(sc), sc->d_acc, c->stream);
  HIPCHK(c, hipGetLastError());
  unsigned long long acc[4];
  int ctl[CTL_WORDS];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(acc, sc->d_acc, sizeof(acc), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ctl, sc->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
  out->steps = sc->steps;
  out->rebuilds = ctl[CTL_REBUILDS] - sc->rebuild_base;
  out->n_active = (int)acc[0];
  out->errors = (uint32_t)ctl[CTL_ERR];
  const double na = acc[0] ? (double)acc[0] : 1.0;
  out->mean_nb = acc[1] / na;
  out->mean_wall = acc[2] / na;
  out->mean_contacts = acc[3] / na;
  return GDEM_OK;
}

}  // namespace
}  // namespace gdem

using namespace gdem;

extern "C" {

int gdem_ctx_create(int hip_device, gdem_ctx** out) {
  if (!out) return GDEM_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GDEM_E_HIP;
  if (hip_device < 0 || hip_device >= ndev) return GDEM_E_ARG;
  auto* c = new gdem_ctx();
  c->device = hip_device;
  if (hipSetDevice(hip_device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GDEM_E_HIP;
  }
  // the allocation-free download's staging array, once per device
  if (noalloc_init(hip_device) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GDEM_E_HIP;
  }
  *out = c;
  return GDEM_OK;
}

void gdem_ctx_destroy(gdem_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int gdem_sync(gdem_ctx* c) {
  if (!c) return GDEM_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GDEM_OK;
}

const char* gdem_last_error(gdem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gdem_device_arch(int dev, char* buf, int len, int* cus) {
  if (!buf || len <= 0) return GDEM_E_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GDEM_E_HIP;
  std::snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
  if (cus) *cus = prop.multiProcessorCount;
  return GDEM_OK;
}

int gdem_material(int n, const double* r, double density, double young, double cor, double* m_out,
                  double* k_out, double* eta_out) {
  if (n < 0 || !r || !m_out || !k_out || !eta_out || !(cor > 0.0)) return GDEM_E_ARG;
  for (int i = 0; i < n; ++i) {
    const double ri = r[i];
    const double m = density * 3.14 * ri * ri * ri * 4. / 3.;
    const double k = 4. / 3. * (young * 0.5) * sqrt(ri * 0.5) * sqrt(0.05 * ri);
    m_out[i] = m;
    k_out[i] = k;
    eta_out[i] = -2. * log(cor) * sqrt(k / (3.1415 * 3.1415 + log(cor) * log(cor)));
  }
  return GDEM_OK;
}

int gdem_bbox_rule(const gdem_mesh_desc* mesh, const double inlet[6], double bbox[6]) {
  if (!inlet || !bbox) return GDEM_E_ARG;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  if (mesh && mesh->ntri > 0) {
    for (int t = 0; t < mesh->ntri; ++t)
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
          const double q = mesh->verts[t * 9 + v * 3 + a];
          lo[a] = q < lo[a] ? q : lo[a];
          hi[a] = q > hi[a] ? q : hi[a];
        }
  }
  for (int a = 0; a < 3; ++a) {
    const double mn = lo[a] < inlet[2 * a] ? lo[a] : inlet[2 * a];
    const double mx = hi[a] > inlet[2 * a + 1] ? hi[a] : inlet[2 * a + 1];
    bbox[2 * a] = mn * 1.1;
    bbox[2 * a + 1] = mx * 1.1;
  }
  return GDEM_OK;
}

int gdem_params_default(double max_r, gdem_params* p) {
  if (!p || !(max_r > 0)) return GDEM_E_ARG;
  std::memset(p, 0, sizeof(*p));
  p->cell = max_r * 4.;
  p->skin = max_r;
  p->refresh = max_r * 0.5;
  p->precision = GDEM_FP64;
  return GDEM_OK;
}

int gdem_mesh_precompute(const gdem_mesh_desc* md, double* geo, int* ids, int* order) {
  if (!md || md->ntri <= 0 || !md->verts) return GDEM_E_ARG;
  MeshHost mh;
  if (!build_mesh(md->verts, md->ntri, md->tris_per_file, md->nfiles, &mh)) return GDEM_E_ARG;
  for (int t = 0; t < md->ntri; ++t) {
    const TriGeo& g = mh.tri[t];
    if (geo) std::memcpy(geo + (size_t)t * 29, &g, 29 * sizeof(double));
    if (ids) {
      const int v[6] = {g.v0, g.v1, g.v2, g.e0, g.e1, g.e2};
      std::memcpy(ids + (size_t)t * 6, v, sizeof(v));
    }
    if (order) order[t] = mh.order[t];
  }
  return GDEM_OK;
}

int gdem_scene_create(gdem_ctx* c, const gdem_particles_desc* pd, const gdem_mesh_desc* md,
                      const gdem_params* pp, gdem_scene** out) {
  if (!c || !pd || !pp || !out) return GDEM_E_ARG;
  *out = nullptr;
  if (pd->n <= 0 || !pd->x || !pd->r || !pd->m || !pd->k || !pd->etaconst)
    return fail(c, GDEM_E_ARG, "particles: n, x, r, m, k and etaconst are required");
  if (!(pp->dt > 0) || !(pp->cell > 0) || !(pp->skin >= 0) || !(pp->refresh >= 0))
    return fail(c, GDEM_E_ARG, "params: dt, cell must be > 0, skin and refresh >= 0");
  if (pp->precision != GDEM_FP64 && pp->precision != GDEM_MIXED)
    return fail(c, GDEM_E_ARG, "params: unknown precision");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, GDEM_E_HIP, "hipSetDevice");
  std::unique_ptr<gdem_scene> sc(new gdem_scene());
  sc->ctx = c;
  SimParams& P = sc->P;
  P.N = pd->n;
  P.dt = pp->dt;
  P.mu = pp->mu;
  double maxr = 0;
  for (int i = 0; i < pd->n; ++i) maxr = pd->r[i] > maxr ? pd->r[i] : maxr;
  for (int a = 0; a < 3; ++a) {
    P.g[a] = pp->g[a];
    P.bmin[a] = pp->bbox[2 * a];
    P.bmax[a] = pp->bbox[2 * a + 1];
    if (!(P.bmax[a] > P.bmin[a])) return fail(c, GDEM_E_ARG, "params: empty bounding box");
    P.invd[a] = 1. / pp->cell;
    // BoundingBox.cu:394-396
    const double cells = ceil((P.bmax[a] - P.bmin[a]) / pp->cell) + 2;
    if (cells > 1e6) return fail(c, GDEM_E_ARG, "params: cell grid too fine");
    P.size[a] = (int)cells;
  }
  const double ncell = (double)P.size[0] * P.size[1] * P.size[2];
  if (ncell > 1.5e9) return fail(c, GDEM_E_ARG, "params: cell grid has too many cells");
  if (pp->cell < 2 * maxr + pp->skin)
    return fail(c, GDEM_E_ARG, "params: cell must be >= 2*max(r) + skin");
  P.use_morton = (P.size[0] <= 1024 && P.size[1] <= 1024 && P.size[2] <= 1024) ? 1 : 0;
  P.skin = pp->skin;
  P.threshSq = pp->refresh * pp->refresh;
  P.nb_cap = pp->nb_cap > 0 ? pp->nb_cap : 48;
  P.wall_cap = pp->wall_cap > 0 ? pp->wall_cap : 24;
  P.hist_cap = pp->hist_cap > 0 ? pp->hist_cap : 16;
  P.whist_cap = pp->whist_cap > 0 ? pp->whist_cap : 8;
  if (P.nb_cap > 64 || P.wall_cap > 48 || P.hist_cap > 64 || P.whist_cap > 64)
    return fail(c, GDEM_E_ARG, "params: slot capacities exceed 64 (pairs/histories) or 48 (walls)");
  sc->precision = pp->precision;
  sc->graph_steps = pp->graph_steps > 0 ? pp->graph_steps : 0;
  sc->chain = pp->graph_steps < 0;
  // every active particle must lie in the box (the cell search assumes it)
  for (int i = 0; i < pd->n; ++i) {
    if (pd->is_active && pd->is_active[i] != 1) continue;
    if (!(pd->r[i] > 0)) return fail(c, GDEM_E_ARG, "particles: radius must be > 0");
    for (int a = 0; a < 3; ++a) {
      const double q = pd->x[3 * i + a];
      if (!(q >= P.bmin[a] && q <= P.bmax[a]))
        return fail(c, GDEM_E_ARG, "particles: active particle outside the bounding box");
    }
  }
  // mesh
  const int ntri = md ? md->ntri : 0;
  if (ntri > 0 && !md->verts) return fail(c, GDEM_E_ARG, "mesh: verts missing");
  if (!build_mesh(ntri > 0 ? md->verts : nullptr, ntri, md ? md->tris_per_file : nullptr,
                  md ? md->nfiles : 0, &sc->mesh))
    return fail(c, GDEM_E_ARG, "mesh: tris_per_file does not add up to ntri");
  P.ntri = ntri;
  P.nnode = (int)sc->mesh.nodes.size();
  HIPCHK(c, sc->arena.alloc(&sc->d_tri, std::max(ntri, 1)));
  HIPCHK(c, sc->arena.alloc(&sc->d_bvh, std::max(P.nnode, 1)));
  HIPCHK(c, sc->arena.alloc(&sc->d_tbox, (size_t)std::max(ntri, 1) * 6));
  // static per-cell wall table over the particle grid (the mesh never moves)
  P.wall_table = 0;
  if (ntri > 0 && pp->wall_broadphase != 1) {
    WallCells wc;
    const size_t max_entries = (size_t)1 << 27;
    const bool ok = ncell <= 4.0e8 &&
                    build_wall_cells(sc->mesh, P.bmin, pp->cell, P.size, maxr + pp->skin, max_entries, &wc);
    if (!ok && pp->wall_broadphase == 2)
      return fail(c, GDEM_E_ARG, "wall table does not fit; use wall_broadphase = 0 or 1");
    if (ok) {
      HIPCHK(c, sc->arena.alloc(&sc->d_wc_off, wc.off.size()));
      HIPCHK(c, sc->arena.alloc(&sc->d_wc_tri, std::max<size_t>(wc.tri.size(), 1)));
      HIPCHK(c, hipMemcpy(sc->d_wc_off, wc.off.data(), wc.off.size() * 4, hipMemcpyHostToDevice));
      if (!wc.tri.empty())
        HIPCHK(c, hipMemcpy(sc->d_wc_tri, wc.tri.data(), wc.tri.size() * 4, hipMemcpyHostToDevice));
      P.wall_table = 1;
    }
  }
  if (ntri > 0) {
    std::vector<double> tb((size_t)ntri * 6);
    for (int t = 0; t < ntri; ++t)
      for (int a = 0; a < 3; ++a) {
        tb[(size_t)t * 6 + a] = sc->mesh.aabb_lo[(size_t)t * 3 + a];
        tb[(size_t)t * 6 + 3 + a] = sc->mesh.aabb_hi[(size_t)t * 3 + a];
      }
    HIPCHK(c, hipMemcpy(sc->d_tbox, tb.data(), tb.size() * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(c, sc->arena.alloc(&sc->d_ctl, CTL_WORDS));
  HIPCHK(c, sc->arena.alloc(&sc->d_acc, 4));
  if (ntri > 0) {
    HIPCHK(c, hipMemcpy(sc->d_tri, sc->mesh.tri.data(), ntri * sizeof(TriGeo), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(sc->d_bvh, sc->mesh.nodes.data(), P.nnode * sizeof(BvhNode), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipEventCreate(&sc->t0));
  HIPCHK(c, hipEventCreate(&sc->t1));
  HIPCHK(c, hipEventCreateWithFlags(&sc->abort_ev, hipEventDisableTiming));
  HIPCHK(c, hipHostMalloc((void**)&sc->h_abort, sizeof(uint32_t), hipHostMallocDefault));
  *sc->h_abort = 0;
  HIPCHK(c, hipHostMalloc((void**)&sc->h_lstat, LSTAT_WORDS * sizeof(int), hipHostMallocDefault));
  for (int k = 0; k < LSTAT_WORDS; ++k) sc->h_lstat[k] = 0;
  int rc = sc->precision == GDEM_FP64 ? create_impl<double>(sc.get(), pd) : create_impl<float>(sc.get(), pd);
  if (rc) return rc;
  *out = sc.release();
  return GDEM_OK;
}

int gdem_scene_step(gdem_scene* sc, int nsteps) {
  if (!sc || nsteps < 0) return GDEM_E_ARG;
  if (nsteps == 0) return GDEM_OK;
  return sc->precision == GDEM_FP64 ? step_impl<double>(sc, nsteps) : step_impl<float>(sc, nsteps);
}

int gdem_scene_status(gdem_scene* sc, gdem_stats* st) {
  if (!sc) return GDEM_E_ARG;
  gdem_stats tmp;
  gdem_stats* o = st ? st : &tmp;
  int rc = sc->precision == GDEM_FP64 ? stats_impl<double>(sc, o) : stats_impl<float>(sc, o);
  if (rc) return rc;
  if (o->errors & (GDEM_ERR_NB_OVERFLOW | GDEM_ERR_WALL_OVERFLOW | GDEM_ERR_HIST_OVERFLOW |
                   GDEM_ERR_WHIST_OVERFLOW | GDEM_ERR_VE_OVERFLOW))
    return fail(sc->ctx, GDEM_E_SLOT_OVERFLOW, "slot overflow (raise the capacities in gdem_params)");
  if (o->errors & GDEM_ERR_NAN) return fail(sc->ctx, GDEM_E_NAN, "non-finite particle state");
  if (o->errors & GDEM_ERR_BARRIER) {
    // reported once: the scene rebuilds through the launch chain from now on
    gdem_ctx* c = sc->ctx;
    int err = 0;
    HIPCHK(c, hipMemcpy(&err, sc->d_ctl + CTL_ERR, sizeof(int), hipMemcpyDeviceToHost));
    err &= ~(int)GDEM_ERR_BARRIER;
    HIPCHK(c, hipMemcpy(sc->d_ctl + CTL_ERR, &err, sizeof(int), hipMemcpyHostToDevice));
    return fail(sc->ctx, GDEM_E_HIP,
                "list rebuild: a grid barrier timed out; the scene now rebuilds through per-phase kernels");
  }
  return GDEM_OK;
}

int gdem_scene_frame(gdem_scene* sc, double length_factor, float* xyz_out, float* r_out, int* n_active) {
  if (!sc || !xyz_out || !r_out || !n_active) return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? frame_impl<double>(sc, length_factor, xyz_out, r_out, n_active)
                                    : frame_impl<float>(sc, length_factor, xyz_out, r_out, n_active);
}

int gdem_scene_download_noalloc(gdem_scene* sc, gdem_particles_out* out) {
  if (!sc || !out) return GDEM_E_ARG;
  gdem_ctx* c = sc->ctx;
  const size_t N = (size_t)sc->P.N;
  if (N > (size_t)NOALLOC_MAX_N)
    return fail(c, GDEM_E_ARG, "gdem_scene_download_noalloc: more than 262144 particles (use gdem_scene_download)");
  const NoallocStage S = noalloc_stage(c->device);
  if (!S.x) return fail(c, GDEM_E_STATE, "gdem_scene_download_noalloc: staging not set up on this device");
  if (sc->precision == GDEM_FP64)
    launch_download<double>(sc->P, sc->b64, S.x, S.v, S.w, S.act, c->stream);
  else
    launch_download<float>(sc->P, sc->b32, S.x, S.v, S.w, S.act, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out->x) HIPCHK(c, hipMemcpy(out->x, S.x, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->v) HIPCHK(c, hipMemcpy(out->v, S.v, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->angv) HIPCHK(c, hipMemcpy(out->angv, S.w, 3 * N * 8, hipMemcpyDeviceToHost));
  if (out->is_active) HIPCHK(c, hipMemcpy(out->is_active, S.act, N * 4, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

int gdem_scene_download(gdem_scene* sc, gdem_particles_out* out) {
  if (!sc || !out) return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? download_impl<double>(sc, out) : download_impl<float>(sc, out);
}

void gdem_scene_destroy(gdem_scene* sc) {
  if (!sc) return;
  (void)hipStreamSynchronize(sc->ctx->stream);
  delete sc;
}

int gdem_scene_debug(gdem_scene* sc, gdem_debug_view* vw) {
  if (!sc || !vw) return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? debug_impl<double>(sc, vw) : debug_impl<float>(sc, vw);
}

int gdem_scene_debug_order(gdem_scene* sc, gdem_order_view* vw) {
  if (!sc || !vw) return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? debug_order_impl<double>(sc, vw) : debug_order_impl<float>(sc, vw);
}

int gdem_scene_step_record(gdem_scene* sc, gdem_contact_view* vw) {
  if (!sc || !vw || vw->rec_cap <= 0) return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? record_impl<double>(sc, vw) : record_impl<float>(sc, vw);
}

int gdem_scene_set_history(gdem_scene* sc, const gdem_history_view* vw) {
  if (!sc || !vw || !vw->nhist || !vw->hist_key || !vw->hist_delta || !vw->nwhist || !vw->whist_key ||
      !vw->whist_delta)
    return GDEM_E_ARG;
  return sc->precision == GDEM_FP64 ? set_history_impl<double>(sc, vw) : set_history_impl<float>(sc, vw);
}

int gdem_scene_set_timing(gdem_scene* sc, int on) {
  if (!sc) return GDEM_E_ARG;
  sc->timing = on > 0 ? on : 0;
  return GDEM_OK;
}

int gdem_scene_timing(gdem_scene* sc, double* total_ms, double* force_ms, int* launches) {
  if (!sc) return GDEM_E_ARG;
  if (total_ms) *total_ms = sc->last_total_ms;
  if (force_ms) *force_ms = sc->last_force_ms;
  if (launches) *launches = sc->last_force_launches;
  return GDEM_OK;
}

int gdem_radix_sort(gdem_ctx* c, int n, const uint32_t* keys, uint32_t* keys_out, int* order_out) {
  if (!c || n < 0 || (n > 0 && (!keys || !keys_out || !order_out))) return GDEM_E_ARG;
  if (n == 0) return GDEM_OK;
  DevArena tmp;
  uint32_t *k0, *k1, *hist;
  int *i0, *i1, *ctl;
  const int blocks = std::max(1, std::min(1024, (n + 2047) / 2048));
  const int chunk = ((n + blocks - 1) / blocks + 255) / 256 * 256;
  HIPCHK(c, tmp.alloc(&k0, n));
  HIPCHK(c, tmp.alloc(&k1, n));
  HIPCHK(c, tmp.alloc(&i0, n));
  HIPCHK(c, tmp.alloc(&i1, n));
  HIPCHK(c, tmp.alloc(&hist, (size_t)RS_DIGITS * blocks + RS_DIGITS));
  HIPCHK(c, tmp.alloc(&ctl, CTL_WORDS));
  std::vector<int> iota(n);
  for (int i = 0; i < n; ++i) iota[i] = i;
  HIPCHK(c, hipMemcpy(k0, keys, (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(i0, iota.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  launch_radix_sort(n, blocks, chunk, k0, i0, k1, i1, nullptr, hist, ctl, false, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(keys_out, RS_FINAL ? k1 : k0, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(order_out, RS_FINAL ? i1 : i0, (size_t)n * 4, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

int gdem_tri_distance(gdem_ctx* c, const gdem_mesh_desc* md, int tri, int np, const double* pts,
                      const double* rsq, double* out, int* hit) {
  if (!c || !md || md->ntri <= 0 || tri < 0 || tri >= md->ntri || np <= 0 || !pts || !rsq || !out || !hit)
    return GDEM_E_ARG;
  MeshHost mh;
  if (!build_mesh(md->verts, md->ntri, md->tris_per_file, md->nfiles, &mh)) return GDEM_E_ARG;
  DevArena tmp;
  TriGeo* t;
  double *p, *r, *o;
  int* h;
  HIPCHK(c, tmp.alloc(&t, 1));
  HIPCHK(c, tmp.alloc(&p, (size_t)np * 3));
  HIPCHK(c, tmp.alloc(&r, (size_t)np));
  HIPCHK(c, tmp.alloc(&o, (size_t)np * 4));
  HIPCHK(c, tmp.alloc(&h, (size_t)np));
  HIPCHK(c, hipMemcpy(t, &mh.tri[tri], sizeof(TriGeo), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(p, pts, (size_t)np * 24, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(r, rsq, (size_t)np * 8, hipMemcpyHostToDevice));
  launch_tri_probe(t, np, p, r, o, h, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(out, o, (size_t)np * 32, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(hit, h, (size_t)np * 4, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

int gdem_copy_probe(gdem_ctx* c, int64_t bytes, int reps, double* ms_out) {
  if (!c || bytes < 16 || reps <= 0) return GDEM_E_ARG;
  DevArena tmp;
  double4 *a, *b;
  const size_t n = (size_t)bytes / sizeof(double4);
  HIPCHK(c, tmp.alloc(&a, n));
  HIPCHK(c, tmp.alloc(&b, n));
  HIPCHK(c, hipMemsetAsync(a, 0, n * sizeof(double4), c->stream));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  launch_copy(a, b, n, c->stream);
  HIPCHK(c, hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) launch_copy(a, b, n, c->stream);
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_out) *ms_out = ms / reps;
  return GDEM_OK;
}

int gdem_fp64_probe(gdem_ctx* c, int n, const double* a, const double* b, double* sq, double* dv) {
  if (!c || n <= 0 || !a || !b || !sq || !dv) return GDEM_E_ARG;
  DevArena tmp;
  double *da, *db, *ds, *dd;
  HIPCHK(c, tmp.alloc(&da, n));
  HIPCHK(c, tmp.alloc(&db, n));
  HIPCHK(c, tmp.alloc(&ds, n));
  HIPCHK(c, tmp.alloc(&dd, n));
  HIPCHK(c, hipMemcpy(da, a, (size_t)n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(db, b, (size_t)n * 8, hipMemcpyHostToDevice));
  launch_fp64_probe(n, da, db, ds, dd, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(sq, ds, (size_t)n * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(dv, dd, (size_t)n * 8, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

}  // extern "C"
