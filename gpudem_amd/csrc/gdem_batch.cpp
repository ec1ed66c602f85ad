// C ABI of the batched engine: one set of device arrays shared by many
// independent scenes, scene admission against fixed capacities (the HBM
// budget), eviction, and stepping every running scene together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_batch.h"
#include "gdem_internal.h"
#include "gdem_mesh.h"

namespace gdem {
namespace {

// First-fit allocator of index ranges, lowest address first, so the live
// ranges stay packed at the bottom and the launch grid (which covers up to
// the highest allocated slot) stays small.
class RangeAlloc {
 public:
  void reset(int64_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  bool alloc(int64_t n, int64_t* start) {
    if (n <= 0) {
      *start = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *start = it->first;
      const int64_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(int64_t start, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(start, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  // one past the highest allocated index
  int64_t high_water() const {
    if (free_.empty()) return cap_;
    auto last = std::prev(free_.end());
    return (last->first + last->second == cap_) ? last->first : cap_;
  }
  int64_t free_total() const {
    int64_t t = 0;
    for (auto& f : free_) t += f.second;
    return t;
  }

 private:
  int64_t cap_ = 0;
  std::map<int64_t, int64_t> free_;
};

struct Slot {
  bool used = false;
  int n = 0;
  int64_t seg0 = 0, nseg = 0;   // particle slots
  int64_t cell0 = 0, ncell1 = 0;  // cells (+1 wall-table end entry)
  int64_t tri0 = 0, ntri = 0;
  int64_t wt0 = 0, nwt = 0;
  int64_t start = 0, end = 0;
  int64_t bytes = 0;
};

constexpr int64_t kGridGranule = 16;  // workgroups: graphs are keyed by the grid rounded up to this
constexpr int64_t kSlotGranule = 64;  // particle slots per wave

}  // namespace
}  // namespace gdem

struct gdem_batch {
  gdem_ctx* ctx = nullptr;
  int precision = GDEM_FP64;
  int nslots = 0;
  int64_t pcap = 0, ccap = 0, tcap = 0, wcap = 0;
  int nb_cap = 48, wall_cap = 24, hist_cap = 16, whist_cap = 8;
  int graph_steps = 0;
  int64_t bytes_total = 0, bytes_per_particle = 0;
  gdem::DevArena arena;
  gdem::BBufs<double> b64{};
  gdem::BBufs<float> b32{};
  gdem::BScene* d_scenes = nullptr;
  gdem::SlotCtl* d_sctl = nullptr;
  int* d_ctl = nullptr;
  int* d_wave = nullptr;
  int4* d_due = nullptr;
  gdem::TriGeo* d_tri = nullptr;
  double* d_tbox = nullptr;
  int* d_wc_off = nullptr;
  int* d_wc_tri = nullptr;
  unsigned long long* d_acc = nullptr;
  gdem::RangeAlloc parts, cells, tris, wtab;
  std::vector<gdem::Slot> slots;
  int64_t step = 0;
  std::map<int, hipGraphExec_t> graphs;  // by grid
  hipEvent_t t0 = nullptr, t1 = nullptr, up = nullptr;
  std::vector<std::vector<char>> pending;  // host sources of queued uploads

  ~gdem_batch() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (up) (void)hipEventDestroy(up);
  }
};

namespace gdem {
namespace {

template <class H>
BBufs<H>& bbufs(gdem_batch* b);
template <>
BBufs<double>& bbufs<double>(gdem_batch* b) { return b->b64; }
template <>
BBufs<float>& bbufs<float>(gdem_batch* b) { return b->b32; }

int64_t particle_bytes(int hbytes, int nb_cap, int wall_cap, int hist_cap, int whist_cap) {
  const int64_t h4 = 4 * hbytes;
  return 6 * 32               // pos, vel, angv (two halves)
         + 2 * h4 + 2 * 4     // material, ids
         + 2 * (int64_t)hist_cap * (4 + h4) + 2 * 4   // pair histories
         + 2 * (int64_t)whist_cap * (4 + h4) + 2 * 4  // wall histories
         + 32                  // reference positions
         + (int64_t)nb_cap * 4 + 4 + (int64_t)wall_cap * 4 + 4  // candidate lists
         + 3 * 4;              // binning
}
constexpr int64_t kCellBytes = 4 + 8 + 4;    // counter, (first, end), wall-table offset
constexpr int64_t kTriBytes = 256 + 6 * 8;   // geometry + box

// Queue a host->device copy whose source stays alive until the upload event
// has passed.
int upload(gdem_batch* b, void* dst, const void* src, size_t bytes) {
  if (!bytes) return GDEM_OK;
  b->pending.emplace_back((const char*)src, (const char*)src + bytes);
  HIPCHK(b->ctx, hipMemcpyAsync(dst, b->pending.back().data(), bytes, hipMemcpyHostToDevice, b->ctx->stream));
  return GDEM_OK;
}
void reap_uploads(gdem_batch* b) {
  if (!b->pending.empty() && hipEventQuery(b->up) == hipSuccess) b->pending.clear();
}

template <class H>
int alloc_batch(gdem_batch* b) {
  gdem_ctx* c = b->ctx;
  auto& A = b->arena;
  BBufs<H>& B = bbufs<H>(b);
  const size_t N = (size_t)b->pcap;
  for (int h = 0; h < 2; ++h) {
    HIPCHK(c, A.alloc(&B.pos[h], N));
    HIPCHK(c, A.alloc(&B.vel[h], N));
    HIPCHK(c, A.alloc(&B.angv[h], N));
    HIPCHK(c, A.alloc(&B.mat[h], N));
    HIPCHK(c, A.alloc(&B.pid[h], N));
    HIPCHK(c, A.alloc(&B.hkey[h], N * b->hist_cap));
    HIPCHK(c, A.alloc(&B.hd[h], N * b->hist_cap));
    HIPCHK(c, A.alloc(&B.nh[h], N));
    HIPCHK(c, A.alloc(&B.wkey[h], N * b->whist_cap));
    HIPCHK(c, A.alloc(&B.wd[h], N * b->whist_cap));
    HIPCHK(c, A.alloc(&B.nw[h], N));
    HIPCHK(c, hipMemsetAsync(B.nh[h], 0, N * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(B.nw[h], 0, N * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(B.pid[h], 0xff, N * 4, c->stream));
  }
  HIPCHK(c, A.alloc(&B.ref, N));
  HIPCHK(c, A.alloc(&B.nbr, N * b->nb_cap));
  HIPCHK(c, A.alloc(&B.nnb, N));
  HIPCHK(c, A.alloc(&B.wl, N * b->wall_cap));
  HIPCHK(c, A.alloc(&B.nwl, N));
  HIPCHK(c, A.alloc(&B.bin_cell, N));
  HIPCHK(c, A.alloc(&B.bin_rank, N));
  HIPCHK(c, A.alloc(&B.cell_of, N));
  HIPCHK(c, A.alloc(&B.ccount, (size_t)b->ccap));
  HIPCHK(c, A.alloc(&B.cell_tab, (size_t)b->ccap));
  HIPCHK(c, hipMemsetAsync(B.nnb, 0, N * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(B.nwl, 0, N * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(B.ccount, 0, (size_t)b->ccap * 4, c->stream));
  B.tri = b->d_tri;
  B.tbox = b->d_tbox;
  B.wc_off = b->d_wc_off;
  B.wc_tri = b->d_wc_tri;
  B.wave_scene = b->d_wave;
  B.due = b->d_due;
  B.scenes = b->d_scenes;
  B.sctl = b->d_sctl;
  B.ctl = b->d_ctl;
  B.cap = (int)b->pcap;
  B.nb_cap = b->nb_cap;
  B.wall_cap = b->wall_cap;
  B.hist_cap = b->hist_cap;
  B.whist_cap = b->whist_cap;
  B.nslots = b->nslots;
  return GDEM_OK;
}

template <class H>
int upload_particles(gdem_batch* b, const Slot& sl, const gdem_particles_desc* pd) {
  using H4 = typename Vec4<H>::type;
  BBufs<H>& B = bbufs<H>(b);
  const int64_t ns = sl.nseg;
  std::vector<double4> pos(ns), vel(ns), angv(ns);
  std::vector<H4> mat(ns);
  std::vector<int> pid(ns, -1), zero(ns, 0);
  for (int64_t i = 0; i < ns; ++i) {
    if (i >= pd->n) {
      pos[i] = make_double4(0., 0., 0., -1.);
      vel[i] = make_double4(0., 0., 0., 0.);
      angv[i] = make_double4(0., 0., 0., 0.);
      mat[i] = H4{};
      continue;
    }
    const double r = pd->r[i], m = pd->m[i];
    const bool act = pd->is_active ? pd->is_active[i] == 1 : true;
    pos[i] = make_double4(pd->x[3 * i], pd->x[3 * i + 1], pd->x[3 * i + 2], act ? r : -r);
    const double* v = pd->v ? pd->v + 3 * i : nullptr;
    const double* w = pd->angv ? pd->angv + 3 * i : nullptr;
    // ParticleSystem.cu:294-335: invm = 1/m, sqrtm = sqrt(m), moi = 2/5 m r^2
    vel[i] = make_double4(v ? v[0] : 0., v ? v[1] : 0., v ? v[2] : 0., 1. / m);
    angv[i] = make_double4(w ? w[0] : 0., w ? w[1] : 0., w ? w[2] : 0., 0.);
    const double moi = 2. / 5. * m * (r * r);
    mat[i].x = (H)pd->k[i];
    mat[i].y = (H)pd->etaconst[i];
    mat[i].z = (H)sqrt(m);
    mat[i].w = (H)(1. / moi);
    pid[i] = (int)i;
  }
  // a new scene starts with q = 0, so its state lives in half (STEP & 1)
  const int cur = (int)(b->step & 1);
  int rc = 0;
  rc |= upload(b, B.pos[cur] + sl.seg0, pos.data(), ns * sizeof(double4));
  rc |= upload(b, B.vel[cur] + sl.seg0, vel.data(), ns * sizeof(double4));
  rc |= upload(b, B.angv[cur] + sl.seg0, angv.data(), ns * sizeof(double4));
  rc |= upload(b, B.mat[0] + sl.seg0, mat.data(), ns * sizeof(H4));
  rc |= upload(b, B.pid[0] + sl.seg0, pid.data(), ns * sizeof(int));
  rc |= upload(b, B.nh[0] + sl.seg0, zero.data(), ns * sizeof(int));
  rc |= upload(b, B.nw[0] + sl.seg0, zero.data(), ns * sizeof(int));
  return rc ? GDEM_E_HIP : GDEM_OK;
}

template <class H>
int batch_step_impl(gdem_batch* b, int nsteps) {
  gdem_ctx* c = b->ctx;
  hipStream_t st = c->stream;
  reap_uploads(b);
  const int64_t hwm = b->parts.high_water();
  if (hwm == 0) return GDEM_OK;  // no scene: nothing runs and the batch clock stands still
  int64_t grid = (hwm + 255) / 256;
  grid = (grid + kGridGranule - 1) / kGridGranule * kGridGranule;
  grid = std::min<int64_t>(grid, b->pcap / 256);
  BBufs<H>& B = bbufs<H>(b);
  HIPCHK(c, hipEventRecord(b->t0, st));
  int left = nsteps;
  if (b->graph_steps > 0 && nsteps >= b->graph_steps) {
    hipGraphExec_t& ex = b->graphs[(int)grid];
    if (!ex) {
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < b->graph_steps; ++i) launch_batch_step<H, H>(B, (int)grid, st);
      HIPCHK(c, hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        b->graphs.erase((int)grid);
        HIPCHK(c, e);
      }
    }
    while (left >= b->graph_steps) {
      HIPCHK(c, hipGraphLaunch(ex, st));
      left -= b->graph_steps;
    }
  }
  for (int i = 0; i < left; ++i) launch_batch_step<H, H>(B, (int)grid, st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(b->t1, st));
  b->step += nsteps;
  return GDEM_OK;
}

template <class H>
int batch_download_impl(gdem_batch* b, int sid, gdem_particles_out* out) {
  gdem_ctx* c = b->ctx;
  const Slot& sl = b->slots[sid];
  const size_t n = (size_t)sl.n;
  DevArena tmp;
  double *x, *v, *w;
  int* a;
  HIPCHK(c, tmp.alloc(&x, 3 * n));
  HIPCHK(c, tmp.alloc(&v, 3 * n));
  HIPCHK(c, tmp.alloc(&w, 3 * n));
  HIPCHK(c, tmp.alloc(&a, n));
  launch_batch_download<H>(bbufs<H>(b), sid, (int)sl.nseg, x, v, w, a, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out->x) HIPCHK(c, hipMemcpy(out->x, x, 3 * n * 8, hipMemcpyDeviceToHost));
  if (out->v) HIPCHK(c, hipMemcpy(out->v, v, 3 * n * 8, hipMemcpyDeviceToHost));
  if (out->angv) HIPCHK(c, hipMemcpy(out->angv, w, 3 * n * 8, hipMemcpyDeviceToHost));
  if (out->is_active) HIPCHK(c, hipMemcpy(out->is_active, a, n * 4, hipMemcpyDeviceToHost));
  return GDEM_OK;
}

// Contact histories of one scene in its particles' own ids: the slot rows of
// the scene's segment, the ids of its slots and the buffer parity are
// copied back and translated on the host (pair keys are partner slots on the
// device, wall keys global triangle ids).
template <class H>
int batch_history_impl(gdem_batch* b, int sid, const gdem_history_out* out) {
  gdem_ctx* c = b->ctx;
  BBufs<H>& B = bbufs<H>(b);
  const Slot& sl = b->slots[sid];
  const size_t ns = (size_t)sl.nseg, N = (size_t)b->pcap, n = (size_t)sl.n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  SlotCtl ctl;
  HIPCHK(c, hipMemcpy(&ctl, b->d_sctl + sid, sizeof(ctl), hipMemcpyDeviceToHost));
  const int q = ctl.q;
  std::vector<int> pid(ns), nh(ns), nw(ns), hk((size_t)b->hist_cap * ns), wk((size_t)b->whist_cap * ns);
  std::vector<Del3<H>> hd((size_t)b->hist_cap * ns), wd((size_t)b->whist_cap * ns);
  HIPCHK(c, hipMemcpy(pid.data(), B.pid[q] + sl.seg0, ns * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(nh.data(), B.nh[q] + sl.seg0, ns * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(nw.data(), B.nw[q] + sl.seg0, ns * 4, hipMemcpyDeviceToHost));
  for (int k = 0; k < b->hist_cap; ++k) {
    HIPCHK(c, hipMemcpy(hk.data() + k * ns, B.hkey[q] + k * N + sl.seg0, ns * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hd.data() + k * ns, B.hd[q] + k * N + sl.seg0, ns * sizeof(Del3<H>), hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < b->whist_cap; ++k) {
    HIPCHK(c, hipMemcpy(wk.data() + k * ns, B.wkey[q] + k * N + sl.seg0, ns * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(wd.data() + k * ns, B.wd[q] + k * N + sl.seg0, ns * sizeof(Del3<H>), hipMemcpyDeviceToHost));
  }
  for (size_t i = 0; i < ns; ++i) {
    const int id = pid[i];
    if (id < 0 || (size_t)id >= n) continue;
    out->nhist[id] = nh[i];
    for (int k = 0; k < nh[i]; ++k) {
      const int j = hk[k * ns + i] - (int)sl.seg0;  // partner slot -> its id
      const size_t o = (size_t)id * b->hist_cap + k;
      out->hist_key[o] = (j >= 0 && (size_t)j < ns) ? pid[j] : -1;
      const Del3<H> d = hd[k * ns + i];
      out->hist_delta[o * 3] = (double)d.x;
      out->hist_delta[o * 3 + 1] = (double)d.y;
      out->hist_delta[o * 3 + 2] = (double)d.z;
    }
    out->nwhist[id] = nw[i];
    for (int k = 0; k < nw[i]; ++k) {
      const size_t o = (size_t)id * b->whist_cap + k;
      out->whist_key[o] = wk[k * ns + i] - (int)sl.tri0;
      const Del3<H> d = wd[k * ns + i];
      out->whist_delta[o * 3] = (double)d.x;
      out->whist_delta[o * 3 + 1] = (double)d.y;
      out->whist_delta[o * 3 + 2] = (double)d.z;
    }
  }
  return GDEM_OK;
}

// Histories into a scene admitted and not stepped yet: its slots still hold
// the particles in their own order (slot seg0 + id, buffer parity 0), so
// the rows are written straight in device form.
template <class H>
int batch_set_history_impl(gdem_batch* b, int sid, const gdem_history_view* in) {
  BBufs<H>& B = bbufs<H>(b);
  const Slot& sl = b->slots[sid];
  const size_t ns = (size_t)sl.nseg, N = (size_t)b->pcap, n = (size_t)sl.n;
  std::vector<int> nh(ns, 0), nw(ns, 0), hk((size_t)b->hist_cap * ns, -1), wk((size_t)b->whist_cap * ns, -1);
  std::vector<Del3<H>> hd((size_t)b->hist_cap * ns, Del3<H>{}), wd((size_t)b->whist_cap * ns, Del3<H>{});
  for (size_t i = 0; i < n; ++i) {
    nh[i] = in->nhist[i];
    for (int k = 0; k < nh[i]; ++k) {
      const size_t o = i * b->hist_cap + k;
      const int j = in->hist_key[o];
      hk[k * ns + i] = (j >= 0 && (size_t)j < n) ? (int)sl.seg0 + j : -1;
      hd[k * ns + i] = Del3<H>{(H)in->hist_delta[o * 3], (H)in->hist_delta[o * 3 + 1], (H)in->hist_delta[o * 3 + 2]};
    }
    nw[i] = in->nwhist[i];
    for (int k = 0; k < nw[i]; ++k) {
      const size_t o = i * b->whist_cap + k;
      wk[k * ns + i] = in->whist_key[o] + (int)sl.tri0;
      wd[k * ns + i] = Del3<H>{(H)in->whist_delta[o * 3], (H)in->whist_delta[o * 3 + 1], (H)in->whist_delta[o * 3 + 2]};
    }
  }
  int rc = 0;
  rc |= upload(b, B.nh[0] + sl.seg0, nh.data(), ns * 4);
  rc |= upload(b, B.nw[0] + sl.seg0, nw.data(), ns * 4);
  for (int k = 0; k < b->hist_cap; ++k) {
    rc |= upload(b, B.hkey[0] + k * N + sl.seg0, hk.data() + k * ns, ns * 4);
    rc |= upload(b, B.hd[0] + k * N + sl.seg0, hd.data() + k * ns, ns * sizeof(Del3<H>));
  }
  for (int k = 0; k < b->whist_cap; ++k) {
    rc |= upload(b, B.wkey[0] + k * N + sl.seg0, wk.data() + k * ns, ns * 4);
    rc |= upload(b, B.wd[0] + k * N + sl.seg0, wd.data() + k * ns, ns * sizeof(Del3<H>));
  }
  if (rc) return fail(b->ctx, GDEM_E_HIP, "batch: history upload failed");
  HIPCHK(b->ctx, hipEventRecord(b->up, b->ctx->stream));
  return GDEM_OK;
}

template <class H>
int batch_counts_impl(gdem_batch* b, int sid, unsigned long long acc[4]) {
  gdem_ctx* c = b->ctx;
  HIPCHK(c, hipMemsetAsync(b->d_acc, 0, 4 * sizeof(unsigned long long), c->stream));
  launch_batch_counts<H>(bbufs<H>(b), sid, (int)b->slots[sid].nseg, b->d_acc, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(acc, b->d_acc, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  return GDEM_OK;
}

bool valid_slot(gdem_batch* b, int sid) { return b && sid >= 0 && sid < b->nslots && b->slots[sid].used; }

}  // namespace
}  // namespace gdem

using namespace gdem;

extern "C" {

int gdem_batch_create(gdem_ctx* c, const gdem_batch_params* bp, gdem_batch** out) {
  if (!c || !bp || !out) return GDEM_E_ARG;
  *out = nullptr;
  if (bp->precision != GDEM_FP64 && bp->precision != GDEM_MIXED)
    return fail(c, GDEM_E_ARG, "batch: unknown precision");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, GDEM_E_HIP, "hipSetDevice");
  std::unique_ptr<gdem_batch> b(new gdem_batch());
  b->ctx = c;
  b->precision = bp->precision;
  b->nslots = bp->max_scenes > 0 ? bp->max_scenes : 256;
  b->nb_cap = bp->nb_cap > 0 ? bp->nb_cap : 48;
  b->wall_cap = bp->wall_cap > 0 ? bp->wall_cap : 24;
  b->hist_cap = bp->hist_cap > 0 ? bp->hist_cap : 16;
  b->whist_cap = bp->whist_cap > 0 ? bp->whist_cap : 8;
  if (b->nb_cap > 64 || b->wall_cap > 48 || b->hist_cap > 64 || b->whist_cap > 64)
    return fail(c, GDEM_E_ARG, "batch: slot capacities exceed 64 (pairs/histories) or 48 (walls)");
  b->graph_steps = bp->graph_steps > 0 ? bp->graph_steps : 0;
  const int hbytes = b->precision == GDEM_FP64 ? 8 : 4;
  b->bytes_per_particle = particle_bytes(hbytes, b->nb_cap, b->wall_cap, b->hist_cap, b->whist_cap);
  int64_t pc = bp->particle_capacity;
  if (pc <= 0) {
    if (bp->hbm_budget <= 0) return fail(c, GDEM_E_ARG, "batch: give particle_capacity or hbm_budget");
    // per particle slot: its own arrays plus the default cell, triangle and
    // wall-table shares
    const int64_t per = b->bytes_per_particle + 3 * kCellBytes + kTriBytes / 16 + 3 * 4 * 4;
    pc = bp->hbm_budget / per;
  }
  const int64_t gran = 256 * kGridGranule;
  pc = (pc + gran - 1) / gran * gran;
  if (pc > ((int64_t)1 << 30)) return fail(c, GDEM_E_ARG, "batch: particle capacity above 2^30");
  b->pcap = pc;
  b->ccap = bp->cell_capacity > 0 ? bp->cell_capacity : 3 * pc + 65536;
  b->tcap = bp->tri_capacity > 0 ? bp->tri_capacity : pc / 16 + 4096;
  b->wcap = bp->wall_capacity > 0 ? bp->wall_capacity : 4 * b->ccap;
  if (b->ccap > ((int64_t)1 << 31) - 1 || b->wcap > ((int64_t)1 << 31) - 1)
    return fail(c, GDEM_E_ARG, "batch: cell or wall-table capacity above 2^31");
  b->parts.reset(pc);
  b->cells.reset(b->ccap);
  b->tris.reset(b->tcap);
  b->wtab.reset(b->wcap);
  b->slots.assign(b->nslots, Slot());
  auto& A = b->arena;
  HIPCHK(c, A.alloc(&b->d_scenes, b->nslots));
  HIPCHK(c, A.alloc(&b->d_sctl, b->nslots));
  HIPCHK(c, A.alloc(&b->d_ctl, BC_WORDS));
  HIPCHK(c, A.alloc(&b->d_wave, (size_t)pc / 64));
  HIPCHK(c, A.alloc(&b->d_due, (size_t)b->nslots * 2));
  HIPCHK(c, A.alloc(&b->d_tri, (size_t)b->tcap));
  HIPCHK(c, A.alloc(&b->d_tbox, (size_t)b->tcap * 6));
  HIPCHK(c, A.alloc(&b->d_wc_off, (size_t)b->ccap));
  HIPCHK(c, A.alloc(&b->d_wc_tri, (size_t)b->wcap));
  HIPCHK(c, A.alloc(&b->d_acc, 4));
  HIPCHK(c, hipMemsetAsync(b->d_sctl, 0, sizeof(SlotCtl) * b->nslots, c->stream));
  HIPCHK(c, hipMemsetAsync(b->d_scenes, 0, sizeof(BScene) * b->nslots, c->stream));
  HIPCHK(c, hipMemsetAsync(b->d_ctl, 0, sizeof(int) * BC_WORDS, c->stream));
  HIPCHK(c, hipMemsetAsync(b->d_wave, 0xff, (size_t)pc / 64 * 4, c->stream));
  const int rc = b->precision == GDEM_FP64 ? alloc_batch<double>(b.get()) : alloc_batch<float>(b.get());
  if (rc) return rc;
  HIPCHK(c, hipEventCreate(&b->t0));
  HIPCHK(c, hipEventCreate(&b->t1));
  HIPCHK(c, hipEventCreate(&b->up));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  b->bytes_total = pc * b->bytes_per_particle + b->ccap * kCellBytes + b->tcap * kTriBytes + b->wcap * 4 +
                   (int64_t)b->nslots * (sizeof(BScene) + sizeof(SlotCtl) + 2 * 16) + pc / 64 * 4;
  *out = b.release();
  return GDEM_OK;
}

void gdem_batch_destroy(gdem_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
}

int gdem_batch_admit(gdem_batch* b, const gdem_particles_desc* pd, const gdem_mesh_desc* md,
                     const gdem_params* pp, int nsteps, int* scene_out) {
  if (!b || !pd || !pp || !scene_out || nsteps < 0) return GDEM_E_ARG;
  gdem_ctx* c = b->ctx;
  *scene_out = -1;
  if (pd->n <= 0 || !pd->x || !pd->r || !pd->m || !pd->k || !pd->etaconst)
    return fail(c, GDEM_E_ARG, "particles: n, x, r, m, k and etaconst are required");
  if (!(pp->dt > 0) || !(pp->cell > 0) || !(pp->skin >= 0) || !(pp->refresh >= 0))
    return fail(c, GDEM_E_ARG, "params: dt, cell must be > 0, skin and refresh >= 0");
  if (b->step + nsteps > ((int64_t)1 << 31) - 2) return fail(c, GDEM_E_ARG, "batch: step counter would overflow");
  BScene S{};
  S.dt = pp->dt;
  S.mu = pp->mu;
  double maxr = 0;
  for (int i = 0; i < pd->n; ++i) maxr = pd->r[i] > maxr ? pd->r[i] : maxr;
  for (int a = 0; a < 3; ++a) {
    S.g[a] = pp->g[a];
    S.bmin[a] = pp->bbox[2 * a];
    S.bmax[a] = pp->bbox[2 * a + 1];
    if (!(S.bmax[a] > S.bmin[a])) return fail(c, GDEM_E_ARG, "params: empty bounding box");
    S.invd[a] = 1. / pp->cell;
    // BoundingBox.cu:394-396
    const double cells = ceil((S.bmax[a] - S.bmin[a]) / pp->cell) + 2;
    if (cells > 1e6) return fail(c, GDEM_E_ARG, "params: cell grid too fine");
    S.size[a] = (int)cells;
  }
  const double ncell_d = (double)S.size[0] * S.size[1] * S.size[2];
  if (ncell_d + 1 > (double)b->ccap) return fail(c, GDEM_E_ARG, "scene: cell grid larger than the batch's cell capacity");
  if (pp->cell < 2 * maxr + pp->skin) return fail(c, GDEM_E_ARG, "params: cell must be >= 2*max(r) + skin");
  S.skin = pp->skin;
  S.threshSq = pp->refresh * pp->refresh;
  for (int i = 0; i < pd->n; ++i) {
    if (pd->is_active && pd->is_active[i] != 1) continue;
    if (!(pd->r[i] > 0)) return fail(c, GDEM_E_ARG, "particles: radius must be > 0");
    for (int a = 0; a < 3; ++a) {
      const double q = pd->x[3 * i + a];
      if (!(q >= S.bmin[a] && q <= S.bmax[a]))
        return fail(c, GDEM_E_ARG, "particles: active particle outside the bounding box");
    }
  }
  const int ntri = md ? md->ntri : 0;
  if (ntri > 0 && !md->verts) return fail(c, GDEM_E_ARG, "mesh: verts missing");
  int sid = -1;
  for (int s = 0; s < b->nslots; ++s)
    if (!b->slots[s].used) {
      sid = s;
      break;
    }
  if (sid < 0) return fail(c, GDEM_E_BUSY, "batch: no free scene slot");
  MeshHost mesh;
  if (!build_mesh(ntri > 0 ? md->verts : nullptr, ntri, md ? md->tris_per_file : nullptr, md ? md->nfiles : 0,
                  &mesh))
    return fail(c, GDEM_E_ARG, "mesh: tris_per_file does not add up to ntri");
  Slot sl;
  sl.n = pd->n;
  sl.nseg = (pd->n + kSlotGranule - 1) / kSlotGranule * kSlotGranule;
  sl.ncell1 = (int64_t)ncell_d + 1;
  sl.ntri = ntri;
  // a scene larger than the batch's whole capacity can never be admitted:
  // GDEM_E_OOM, not the retry-later GDEM_E_BUSY
  if (sl.nseg > b->pcap || sl.ntri > b->tcap)
    return fail(c, GDEM_E_OOM, "batch: scene larger than the batch's total particle or triangle capacity");
  // reserve the ranges (all or nothing)
  bool ok = b->parts.alloc(sl.nseg, &sl.seg0);
  if (ok && !b->cells.alloc(sl.ncell1, &sl.cell0)) {
    b->parts.release(sl.seg0, sl.nseg);
    ok = false;
  }
  if (ok && !b->tris.alloc(sl.ntri, &sl.tri0)) {
    b->parts.release(sl.seg0, sl.nseg);
    b->cells.release(sl.cell0, sl.ncell1);
    ok = false;
  }
  if (!ok) return fail(c, GDEM_E_BUSY, "batch: not enough free particle, cell or triangle capacity");
  auto undo = [&]() {
    b->parts.release(sl.seg0, sl.nseg);
    b->cells.release(sl.cell0, sl.ncell1);
    b->tris.release(sl.tri0, sl.ntri);
  };
  // static per-cell wall table of this scene's grid, in global cell and
  // triangle ids
  std::vector<int> off((size_t)sl.ncell1, 0), wtri;
  if (ntri > 0) {
    WallCells wc;
    if (!build_wall_cells(mesh, S.bmin, pp->cell, S.size, maxr + pp->skin, (size_t)b->wcap, &wc)) {
      undo();
      return fail(c, GDEM_E_OOM, "batch: wall table larger than the batch's total wall capacity");
    }
    sl.nwt = (int64_t)wc.tri.size();
    if (!b->wtab.alloc(sl.nwt, &sl.wt0)) {
      undo();
      return fail(c, GDEM_E_BUSY, "batch: not enough free wall-table capacity");
    }
    for (size_t k = 0; k < off.size(); ++k) off[k] = (int)(wc.off[k] + sl.wt0);
    wtri.resize(wc.tri.size());
    for (size_t k = 0; k < wtri.size(); ++k) wtri[k] = (int)(wc.tri[k] + sl.tri0);
  }
  sl.used = true;
  sl.start = b->step;
  sl.end = b->step + nsteps;
  sl.bytes = sl.nseg * b->bytes_per_particle + sl.ncell1 * kCellBytes + sl.ntri * kTriBytes + sl.nwt * 4;
  S.cell_base = (int)sl.cell0;
  S.ncell = (int)(sl.ncell1 - 1);
  S.seg0 = (int)sl.seg0;
  S.nseg = (int)sl.nseg;
  S.n = pd->n;
  S.ntri = ntri;
  S.wall_table = ntri > 0 ? 1 : 0;
  S.start_step = (int)sl.start;
  S.end_step = (int)sl.end;
  int rc = 0;
  rc |= upload(b, b->d_scenes + sid, &S, sizeof(S));
  if (ntri > 0) {
    // triangles in the batch's global ids: geometry and boxes
    std::vector<double> tb((size_t)ntri * 6);
    for (int t = 0; t < ntri; ++t)
      for (int a = 0; a < 3; ++a) {
        tb[(size_t)t * 6 + a] = mesh.aabb_lo[(size_t)t * 3 + a];
        tb[(size_t)t * 6 + 3 + a] = mesh.aabb_hi[(size_t)t * 3 + a];
      }
    rc |= upload(b, b->d_tri + sl.tri0, mesh.tri.data(), (size_t)ntri * sizeof(TriGeo));
    rc |= upload(b, b->d_tbox + sl.tri0 * 6, tb.data(), tb.size() * 8);
    rc |= upload(b, b->d_wc_off + sl.cell0, off.data(), off.size() * 4);
    rc |= upload(b, b->d_wc_tri + sl.wt0, wtri.data(), wtri.size() * 4);
  }
  rc |= b->precision == GDEM_FP64 ? upload_particles<double>(b, sl, pd) : upload_particles<float>(b, sl, pd);
  SlotCtl ctl{};
  rc |= upload(b, b->d_sctl + sid, &ctl, sizeof(ctl));
  std::vector<int> wave((size_t)(sl.nseg / kSlotGranule), sid);
  rc |= upload(b, b->d_wave + sl.seg0 / kSlotGranule, wave.data(), wave.size() * 4);
  if (!rc) {
    // the lists are built by the first batch step's rebuild phase
    const int par = (int)((b->step + 1) & 1);
    if (b->precision == GDEM_FP64) launch_batch_admit<double>(b->b64, sid, par, c->stream);
    else launch_batch_admit<float>(b->b32, sid, par, c->stream);
    if (hipGetLastError() != hipSuccess) rc |= GDEM_E_HIP;
  }
  if (rc) {
    (void)hipStreamSynchronize(c->stream);
    undo();
    b->wtab.release(sl.wt0, sl.nwt);
    return fail(c, GDEM_E_HIP, "batch: upload failed");
  }
  HIPCHK(c, hipEventRecord(b->up, c->stream));
  b->slots[sid] = sl;
  *scene_out = sid;
  return GDEM_OK;
}

int gdem_batch_step(gdem_batch* b, int nsteps) {
  if (!b || nsteps < 0) return GDEM_E_ARG;
  if (nsteps == 0) return GDEM_OK;
  if (b->step + nsteps > ((int64_t)1 << 31) - 2) return fail(b->ctx, GDEM_E_ARG, "batch: step counter would overflow");
  return b->precision == GDEM_FP64 ? batch_step_impl<double>(b, nsteps) : batch_step_impl<float>(b, nsteps);
}

int gdem_batch_poll(gdem_batch* b, int* ids, int max, int* n_out) {
  if (!b || !n_out || (max > 0 && !ids)) return GDEM_E_ARG;
  int n = 0;
  for (int s = 0; s < b->nslots; ++s)
    if (b->slots[s].used && b->slots[s].end <= b->step) {
      if (n < max) ids[n] = s;
      ++n;
    }
  *n_out = n;
  return GDEM_OK;
}

int gdem_batch_download(gdem_batch* b, int sid, gdem_particles_out* out) {
  if (!valid_slot(b, sid) || !out) return GDEM_E_ARG;
  return b->precision == GDEM_FP64 ? batch_download_impl<double>(b, sid, out)
                                   : batch_download_impl<float>(b, sid, out);
}

int gdem_batch_history(gdem_batch* b, int sid, const gdem_history_out* out) {
  if (!valid_slot(b, sid) || !out || !out->nhist || !out->hist_key || !out->hist_delta || !out->nwhist ||
      !out->whist_key || !out->whist_delta)
    return GDEM_E_ARG;
  return b->precision == GDEM_FP64 ? batch_history_impl<double>(b, sid, out) : batch_history_impl<float>(b, sid, out);
}

int gdem_batch_set_history(gdem_batch* b, int sid, const gdem_history_view* in) {
  if (!valid_slot(b, sid) || !in || !in->nhist || !in->hist_key || !in->hist_delta || !in->nwhist ||
      !in->whist_key || !in->whist_delta)
    return GDEM_E_ARG;
  const Slot& sl = b->slots[sid];
  if (b->step != sl.start)
    return fail(b->ctx, GDEM_E_STATE, "batch: histories are set between admission and the next batch step");
  for (int i = 0; i < sl.n; ++i) {
    if (in->nhist[i] < 0 || in->nhist[i] > b->hist_cap || in->nwhist[i] < 0 || in->nwhist[i] > b->whist_cap)
      return fail(b->ctx, GDEM_E_SLOT_OVERFLOW, "history row longer than the batch's capacities");
    for (int k = 0; k < in->nwhist[i]; ++k) {
      const int t = in->whist_key[(size_t)i * b->whist_cap + k];
      if (t < 0 || t >= sl.ntri) return fail(b->ctx, GDEM_E_ARG, "wall history key is not a triangle of the scene");
    }
  }
  return b->precision == GDEM_FP64 ? batch_set_history_impl<double>(b, sid, in)
                                   : batch_set_history_impl<float>(b, sid, in);
}

int gdem_batch_scene_status(gdem_batch* b, int sid, gdem_batch_scene_stats* st) {
  if (!valid_slot(b, sid) || !st) return GDEM_E_ARG;
  unsigned long long acc[4];
  const int rc = b->precision == GDEM_FP64 ? batch_counts_impl<double>(b, sid, acc)
                                           : batch_counts_impl<float>(b, sid, acc);
  if (rc) return rc;
  SlotCtl ctl;
  HIPCHK(b->ctx, hipMemcpy(&ctl, b->d_sctl + sid, sizeof(ctl), hipMemcpyDeviceToHost));
  const Slot& sl = b->slots[sid];
  st->n = sl.n;
  st->start_step = sl.start;
  st->end_step = sl.end;
  st->steps_done = std::min(b->step, sl.end) - sl.start;
  st->rebuilds = ctl.rebuilds;
  st->errors = ctl.err;
  st->n_active = (int)acc[0];
  const double na = acc[0] ? (double)acc[0] : 1.0;
  st->mean_nb = acc[1] / na;
  st->mean_wall = acc[2] / na;
  st->mean_contacts = acc[3] / na;
  st->bytes = sl.bytes;
  if (ctl.err & (GDEM_ERR_NB_OVERFLOW | GDEM_ERR_WALL_OVERFLOW | GDEM_ERR_HIST_OVERFLOW |
                 GDEM_ERR_WHIST_OVERFLOW | GDEM_ERR_VE_OVERFLOW))
    return fail(b->ctx, GDEM_E_SLOT_OVERFLOW, "slot overflow (raise the batch capacities)");
  if (ctl.err & GDEM_ERR_NAN) return fail(b->ctx, GDEM_E_NAN, "non-finite particle state");
  return GDEM_OK;
}

int gdem_batch_evict(gdem_batch* b, int sid) {
  if (!valid_slot(b, sid)) return GDEM_E_ARG;
  gdem_ctx* c = b->ctx;
  Slot& sl = b->slots[sid];
  std::vector<int> wave((size_t)(sl.nseg / kSlotGranule), -1);
  SlotCtl ctl{};
  int rc = 0;
  rc |= upload(b, b->d_wave + sl.seg0 / kSlotGranule, wave.data(), wave.size() * 4);
  rc |= upload(b, b->d_sctl + sid, &ctl, sizeof(ctl));
  // a rebuild the scene requested at its last step must not run
  if (!rc) {
    if (b->precision == GDEM_FP64) launch_batch_evict<double>(b->b64, sid, c->stream);
    else launch_batch_evict<float>(b->b32, sid, c->stream);
    if (hipGetLastError() != hipSuccess) rc |= GDEM_E_HIP;
  }
  if (rc) return fail(c, GDEM_E_HIP, "batch: eviction upload failed");
  HIPCHK(c, hipEventRecord(b->up, c->stream));
  b->parts.release(sl.seg0, sl.nseg);
  b->cells.release(sl.cell0, sl.ncell1);
  b->tris.release(sl.tri0, sl.ntri);
  b->wtab.release(sl.wt0, sl.nwt);
  sl = Slot();
  return GDEM_OK;
}

int gdem_batch_info(gdem_batch* b, int64_t* step, int* live, int64_t* free_particles, int64_t* bytes) {
  if (!b) return GDEM_E_ARG;
  if (step) *step = b->step;
  if (live) {
    int n = 0;
    for (auto& s : b->slots) n += s.used ? 1 : 0;
    *live = n;
  }
  if (free_particles) *free_particles = b->parts.free_total();
  if (bytes) *bytes = b->bytes_total;
  return GDEM_OK;
}

int gdem_batch_timing(gdem_batch* b, double* ms) {
  if (!b || !ms) return GDEM_E_ARG;
  HIPCHK(b->ctx, hipEventSynchronize(b->t1));
  float m = 0;
  HIPCHK(b->ctx, hipEventElapsedTime(&m, b->t0, b->t1));
  *ms = m;
  return GDEM_OK;
}

}  // extern "C"
