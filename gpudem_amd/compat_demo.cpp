// A minimal C++ caller of include/gdem_compat.h, shaped like the reference's
// main.cu step loop (main.cu:206-373): host particle arrays with the
// ParticleSys<HostMemory> member names, a BoundingBox-like grid, the walls
// from ASCII STL files, then gdem_setup / gdem_advance / gdem_fetch /
// gdem_teardown in place of the device calls.
//
// usage: gdem_compat_demo <input.txt> <walls.stl> <output.txt>
// input.txt: one header line
//   N steps out_every precision dt mu gx gy gz minx maxx miny maxy minz maxz dx skin refresh
// then N lines "x y z vx vy vz wx wy wz r m k etaconst active" (%.17g).
// output.txt: N lines "x y z vx vy vz wx wy wz active" after `steps` steps,
// fetched every `out_every` steps as main.cu's output loop does.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../include/gdem_compat.h"

namespace {

// the ParticleSys<HostMemory> members gdem_compat reads and writes
struct HostParticles {
  int N = 0;
  double g[3] = {0, 0, 0};
  std::vector<double> xs, vs, ws, rs, ms, ks, es;
  std::vector<int> act;
  double *x = nullptr, *v = nullptr, *angv = nullptr, *r = nullptr, *m = nullptr, *k = nullptr,
         *etaconst = nullptr;
  int* isActive = nullptr;
  void bind() {
    x = xs.data();
    v = vs.data();
    angv = ws.data();
    r = rs.data();
    m = ms.data();
    k = ks.data();
    etaconst = es.data();
    isActive = act.data();
  }
};

// the BoundingBox members gdem_compat reads
struct Box {
  double minx, maxx, miny, maxy, minz, maxz, dx, skinR, refreshThresh;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s input.txt walls.stl output.txt\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "r");
  if (!in) return 2;
  HostParticles ps;
  Box box{};
  int steps = 0, out_every = 0, precision = GDEM_FP64;
  double dt = 0, mu = 0;
  if (std::fscanf(in, "%d %d %d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &ps.N, &steps,
                  &out_every, &precision, &dt, &mu, &ps.g[0], &ps.g[1], &ps.g[2], &box.minx, &box.maxx, &box.miny,
                  &box.maxy, &box.minz, &box.maxz, &box.dx, &box.skinR, &box.refreshThresh) != 18)
    return 2;
  const int N = ps.N;
  ps.xs.resize(3 * N);
  ps.vs.resize(3 * N);
  ps.ws.resize(3 * N);
  ps.rs.resize(N);
  ps.ms.resize(N);
  ps.ks.resize(N);
  ps.es.resize(N);
  ps.act.resize(N);
  for (int i = 0; i < N; ++i) {
    double* f[13] = {&ps.xs[3 * i], &ps.xs[3 * i + 1], &ps.xs[3 * i + 2], &ps.vs[3 * i], &ps.vs[3 * i + 1],
                     &ps.vs[3 * i + 2], &ps.ws[3 * i], &ps.ws[3 * i + 1], &ps.ws[3 * i + 2], &ps.rs[i],
                     &ps.ms[i], &ps.ks[i], &ps.es[i]};
    for (double* p : f)
      if (std::fscanf(in, "%lf", p) != 1) return 2;
    if (std::fscanf(in, "%d", &ps.act[i]) != 1) return 2;
  }
  std::fclose(in);
  ps.bind();
  try {
    gdem_compat::GdemRun run = gdem_compat::gdem_setup(&ps, {std::string(argv[2])}, &box, dt, mu, precision);
    for (int done = 0; done < steps;) {
      const int n = std::min(out_every, steps - done);
      gdem_compat::gdem_advance(run, n);
      gdem_compat::gdem_fetch(run, &ps);  // the reference writes its output files here
      done += n;
    }
    gdem_compat::gdem_teardown(run);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gdem_compat_demo: %s\n", e.what());
    return 1;
  }
  FILE* out = std::fopen(argv[3], "w");
  if (!out) return 2;
  for (int i = 0; i < N; ++i)
    std::fprintf(out, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d\n", ps.x[3 * i], ps.x[3 * i + 1],
                 ps.x[3 * i + 2], ps.v[3 * i], ps.v[3 * i + 1], ps.v[3 * i + 2], ps.angv[3 * i], ps.angv[3 * i + 1],
                 ps.angv[3 * i + 2], ps.isActive[i]);
  std::fclose(out);
  return 0;
}
