/*
 * gdem_compat.h -- the reference's step-loop calls on top of the gdem C ABI.
 *
 * bi-rugaoishii/gpudem's main.cu drives its device path through free
 * functions on ParticleSys<DeviceMemory>, BoundingBox, TriangleMesh and BVH:
 * allocate / copyToDevice (ParticleSystem.h:78, 91), initialize_BoundingBox /
 * copyToDeviceBox (BoundingBox.h:160, 179), initializeBVH / copyToDeviceBVH
 * (BVH.h:47, 54), deviceMallocCopyTriangleMesh (TriangleMesh.h:184), then
 * device_dem_verlet_verlet once per step (device_dem.h:112) and
 * copyFromDevice before each output (ParticleSystem.h:99).  The four
 * functions below replace those calls one for one.  They are templates over
 * the caller's particle and box types, so the header needs none of the
 * reference's headers: any type with the ParticleSys<HostMemory> members
 * (N, x, v, angv, r, m, k, etaconst, isActive, g) and the BoundingBox members
 * (minx..maxz, dx, skinR, refreshThresh) works.  Header-only C++17; link
 * libgdem.so.
 */
#ifndef GDEM_COMPAT_H_
#define GDEM_COMPAT_H_

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gdem.h"

namespace gdem_compat {

struct GdemRun {
  gdem_ctx* ctx = nullptr;
  gdem_scene* scene = nullptr;
};

inline void check(int rc, gdem_ctx* ctx, const char* what) {
  if (rc != GDEM_OK)
    throw std::runtime_error(std::string(what) + ": " + (ctx ? gdem_last_error(ctx) : "no context"));
}

// Facet vertices of ASCII STL files in file order, one triangle count per
// file: the input load_ascii_stl_double (TriangleMesh.cu:363-461) reads
// before its per-file vertex merge, which gdem repeats itself together with
// the triangle precompute.
inline void read_stl_facets(const std::vector<std::string>& paths, std::vector<double>& verts,
                            std::vector<int>& per_file) {
  for (const std::string& p : paths) {
    FILE* fp = std::fopen(p.c_str(), "r");
    if (!fp) throw std::runtime_error("cannot open STL file " + p);
    char line[512];
    int n = 0;
    double x, y, z;
    while (std::fgets(line, sizeof line, fp))
      if (std::sscanf(line, " vertex %lf %lf %lf", &x, &y, &z) == 3) {
        verts.push_back(x);
        verts.push_back(y);
        verts.push_back(z);
        ++n;
      }
    std::fclose(fp);
    per_file.push_back(n / 3);
  }
}

// Replaces allocate + copyToDevice + initialize_BoundingBox + copyToDeviceBox
// + initializeBVH/buildBVH/copyToDeviceBVH + deviceMallocCopyTriangleMesh
// (main.cu:206-222).  The grid, skin and refresh distance are the box's
// (BoundingBox.cu:381-388); precision is GDEM_FP64 (the reference's
// arithmetic, bit for bit) or GDEM_MIXED.
template <class PS, class Box>
GdemRun gdem_setup(const PS* ps, const std::vector<std::string>& wall_files, const Box* box, double dt, double mu,
                   int precision, int device = 0) {
  GdemRun run;
  check(gdem_ctx_create(device, &run.ctx), nullptr, "gdem_ctx_create");
  std::vector<double> verts;
  std::vector<int> per_file;
  read_stl_facets(wall_files, verts, per_file);
  gdem_mesh_desc m{(int)(verts.size() / 9), verts.data(), (int)per_file.size(), per_file.data()};
  gdem_particles_desc d{ps->N, ps->x, ps->v, ps->angv, ps->r, ps->m, ps->k, ps->etaconst, ps->isActive};
  gdem_params prm;
  std::memset(&prm, 0, sizeof prm);
  prm.dt = dt;
  prm.mu = mu;
  for (int a = 0; a < 3; ++a) prm.g[a] = ps->g[a];
  prm.bbox[0] = box->minx;
  prm.bbox[1] = box->maxx;
  prm.bbox[2] = box->miny;
  prm.bbox[3] = box->maxy;
  prm.bbox[4] = box->minz;
  prm.bbox[5] = box->maxz;
  prm.cell = box->dx;
  prm.skin = box->skinR;
  prm.refresh = box->refreshThresh;
  prm.precision = precision;
  const int rc = gdem_scene_create(run.ctx, &d, &m, &prm, &run.scene);
  if (rc != GDEM_OK) {
    const std::string msg = gdem_last_error(run.ctx);
    gdem_ctx_destroy(run.ctx);
    throw std::runtime_error("gdem_scene_create: " + msg);
  }
  return run;
}

// Replaces `steps` calls of device_dem_verlet_verlet (device_dem.h:112,
// main.cu:263, 270) between two outputs; a slot overflow or a non-finite
// state raises here instead of corrupting memory.
inline void gdem_advance(GdemRun& run, int steps) {
  check(gdem_scene_step(run.scene, steps), run.ctx, "gdem_scene_step");
  gdem_stats st;
  check(gdem_scene_status(run.scene, &st), run.ctx, "gdem_scene_status");
}

// Replaces copyFromDevice(&dps, &ps) before an output (main.cu:278-279):
// positions, velocities, spins and activity in original particle order.
template <class PS>
void gdem_fetch(GdemRun& run, PS* ps) {
  gdem_particles_out o{ps->x, ps->v, ps->angv, ps->isActive};
  check(gdem_scene_download(run.scene, &o), run.ctx, "gdem_scene_download");
}

// Replaces deallocate + the free_* calls (main.cu:358-373).
inline void gdem_teardown(GdemRun& run) {
  gdem_scene_destroy(run.scene);
  gdem_ctx_destroy(run.ctx);
  run.scene = nullptr;
  run.ctx = nullptr;
}

}  // namespace gdem_compat

#endif /* GDEM_COMPAT_H_ */
