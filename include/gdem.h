/*
 * gdem -- MI355X-native DEM contact engine, C ABI.
 *
 * This is the drop-in boundary for the per-step sphere contact path of
 * bi-rugaoishii/gpudem.  Each entry point names the reference interface it
 * replaces.  All pointers are plain host pointers borrowed for the duration of
 * the call; device memory is owned by the library.  Every function returns
 * GDEM_OK (0) or a negative GDEM_E_* code; no exceptions cross the ABI.
 * A context is single-threaded (callers synchronise externally); distinct
 * contexts are independent.
 */
#ifndef GDEM_H_
#define GDEM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDEM_OK 0
#define GDEM_E_ARG (-1)           /* invalid argument or shape */
#define GDEM_E_HIP (-2)           /* HIP runtime error (see gdem_last_error) */
#define GDEM_E_OOM (-3)           /* device allocation failed or over budget */
#define GDEM_E_SLOT_OVERFLOW (-4) /* a neighbour / contact slot row overflowed */
#define GDEM_E_NAN (-5)           /* non-finite particle state detected */
#define GDEM_E_STATE (-6)         /* call not valid in the object's state */
#define GDEM_E_BUSY (-7)          /* no free scene slot or capacity now: retry after an eviction */

/* precision of the contact arithmetic */
#define GDEM_FP64 0  /* plain FP64, bit-identical to the reference */
#define GDEM_MIXED 1 /* FP32 forces/history, FP64 relative position, overlap and integration */

/* bits of the error word reported by gdem_scene_status */
#define GDEM_ERR_NB_OVERFLOW 1u
#define GDEM_ERR_WALL_OVERFLOW 2u
#define GDEM_ERR_HIST_OVERFLOW 4u
#define GDEM_ERR_WHIST_OVERFLOW 8u
#define GDEM_ERR_VE_OVERFLOW 16u
#define GDEM_ERR_NAN 32u
#define GDEM_ERR_BARRIER 64u /* a grid barrier of the list rebuild timed out and was cancelled;
                                the scene then rebuilds through per-phase kernels (reported once
                                as GDEM_E_HIP by gdem_scene_status).  The environment variable
                                GDEM_GBAR_SPIN, read at scene creation, lowers the barrier's poll
                                limit to exercise this path. */

typedef struct gdem_ctx gdem_ctx;
typedef struct gdem_scene gdem_scene;

/* Triangle soup as read from one or more ASCII STL files: facet vertices in
 * file order.  Vertices are merged per file on a 1e-9 grid and edge ids are
 * global, as TriangleMesh.cu:363-752 does. */
typedef struct {
  int ntri;
  const double* verts;      /* [ntri*9]: v0xyz v1xyz v2xyz of each facet */
  int nfiles;               /* 0 means one file */
  const int* tris_per_file; /* [nfiles] or NULL */
} gdem_mesh_desc;

/* Host particle arrays in the reference's ParticleSys<HostMemory> layout
 * (ParticleSystemMember_common.def): xyz triples are AoS [n*3]. */
typedef struct {
  int n;
  const double* x;        /* [n*3] positions */
  const double* v;        /* [n*3] velocities, NULL = zero */
  const double* angv;     /* [n*3] angular velocities, NULL = zero */
  const double* r;        /* [n] radii */
  const double* m;        /* [n] masses */
  const double* k;        /* [n] normal stiffness */
  const double* etaconst; /* [n] damping constant */
  const int* is_active;   /* [n], NULL = all active */
} gdem_particles_desc;

/* Step parameters.  The reference derives cell = 4*max r, skin = max r,
 * refresh = 0.5*max r (BoundingBox.cu:381-388); gdem_params_default fills
 * exactly that. */
typedef struct {
  double dt;
  double mu;        /* Coulomb friction coefficient */
  double g[3];      /* gravity */
  double bbox[6];   /* minx, maxx, miny, maxy, minz, maxz */
  double cell;      /* cell-grid spacing */
  double skin;      /* Verlet skin for pair and wall candidates */
  double refresh;   /* displacement that triggers a list rebuild */
  int precision;    /* GDEM_FP64 or GDEM_MIXED */
  int nb_cap;       /* pair candidates per particle (0 = 48) */
  int wall_cap;     /* wall candidates per particle (0 = 24) */
  int hist_cap;     /* pair contact history slots (0 = 16) */
  int whist_cap;    /* wall contact history slots (0 = 8) */
  int graph_steps;  /* 0: direct launches; > 0: steps captured per hipGraph (the list
                       rebuild is one grid-barrier kernel either way);
                       < 0: list rebuild as a chain of per-phase kernels, direct launches
                       (per-phase profiling) */
  int wall_broadphase; /* 0: static per-cell triangle table when it fits, else BVH;
                          1: BVH traversal; 2: per-cell table (error if it does not fit) */
} gdem_params;

/* Host download view (original particle order). Any pointer may be NULL. */
typedef struct {
  double* x;      /* [n*3] */
  double* v;      /* [n*3] */
  double* angv;   /* [n*3] */
  int* is_active; /* [n] */
} gdem_particles_out;

/* Run statistics. */
typedef struct {
  int64_t steps;
  int64_t rebuilds;
  int n_active;
  uint32_t errors;     /* GDEM_ERR_* bits */
  double mean_nb;      /* mean pair-candidate list length over active particles */
  double mean_wall;    /* mean wall-candidate list length */
  double mean_contacts;/* mean live pair + wall history slots per active particle */
} gdem_stats;

/* Debug views (original particle ids / triangle ids) for parity checks. */
typedef struct {
  int* cellx;           /* [n*3] cell coordinates as d_calcCellId computes them */
  int* cell_id;         /* [n] linear cell id */
  uint32_t* morton;     /* [n] morton3D of the cell coordinates (10-bit masks) */
  int* nb;              /* [n*nb_cap] pair candidates as original ids, ascending */
  int* nnb;             /* [n] */
  int* wall;            /* [n*wall_cap] wall candidates (triangle ids), ascending */
  int* nwall;           /* [n] */
  int* hist_key;        /* [n*hist_cap] pair history keys (partner original ids) in slot order */
  double* hist_delta;   /* [n*hist_cap*3] */
  int* nhist;           /* [n] */
  int* whist_key;       /* [n*whist_cap] */
  double* whist_delta;  /* [n*whist_cap*3] */
  int* nwhist;          /* [n] */
} gdem_debug_view;

/* What the last list rebuild stored, slot by slot (slots are the engine's
 * internal order after the stable radix sort): the sort key (morton3 of the
 * cell coordinates d_calcCellId gives, BoundingBox.cu:1007-1023, 1089-1104, or
 * the cell id when an axis has more than 1024 cells; 0xffffffff for an
 * inactive particle), the original id the slot holds (the order
 * cub::DeviceRadixSort::SortPairs leaves, BoundingBox.cu:1235-1242, applied to
 * the slot order of the previous rebuild), the slot's cell id and the (first,
 * end) slot range the cell table holds for that cell (the CSR of
 * dk_build_cellCount / ExclusiveSum / dk_build_pList, BoundingBox.cu:1062-1087,
 * 1257-1280).  Any pointer may be NULL. */
typedef struct {
  uint32_t* key;    /* [n] */
  int* order;       /* [n] original particle id of each slot */
  int* cell;        /* [n] cell id of each slot, -1 for an inactive particle */
  int* cell_first;  /* [n] first slot of the slot's cell, -1 for an inactive particle */
  int* cell_end;    /* [n] one past the last slot of the slot's cell */
  int* n_sorted;    /* [1] active slots (they come first) */
} gdem_order_view;

/* Per-contact record of one step (pairs ascending by partner id, then walls
 * ascending by triangle id): [n*rec_cap] rows. */
typedef struct {
  int rec_cap;
  int* count;    /* [n] */
  int* key;      /* partner original id or triangle id */
  int* kind;     /* 0 pair, 1 wall */
  double* fn;    /* [n*rec_cap*3] normal force */
  double* ft;    /* [n*rec_cap*3] tangential force */
  double* delta; /* [n*rec_cap] overlap */
  double* accel;     /* [n*3] or NULL: linear acceleration of the step (f/m + g, k_integrate) */
  double* ang_accel; /* [n*3] or NULL: angular acceleration of the step (mom/I) */
} gdem_contact_view;

/* Contact histories in original particle order, as gdem_scene_debug returns
 * them (the SAVE_ON history members deltHis*, indHis*, numCont* of
 * ParticleSystemMember_common.def:26-35). */
typedef struct {
  const int* nhist;          /* [n] */
  const int* hist_key;       /* [n*hist_cap] partner original ids in slot order */
  const double* hist_delta;  /* [n*hist_cap*3] */
  const int* nwhist;         /* [n] */
  const int* whist_key;      /* [n*whist_cap] triangle ids */
  const double* whist_delta; /* [n*whist_cap*3] */
} gdem_history_view;

/* Output form of gdem_history_view: rows [n*hist_cap] / [n*whist_cap] in
 * original particle order, pair keys as partner ids, wall keys as triangle
 * ids, displacements [row*3]. */
typedef struct {
  int* nhist;
  int* hist_key;
  double* hist_delta;
  int* nwhist;
  int* whist_key;
  double* whist_delta;
} gdem_history_out;

/* demSettings.json (main.cu:57-135, 227; ParticleSystem.cu:189-352). */
#define GDEM_INLET_SHUFFLE 0
#define GDEM_INLET_FILE 1
typedef struct {
  int n;                /* inlet.numParticle */
  int input_mode;       /* GDEM_INLET_SHUFFLE or GDEM_INLET_FILE */
  double inlet_box[6];  /* shuffle box, or the file's coordinate range */
  double radius, density, cor, mu;
  double young;         /* YoungModulus, 0 when the legacy stiffness "k" is given */
  double k;             /* normal stiffness (ParticleSystem.cu:202 from YoungModulus, or "k") */
  double etaconst;      /* ParticleSystem.cu:301 */
  double m;             /* density*3.14*r^3*4/3 (main.cu:67) */
  double g[3];
  double dt;            /* 2 PI sqrt(m/k)/dtFactor, adjusted to divide outputTiming (main.cu:123-129) */
  int out_step;         /* steps between outputs */
  int steps;            /* (int)(endTime/dt) (main.cu:227) */
  int seed, gpu_on, brute_on;
  int nwalls;           /* walls.filepaths */
} gdem_settings;

/* ---- context ---------------------------------------------------------- */
/* Replaces the implicit CUDA context + legacy stream of main.cu:206-245. */
int gdem_ctx_create(int hip_device, gdem_ctx** out);
void gdem_ctx_destroy(gdem_ctx* ctx);
int gdem_sync(gdem_ctx* ctx);
const char* gdem_last_error(gdem_ctx* ctx);
/* Architecture name of a HIP device (gcnArchName, e.g. "gfx950:sramecc+:xnack-")
 * and its compute-unit count; no counterpart in the reference (the autotuner
 * keys its cached choices by it). */
int gdem_device_arch(int hip_device, char* buf, int len, int* cus);

/* ---- set-up helpers ----------------------------------------------------- */
/* Material of one particle: m = density*3.14*r^3*4/3 (main.cu:67),
 * k = 4/3*(E/2)*sqrt(r/2)*sqrt(0.05 r) (ParticleSystem.cu:202),
 * etaconst = -2 ln(e) sqrt(k/(3.1415^2+ln(e)^2)) (ParticleSystem.cu:301). */
int gdem_material(int n, const double* r, double density, double young, double cor,
                  double* m_out, double* k_out, double* etaconst_out);
/* calc_BoundingBoxLimits (BoundingBox.cu:448-478): min/max of the mesh and
 * inlet boxes, each scaled by 1.1. */
int gdem_bbox_rule(const gdem_mesh_desc* mesh, const double inlet[6], double bbox_out[6]);
/* cell = 4 max r, skin = max r, refresh = 0.5 max r (BoundingBox.cu:381-388). */
int gdem_params_default(double max_r, gdem_params* p);

/* Host-only triangle precompute (TriangleMesh.cu:536-677), for inspection and
 * parity checks; needs no GPU.  geo_out[ntri*29] holds per triangle
 * n[3], d, p0[3], p1[3], p2[3], e01[3], e02[3], e12[3], d00, d01, d11, denom,
 * d00inv, d11inv, d22inv; ids_out[ntri*6] holds the vertex ids v0<v1<v2 and
 * the edge ids e01, e12, e02 (+3*ntri); order_out[ntri] the Morton order of
 * the centroids (stable, as radixSortUint32 leaves it). Any output may be NULL. */
int gdem_mesh_precompute(const gdem_mesh_desc* mesh, double* geo_out, int* ids_out, int* order_out);

/* ---- plain scene --------------------------------------------------------- */
/* Replaces allocate(ParticleSys<DeviceMemory>*) + copyToDevice
 * (ParticleSystem.h:78,91), initialize_BoundingBox + copyToDeviceBox
 * (BoundingBox.h:160,179), initializeBVH/buildBVH/copyToDeviceBVH (BVH.h:47-62),
 * deviceMallocCopyTriangleMesh (TriangleMesh.h:184) and the host list builds of
 * main.cu:151,196.  Copies in; the caller keeps its host arrays. */
int gdem_scene_create(gdem_ctx* ctx, const gdem_particles_desc* particles,
                      const gdem_mesh_desc* mesh, const gdem_params* params, gdem_scene** out);
/* Replaces nsteps calls of device_dem_verlet_verlet (device_dem.h:112,
 * device_dem.cu:1195-1236).  Asynchronous on the context stream. */
int gdem_scene_step(gdem_scene* scene, int nsteps);
/* Waits, then reports errors raised by the kernels (overflow, NaN). */
int gdem_scene_status(gdem_scene* scene, gdem_stats* stats);
/* Replaces copyFromDevice (ParticleSystem.h:99, SAVE_ON members). */
int gdem_scene_download(gdem_scene* scene, gdem_particles_out* out);
/* No counterpart in the reference.  gdem_scene_download without any device
 * allocation, for callers that cannot allocate (code between hipGraph
 * captures, tight serving loops, harnesses that forbid hipMalloc): the
 * gather into original order goes through a fixed staging array that each
 * GPU holds for the life of the process (19 MiB).  Scenes of at most 262,144
 * particles (GDEM_E_ARG beyond).  Calls that overlap on one GPU share the
 * array and race. */
int gdem_scene_download_noalloc(gdem_scene* scene, gdem_particles_out* out);
/* The data of one solver_output frame (write_frame_bin,
 * solver_output.cu:259-316), compacted on the device: the active particles'
 * positions and radii in original id order, each value multiplied by
 * length_factor in FP64 and converted to float32 there, so 16 B per active
 * particle cross to the host instead of the whole state.  xyz_out holds 3 n
 * floats and r_out n floats (n: the scene's particle count); the first
 * *n_active entries are written.  gdem_write_frame_compact writes the file. */
int gdem_scene_frame(gdem_scene* scene, double length_factor, float* xyz_out, float* r_out, int* n_active);
/* Replaces deallocate(ParticleSys<DeviceMemory>*) and the free_* calls.
 * Destroy every scene before the context it was created on. */
void gdem_scene_destroy(gdem_scene* scene);
/* Parity hooks: cell coordinates/keys of the current positions, candidate
 * lists and contact histories, all in original particle ids. */
int gdem_scene_debug(gdem_scene* scene, gdem_debug_view* view);
/* Parity hook: the keys, slot order and cell table the last list rebuild
 * stored (see gdem_order_view). */
int gdem_scene_debug_order(gdem_scene* scene, gdem_order_view* view);
/* Runs one step that also records every contact's forces. */
int gdem_scene_step_record(gdem_scene* scene, gdem_contact_view* view);
/* Replaces the scene's contact histories (snapshot restore: the SAVE_ON
 * history members that copyToDevice carries, ParticleSystem.cu:82-89).
 * GDEM_E_SLOT_OVERFLOW when a row exceeds the scene's capacities. */
int gdem_scene_set_history(gdem_scene* scene, const gdem_history_view* view);
/* Device time of the last gdem_scene_step call (ms) and the summed time of
 * the force-kernel launches it sampled, with the number sampled; measured
 * with HIP events on the context stream. */
int gdem_scene_timing(gdem_scene* scene, double* total_ms, double* force_ms, int* force_launches);
/* every > 0: time the whole call and the force kernel of every `every`-th
 * step (direct launches, no graph); 0 disables timing. */
int gdem_scene_set_timing(gdem_scene* scene, int every);

/* ---- batched and served scenes ------------------------------------------- */
/* No counterpart in the reference, which steps one scene per process.  A
 * batch holds many independent scenes in one set of device arrays sized at
 * creation (an HBM budget); scenes are admitted and evicted while others keep
 * running, and every gdem_batch_step advances all running scenes together in
 * one launch per phase (captured in hipGraphs).  Each scene's trajectory is
 * bit-identical to the same scene run alone through gdem_scene_*. */
typedef struct gdem_batch gdem_batch;

typedef struct {
  int precision;              /* GDEM_FP64 or GDEM_MIXED, for every scene */
  int max_scenes;             /* scene slots (0 = 256) */
  int64_t hbm_budget;         /* device bytes for the batch (used when particle_capacity == 0) */
  int64_t particle_capacity;  /* particle slots over all scenes (each scene rounds up to 64) */
  int64_t cell_capacity;      /* grid cells over all scenes (0 = 3 x particles + 65536) */
  int64_t tri_capacity;       /* triangles over all scenes (0 = particles / 16 + 4096) */
  int64_t wall_capacity;      /* per-cell wall-table entries (0 = 4 x cells) */
  int nb_cap, wall_cap, hist_cap, whist_cap; /* per-particle slots (0 = 48, 24, 16, 8) */
  int graph_steps;            /* > 0: steps per captured hipGraph; 0: direct launches */
} gdem_batch_params;

typedef struct {
  int n;                /* particles */
  int64_t start_step;   /* batch step at admission */
  int64_t end_step;     /* batch step at which the scene stops (start + nsteps) */
  int64_t steps_done;
  int rebuilds;
  uint32_t errors;      /* GDEM_ERR_* bits */
  int n_active;
  double mean_nb, mean_wall, mean_contacts;
  int64_t bytes;        /* device bytes reserved for the scene */
} gdem_batch_scene_stats;

int gdem_batch_create(gdem_ctx* ctx, const gdem_batch_params* params, gdem_batch** out);
/* Destroy every batch before the context it was created on. */
void gdem_batch_destroy(gdem_batch* batch);
/* Admits a scene that will run nsteps batch steps from the next one.  Inputs
 * as for gdem_scene_create (params.precision, capacities and graph_steps are
 * the batch's).  GDEM_E_BUSY when no slot or capacity is free now;
 * GDEM_E_OOM when the scene exceeds the batch's total capacity (it can never
 * be admitted); GDEM_E_ARG when its cell grid exceeds the cell capacity. */
int gdem_batch_admit(gdem_batch* batch, const gdem_particles_desc* particles, const gdem_mesh_desc* mesh,
                     const gdem_params* params, int nsteps, int* scene_out);
/* Enqueues nsteps batch steps (asynchronous on the context stream). */
int gdem_batch_step(gdem_batch* batch, int nsteps);
/* Scenes that have run all their steps and are not evicted yet. */
int gdem_batch_poll(gdem_batch* batch, int* scenes, int max_scenes, int* n_out);
int gdem_batch_download(gdem_batch* batch, int scene, gdem_particles_out* out);
int gdem_batch_scene_status(gdem_batch* batch, int scene, gdem_batch_scene_stats* stats);
/* Snapshot and restore of a batch scene (the SAVE_ON members copyFromDevice
 * and copyToDevice carry, ParticleSystem.cu:82-106), for evicting a running
 * scene and resuming it later: gdem_batch_download gives its state and
 * gdem_batch_history its contact histories at the current batch step;
 * gdem_batch_admit from that state followed by gdem_batch_set_history,
 * before the next gdem_batch_step (GDEM_E_STATE otherwise), resumes it.  A
 * resumed scene continues bit for bit as if it had never left. */
int gdem_batch_history(gdem_batch* batch, int scene, const gdem_history_out* out);
int gdem_batch_set_history(gdem_batch* batch, int scene, const gdem_history_view* in);
/* Frees a scene's slot and ranges (finished or not). */
int gdem_batch_evict(gdem_batch* batch, int scene);
/* Batch step count, live scenes, free particle slots, device bytes held. */
int gdem_batch_info(gdem_batch* batch, int64_t* step, int* live_scenes, int64_t* free_particles,
                    int64_t* bytes);
/* Device time (ms) of the last gdem_batch_step call; waits for it. */
int gdem_batch_timing(gdem_batch* batch, double* ms);

/* ---- settings and output files (host only, no GPU) ------------------------ */
/* Reads demSettings.json.  Keys match case-insensitively (as
 * cJSON_GetObjectItem).  A missing "YoungModulus" falls back to the legacy
 * "k"; a missing "bruteOn"/"gpuOn" means 0. */
int gdem_settings_load(const char* path, gdem_settings* out);
/* i-th walls.filepaths entry, relative paths resolved against the folder of
 * the settings file. */
int gdem_settings_wall(const char* path, int i, char* buf, int len);
/* Initial positions and velocities [n*3]: the file inlet's CSV, or the
 * shuffle inlet drawn exactly as initializeParticles (ParticleSystem.cu:
 * 205-285: std::mt19937 seeded with 1, overlap rejection, 1000 trials). */
int gdem_settings_particles(const char* path, double* x, double* v);
/* solver_output_init (solver_output.cu:14-41) */
int gdem_output_init(const char* dir);
/* write_frame_bin (solver_output.cu:259-316) from host state:
 * dir/result_%012d.bin */
int gdem_write_frame_bin(const char* dir, int step, int n, const double* x, const double* r,
                         const int* is_active, double length_factor);
/* write_frame_bin (solver_output.cu:259-316) from the arrays gdem_scene_frame
 * compacted on the device: the same file */
int gdem_write_frame_compact(const char* dir, int step, int n_active, const float* xyz, const float* r);
/* write_header_text (solver_output.cu:62-85): dir/pid<pid>.txt */
int gdem_write_pid_header(const char* dir, int pid);
/* write_single_text (solver_output.cu:112-172), unit factors */
int gdem_write_pid_row(const char* dir, int step, double dt, int pid, const double x[3], const double v[3],
                       const double a[3], const double angv[3], const double anga[3], int is_active);
/* write_initialPos_csv (solver_output.cu:87-110): dir/initialPos.csv */
int gdem_write_initial_pos(const char* dir, int n, const double* x, const double* r);

/* ---- primitives exposed for parity tests ---------------------------------- */
/* Stable LSD radix sort of (key, iota) on the device: the order
 * cub::DeviceRadixSort::SortPairs gives (BoundingBox.cu:1235-1242). */
int gdem_radix_sort(gdem_ctx* ctx, int n, const uint32_t* keys, uint32_t* keys_out, int* order_out);
/* d_dist_triangle (device_dem.cu:13-205) on the device for np points against
 * one triangle of a mesh: out[np*4] = (nx, ny, nz, dist), hit[np]. */
int gdem_tri_distance(gdem_ctx* ctx, const gdem_mesh_desc* mesh, int tri, int np,
                      const double* points, const double* rsq, double* out, int* hit);
/* Plain device-to-device streaming copy of `bytes` bytes, `reps` times; the
 * average duration per copy (ms).  Calibrates HBM byte counters and gives the
 * achievable copy bandwidth. */
int gdem_copy_probe(gdem_ctx* ctx, int64_t bytes, int reps, double* ms_out);
/* Device FP64 sqrt and division on arrays (checks correct rounding). */
int gdem_fp64_probe(gdem_ctx* ctx, int n, const double* a, const double* b, double* sqrt_out,
                    double* div_out);

#ifdef __cplusplus
}
#endif

#endif /* GDEM_H_ */
