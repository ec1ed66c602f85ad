"""ctypes binding of libgdem.so (the C ABI declared in include/gdem.h).

The classes mirror the reference's host-side objects: ``Scene`` stands for
ParticleSys<DeviceMemory> + BoundingBox + BVH + DeviceTriangleMesh after
``copyToDevice`` (ParticleSystem.cu:82-89) and ``Scene.step`` for repeated
``device_dem_verlet_verlet`` (device_dem.cu:1195-1236).  There is no CPU
fallback: without the built library every call raises.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgdem.so")

GDEM_FP64 = 0
GDEM_MIXED = 1
GDEM_E_BUSY = -7
ERRORS = {-1: "GDEM_E_ARG", -2: "GDEM_E_HIP", -3: "GDEM_E_OOM", -4: "GDEM_E_SLOT_OVERFLOW",
          -5: "GDEM_E_NAN", -6: "GDEM_E_STATE", -7: "GDEM_E_BUSY"}

# exported symbols (checked by the CPU test suite against include/gdem.h)
SYMBOLS = ["gdem_ctx_create", "gdem_ctx_destroy", "gdem_sync", "gdem_last_error", "gdem_device_arch", "gdem_material",
           "gdem_bbox_rule", "gdem_params_default", "gdem_scene_create", "gdem_scene_step",
           "gdem_scene_status", "gdem_scene_download", "gdem_scene_download_noalloc", "gdem_scene_frame", "gdem_scene_destroy",
           "gdem_scene_debug",
           "gdem_scene_step_record", "gdem_scene_debug_order", "gdem_scene_timing", "gdem_scene_set_timing",
           "gdem_radix_sort", "gdem_tri_distance", "gdem_fp64_probe", "gdem_mesh_precompute",
           "gdem_copy_probe", "gdem_batch_create", "gdem_batch_destroy", "gdem_batch_admit",
           "gdem_batch_step", "gdem_batch_poll", "gdem_batch_download", "gdem_batch_scene_status",
           "gdem_batch_evict", "gdem_batch_info", "gdem_batch_timing", "gdem_batch_history", "gdem_batch_set_history", "gdem_scene_set_history",
           "gdem_settings_load", "gdem_settings_wall", "gdem_settings_particles", "gdem_output_init",
           "gdem_write_frame_bin", "gdem_write_frame_compact", "gdem_write_pid_header", "gdem_write_pid_row",
           "gdem_write_initial_pos"]


class GdemError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [("ntri", ctypes.c_int), ("verts", ctypes.c_void_p), ("nfiles", ctypes.c_int),
                ("tris_per_file", ctypes.c_void_p)]


class ParticlesDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("x", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("angv", ctypes.c_void_p), ("r", ctypes.c_void_p), ("m", ctypes.c_void_p),
                ("k", ctypes.c_void_p), ("etaconst", ctypes.c_void_p),
                ("is_active", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("mu", ctypes.c_double), ("g", ctypes.c_double * 3),
                ("bbox", ctypes.c_double * 6), ("cell", ctypes.c_double),
                ("skin", ctypes.c_double), ("refresh", ctypes.c_double),
                ("precision", ctypes.c_int), ("nb_cap", ctypes.c_int), ("wall_cap", ctypes.c_int),
                ("hist_cap", ctypes.c_int), ("whist_cap", ctypes.c_int),
                ("graph_steps", ctypes.c_int), ("wall_broadphase", ctypes.c_int)]


class ParticlesOut(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("angv", ctypes.c_void_p),
                ("is_active", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("rebuilds", ctypes.c_int64), ("n_active", ctypes.c_int),
                ("errors", ctypes.c_uint32), ("mean_nb", ctypes.c_double),
                ("mean_wall", ctypes.c_double), ("mean_contacts", ctypes.c_double)]


class BatchParams(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("max_scenes", ctypes.c_int), ("hbm_budget", ctypes.c_int64),
                ("particle_capacity", ctypes.c_int64), ("cell_capacity", ctypes.c_int64),
                ("tri_capacity", ctypes.c_int64), ("wall_capacity", ctypes.c_int64),
                ("nb_cap", ctypes.c_int), ("wall_cap", ctypes.c_int), ("hist_cap", ctypes.c_int),
                ("whist_cap", ctypes.c_int), ("graph_steps", ctypes.c_int)]


class BatchSceneStats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("start_step", ctypes.c_int64), ("end_step", ctypes.c_int64),
                ("steps_done", ctypes.c_int64), ("rebuilds", ctypes.c_int), ("errors", ctypes.c_uint32),
                ("n_active", ctypes.c_int), ("mean_nb", ctypes.c_double), ("mean_wall", ctypes.c_double),
                ("mean_contacts", ctypes.c_double), ("bytes", ctypes.c_int64)]


class DebugView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("cellx", "cell_id", "morton", "nb", "nnb", "wall", "nwall", "hist_key",
                 "hist_delta", "nhist", "whist_key", "whist_delta", "nwhist")]


class OrderView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("key", "order", "cell", "cell_first", "cell_end", "n_sorted")]


class ContactView(ctypes.Structure):
    _fields_ = [("rec_cap", ctypes.c_int)] + [(n, ctypes.c_void_p) for n in
                                              ("count", "key", "kind", "fn", "ft", "delta", "accel",
                                               "ang_accel")]


class HistoryView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("nhist", "hist_key", "hist_delta", "nwhist", "whist_key", "whist_delta")]


class HistoryOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("nhist", "hist_key", "hist_delta", "nwhist", "whist_key", "whist_delta")]


class Settings(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("input_mode", ctypes.c_int), ("inlet_box", ctypes.c_double * 6),
                ("radius", ctypes.c_double), ("density", ctypes.c_double), ("cor", ctypes.c_double),
                ("mu", ctypes.c_double), ("young", ctypes.c_double), ("k", ctypes.c_double),
                ("etaconst", ctypes.c_double), ("m", ctypes.c_double), ("g", ctypes.c_double * 3),
                ("dt", ctypes.c_double), ("out_step", ctypes.c_int), ("steps", ctypes.c_int),
                ("seed", ctypes.c_int), ("gpu_on", ctypes.c_int), ("brute_on", ctypes.c_int),
                ("nwalls", ctypes.c_int)]


_lib = None


def lib():
    """Loads libgdem.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with gpudem_amd/_build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.gdem_ctx_create.argtypes = [ctypes.c_int, P(vp)]
        L.gdem_ctx_destroy.argtypes = [vp]
        L.gdem_ctx_destroy.restype = None
        L.gdem_sync.argtypes = [vp]
        L.gdem_last_error.argtypes = [vp]
        L.gdem_last_error.restype = ctypes.c_char_p
        L.gdem_device_arch.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int)]
        L.gdem_material.argtypes = [ctypes.c_int, vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, vp, vp, vp]
        L.gdem_bbox_rule.argtypes = [P(MeshDesc), vp, vp]
        L.gdem_params_default.argtypes = [ctypes.c_double, P(Params)]
        L.gdem_scene_create.argtypes = [vp, P(ParticlesDesc), P(MeshDesc), P(Params), P(vp)]
        L.gdem_scene_step.argtypes = [vp, ctypes.c_int]
        L.gdem_scene_status.argtypes = [vp, P(Stats)]
        L.gdem_scene_download.argtypes = [vp, P(ParticlesOut)]
        L.gdem_scene_download_noalloc.argtypes = [vp, P(ParticlesOut)]
        L.gdem_scene_frame.argtypes = [vp, ctypes.c_double, vp, vp, P(ctypes.c_int)]
        L.gdem_scene_destroy.argtypes = [vp]
        L.gdem_scene_destroy.restype = None
        L.gdem_scene_debug.argtypes = [vp, P(DebugView)]
        L.gdem_scene_step_record.argtypes = [vp, P(ContactView)]
        L.gdem_scene_debug_order.argtypes = [vp, P(OrderView)]
        L.gdem_scene_timing.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int)]
        L.gdem_scene_set_timing.argtypes = [vp, ctypes.c_int]
        L.gdem_radix_sort.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.gdem_tri_distance.argtypes = [vp, P(MeshDesc), ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.gdem_fp64_probe.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.gdem_mesh_precompute.argtypes = [P(MeshDesc), vp, vp, vp]
        L.gdem_copy_probe.argtypes = [vp, ctypes.c_int64, ctypes.c_int, P(ctypes.c_double)]
        L.gdem_batch_create.argtypes = [vp, P(BatchParams), P(vp)]
        L.gdem_batch_destroy.argtypes = [vp]
        L.gdem_batch_destroy.restype = None
        L.gdem_batch_admit.argtypes = [vp, P(ParticlesDesc), P(MeshDesc), P(Params), ctypes.c_int,
                                       P(ctypes.c_int)]
        L.gdem_batch_step.argtypes = [vp, ctypes.c_int]
        L.gdem_batch_poll.argtypes = [vp, vp, ctypes.c_int, P(ctypes.c_int)]
        L.gdem_batch_download.argtypes = [vp, ctypes.c_int, P(ParticlesOut)]
        L.gdem_batch_scene_status.argtypes = [vp, ctypes.c_int, P(BatchSceneStats)]
        L.gdem_batch_evict.argtypes = [vp, ctypes.c_int]
        L.gdem_batch_info.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int), P(ctypes.c_int64),
                                      P(ctypes.c_int64)]
        L.gdem_batch_timing.argtypes = [vp, P(ctypes.c_double)]
        L.gdem_batch_history.argtypes = [vp, ctypes.c_int, P(HistoryOut)]
        L.gdem_batch_set_history.argtypes = [vp, ctypes.c_int, P(HistoryView)]
        L.gdem_scene_set_history.argtypes = [vp, P(HistoryView)]
        L.gdem_settings_load.argtypes = [ctypes.c_char_p, P(Settings)]
        L.gdem_settings_wall.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.gdem_settings_particles.argtypes = [ctypes.c_char_p, vp, vp]
        L.gdem_output_init.argtypes = [ctypes.c_char_p]
        L.gdem_write_frame_bin.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                           ctypes.c_double]
        L.gdem_write_frame_compact.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, vp, vp]
        L.gdem_write_pid_header.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.gdem_write_pid_row.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, vp, vp,
                                         vp, vp, vp, ctypes.c_int]
        L.gdem_write_initial_pos.argtypes = [ctypes.c_char_p, ctypes.c_int, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def device_arch(device=0):
    """Architecture name of a HIP device without its feature suffixes (gfx950)."""
    buf = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    rc = lib().gdem_device_arch(int(device), buf, len(buf), ctypes.byref(cus))
    if rc:
        raise GdemError(rc, "gdem_device_arch")
    return buf.value.decode().split(":")[0]


def material(r, density, young, cor):
    """(m, k, etaconst) per particle, as main.cu:67 and ParticleSystem.cu:202,301."""
    r = _f64(r).ravel()
    m, k, e = np.empty_like(r), np.empty_like(r), np.empty_like(r)
    rc = lib().gdem_material(len(r), _ptr(r), density, young, cor, _ptr(m), _ptr(k), _ptr(e))
    if rc:
        raise GdemError(rc, "gdem_material")
    return m, k, e


def mesh_desc(verts, tris_per_file=None):
    v = _f64(verts).reshape(-1, 9)
    tpf = None if tris_per_file is None else np.ascontiguousarray(tris_per_file, dtype=np.int32)
    d = MeshDesc(v.shape[0], _ptr(v), 0 if tpf is None else len(tpf), _ptr(tpf))
    d._keep = (v, tpf)
    return d


def mesh_precompute(verts, tris_per_file=None):
    """Host-side triangle precompute of the library (no GPU needed)."""
    md = mesh_desc(verts, tris_per_file)
    geo = np.empty((md.ntri, 29))
    ids = np.empty((md.ntri, 6), np.int32)
    order = np.empty(md.ntri, np.int32)
    rc = lib().gdem_mesh_precompute(ctypes.byref(md), _ptr(geo), _ptr(ids), _ptr(order))
    if rc:
        raise GdemError(rc, "gdem_mesh_precompute")
    names = ["n", "d", "p0", "p1", "p2", "e01", "e02", "e12"]
    out, o = {}, 0
    for nm in names:
        w = 1 if nm == "d" else 3
        out[nm] = geo[:, o:o + w] if w > 1 else geo[:, o]
        o += w
    for i, nm in enumerate(["d00", "d01", "d11", "denom", "d00inv", "d11inv", "d22inv"]):
        out[nm] = geo[:, 22 + i]
    out["ids"] = ids
    out["order"] = order
    return out


def bbox_rule(verts, inlet):
    """calc_BoundingBoxLimits (BoundingBox.cu:448-478)."""
    md = mesh_desc(verts if verts is not None else np.zeros((0, 9)))
    inl = _f64(inlet)
    out = np.empty(6)
    rc = lib().gdem_bbox_rule(ctypes.byref(md), _ptr(inl), _ptr(out))
    if rc:
        raise GdemError(rc, "gdem_bbox_rule")
    return out


def scene_inputs(x, r, m, k, etaconst, verts, params, v=None, angv=None, active=None, tris_per_file=None):
    """C descriptors of one scene (particles, mesh, params) and the arrays they point into."""
    x = _f64(x, (-1, 3))
    n = x.shape[0]
    arrs = dict(x=x, r=_f64(r).ravel(), m=_f64(m).ravel(), k=_f64(k).ravel(), e=_f64(etaconst).ravel(),
                v=None if v is None else _f64(v, (-1, 3)),
                w=None if angv is None else _f64(angv, (-1, 3)),
                a=None if active is None else np.ascontiguousarray(active, dtype=np.int32))
    pd = ParticlesDesc(n, _ptr(arrs["x"]), _ptr(arrs["v"]), _ptr(arrs["w"]), _ptr(arrs["r"]),
                       _ptr(arrs["m"]), _ptr(arrs["k"]), _ptr(arrs["e"]), _ptr(arrs["a"]))
    md = mesh_desc(verts if verts is not None else np.zeros((0, 9)), tris_per_file)
    p = Params()
    p.dt = params["dt"]
    p.mu = params["mu"]
    p.g = (ctypes.c_double * 3)(*params["g"])
    p.bbox = (ctypes.c_double * 6)(*params["bbox"])
    p.cell = params["cell"]
    p.skin = params["skin"]
    p.refresh = params["refresh"]
    p.precision = params.get("precision", GDEM_FP64)
    p.nb_cap = params.get("nb_cap", 0)
    p.wall_cap = params.get("wall_cap", 0)
    p.hist_cap = params.get("hist_cap", 0)
    p.whist_cap = params.get("whist_cap", 0)
    p.graph_steps = params.get("graph_steps", 0)
    p.wall_broadphase = params.get("wall_broadphase", 0)
    return pd, md, p, arrs


class Context:
    """One HIP device + stream (gdem_ctx)."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        self._scenes = weakref.WeakSet()
        rc = lib().gdem_ctx_create(device, ctypes.byref(self.h))
        if rc:
            raise GdemError(rc, "gdem_ctx_create (no usable HIP device?)")

    def error(self):
        return lib().gdem_last_error(self.h).decode()

    def check(self, rc, what):
        if rc:
            raise GdemError(rc, f"{what}: {self.error()}")

    def sync(self):
        self.check(lib().gdem_sync(self.h), "gdem_sync")

    def radix_sort(self, keys):
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        ko = np.empty_like(k)
        io = np.empty(len(k), np.int32)
        self.check(lib().gdem_radix_sort(self.h, len(k), _ptr(k), _ptr(ko), _ptr(io)), "radix_sort")
        return ko, io

    def tri_distance(self, verts, tri, points, rsq):
        md = mesh_desc(verts)
        p = _f64(points).reshape(-1, 3)
        r = _f64(rsq).reshape(-1)
        out = np.empty((len(p), 4))
        hit = np.empty(len(p), np.int32)
        self.check(lib().gdem_tri_distance(self.h, ctypes.byref(md), int(tri), len(p), _ptr(p), _ptr(r),
                                           _ptr(out), _ptr(hit)), "tri_distance")
        return out, hit

    def copy_probe(self, nbytes, reps=20):
        """Average duration (ms) of a plain device copy of nbytes."""
        ms = ctypes.c_double()
        self.check(lib().gdem_copy_probe(self.h, int(nbytes), int(reps), ctypes.byref(ms)), "copy_probe")
        return ms.value

    def fp64_probe(self, a, b):
        a = _f64(a)
        b = _f64(b)
        s = np.empty_like(a)
        d = np.empty_like(a)
        self.check(lib().gdem_fp64_probe(self.h, len(a), _ptr(a), _ptr(b), _ptr(s), _ptr(d)), "fp64_probe")
        return s, d

    def close(self):
        """Destroys every scene still open on this context, then the context."""
        if self.h:
            for sc in list(self._scenes):
                sc.close()
            lib().gdem_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Device-resident particle system + mesh (gdem_scene).

    x: (n,3) positions; r: (n,) radii; m, k, etaconst: material per particle
    (see ``material``); verts: (ntri,9) triangle soup; params: dict with dt, mu,
    g, bbox, cell, skin, refresh and optional precision, caps, graph_steps.
    """

    def __init__(self, ctx, x, r, m, k, etaconst, verts, params, v=None, angv=None, active=None,
                 tris_per_file=None):
        self.ctx = ctx
        pd, md, p, keep = scene_inputs(x, r, m, k, etaconst, verts, params, v, angv, active, tris_per_file)
        self.n = pd.n
        self.caps = dict(nb=p.nb_cap or 48, wall=p.wall_cap or 24, hist=p.hist_cap or 16,
                         whist=p.whist_cap or 8)
        self.h = ctypes.c_void_p()
        ctx.check(lib().gdem_scene_create(ctx.h, ctypes.byref(pd), ctypes.byref(md), ctypes.byref(p),
                                          ctypes.byref(self.h)), "gdem_scene_create")
        ctx._scenes.add(self)

    def step(self, n=1):
        self.ctx.check(lib().gdem_scene_step(self.h, int(n)), "gdem_scene_step")

    def status(self, check=True):
        st = Stats()
        rc = lib().gdem_scene_status(self.h, ctypes.byref(st))
        if check:
            self.ctx.check(rc, "gdem_scene_status")
        return dict(steps=st.steps, rebuilds=st.rebuilds, n_active=st.n_active, errors=st.errors,
                    mean_nb=st.mean_nb, mean_wall=st.mean_wall, mean_contacts=st.mean_contacts, rc=rc)

    def download(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        w = np.empty((self.n, 3))
        a = np.empty(self.n, np.int32)
        out = ParticlesOut(_ptr(x), _ptr(v), _ptr(w), _ptr(a))
        self.ctx.check(lib().gdem_scene_download(self.h, ctypes.byref(out)), "gdem_scene_download")
        return dict(x=x, v=v, angv=w, active=a)

    def download_noalloc(self, out=None):
        """download() through the library's fixed staging array (no device
        allocation; scenes of at most 262,144 particles).  `out`: a dict of
        preallocated arrays as download() returns, filled in place."""
        if out is None:
            out = dict(x=np.empty((self.n, 3)), v=np.empty((self.n, 3)), angv=np.empty((self.n, 3)),
                       active=np.empty(self.n, np.int32))
        po = ParticlesOut(_ptr(out["x"]), _ptr(out["v"]), _ptr(out["angv"]), _ptr(out["active"]))
        self.ctx.check(lib().gdem_scene_download_noalloc(self.h, ctypes.byref(po)), "gdem_scene_download_noalloc")
        return out

    def frame(self, length_factor=1.0):
        """The active particles' positions (float32, n_active x 3) and radii
        (float32) in original id order, scaled by length_factor, compacted on
        the device (the data of one result_%012d.bin frame)."""
        xyz = np.empty((self.n, 3), np.float32)
        r = np.empty(self.n, np.float32)
        na = ctypes.c_int(0)
        self.ctx.check(lib().gdem_scene_frame(self.h, float(length_factor), _ptr(xyz), _ptr(r), ctypes.byref(na)),
                       "gdem_scene_frame")
        return xyz[:na.value], r[:na.value]

    def debug(self):
        n, c = self.n, self.caps
        d = dict(cellx=np.empty((n, 3), np.int32), cell_id=np.empty(n, np.int32),
                 morton=np.empty(n, np.uint32), nb=np.empty((n, c["nb"]), np.int32),
                 nnb=np.empty(n, np.int32), wall=np.empty((n, c["wall"]), np.int32),
                 nwall=np.empty(n, np.int32), hist_key=np.empty((n, c["hist"]), np.int32),
                 hist_delta=np.empty((n, c["hist"], 3)), nhist=np.empty(n, np.int32),
                 whist_key=np.empty((n, c["whist"]), np.int32),
                 whist_delta=np.empty((n, c["whist"], 3)), nwhist=np.empty(n, np.int32))
        view = DebugView(*[_ptr(d[f]) for f, _ in DebugView._fields_])
        self.ctx.check(lib().gdem_scene_debug(self.h, ctypes.byref(view)), "gdem_scene_debug")
        return d

    def debug_order(self):
        """What the last list rebuild stored per slot: sort key, original id,
        cell id and the cell's (first, end) slot range; plus the number of
        active (sorted-first) slots."""
        n = self.n
        d = dict(key=np.empty(n, np.uint32), order=np.empty(n, np.int32), cell=np.empty(n, np.int32),
                 cell_first=np.empty(n, np.int32), cell_end=np.empty(n, np.int32),
                 n_sorted=np.empty(1, np.int32))
        view = OrderView(*[_ptr(d[f]) for f, _ in OrderView._fields_])
        self.ctx.check(lib().gdem_scene_debug_order(self.h, ctypes.byref(view)), "gdem_scene_debug_order")
        d["n_sorted"] = int(d["n_sorted"][0])
        return d

    def step_record(self, cap=64):
        """One step that records every contact (partner id or triangle id,
        kind, fn, ft, overlap) and every particle's accelerations."""
        n = self.n
        d = dict(count=np.zeros(n, np.int32), key=np.empty((n, cap), np.int32),
                 kind=np.empty((n, cap), np.int32), fn=np.empty((n, cap, 3)), ft=np.empty((n, cap, 3)),
                 delta=np.empty((n, cap)), accel=np.zeros((n, 3)), ang_accel=np.zeros((n, 3)))
        view = ContactView(cap, *[_ptr(d[f]) for f in ("count", "key", "kind", "fn", "ft", "delta", "accel",
                                                        "ang_accel")])
        self.ctx.check(lib().gdem_scene_step_record(self.h, ctypes.byref(view)), "gdem_scene_step_record")
        return d

    def set_history(self, hist):
        """Replaces the contact histories (a dict as debug() returns: nhist,
        hist_key, hist_delta, nwhist, whist_key, whist_delta)."""
        n, c = self.n, self.caps
        arrs = dict(nhist=np.ascontiguousarray(hist["nhist"], np.int32),
                    hist_key=np.ascontiguousarray(hist["hist_key"], np.int32).reshape(n, c["hist"]),
                    hist_delta=np.ascontiguousarray(hist["hist_delta"], np.float64).reshape(n, c["hist"], 3),
                    nwhist=np.ascontiguousarray(hist["nwhist"], np.int32),
                    whist_key=np.ascontiguousarray(hist["whist_key"], np.int32).reshape(n, c["whist"]),
                    whist_delta=np.ascontiguousarray(hist["whist_delta"], np.float64).reshape(n, c["whist"], 3))
        view = HistoryView(*[_ptr(arrs[f]) for f, _ in HistoryView._fields_])
        self.ctx.check(lib().gdem_scene_set_history(self.h, ctypes.byref(view)), "gdem_scene_set_history")

    def snapshot(self):
        """SAVE_ON state: positions, velocities, spins, activity and the
        contact histories, in original particle order."""
        d = self.download()
        dbg = self.debug()
        for f in ("nhist", "hist_key", "hist_delta", "nwhist", "whist_key", "whist_delta"):
            d[f] = dbg[f]
        return d

    def set_timing(self, every=1):
        """Time the next step() calls; every: sample the force kernel of every n-th step (0 = off)."""
        self.ctx.check(lib().gdem_scene_set_timing(self.h, int(every)), "set_timing")

    def timing(self):
        t = ctypes.c_double()
        f = ctypes.c_double()
        k = ctypes.c_int()
        self.ctx.check(lib().gdem_scene_timing(self.h, ctypes.byref(t), ctypes.byref(f), ctypes.byref(k)),
                       "timing")
        return t.value, f.value, k.value

    def close(self):
        if self.h:
            if self.ctx.h:
                lib().gdem_scene_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent scenes stepped together (gdem_batch).

    Capacities are fixed at creation (``particle_capacity`` or ``hbm_budget``
    bytes); ``admit`` returns a scene id or None when no slot or capacity is
    free (GDEM_E_BUSY); ``step`` advances every running scene; finished scenes
    show up in ``poll`` until evicted.
    """

    def __init__(self, ctx, precision=GDEM_FP64, max_scenes=256, hbm_budget=0, particle_capacity=0,
                 cell_capacity=0, tri_capacity=0, wall_capacity=0, nb_cap=0, wall_cap=0, hist_cap=0,
                 whist_cap=0, graph_steps=0):
        self.ctx = ctx
        self.precision = precision
        bp = BatchParams(precision, max_scenes, int(hbm_budget), int(particle_capacity), int(cell_capacity),
                         int(tri_capacity), int(wall_capacity), nb_cap, wall_cap, hist_cap, whist_cap,
                         graph_steps)
        self.h = ctypes.c_void_p()
        self.n_of = {}
        self.caps = dict(hist=hist_cap or 16, whist=whist_cap or 8)
        ctx.check(lib().gdem_batch_create(ctx.h, ctypes.byref(bp), ctypes.byref(self.h)), "gdem_batch_create")
        ctx._scenes.add(self)

    def admit(self, x, r, m, k, etaconst, verts, params, nsteps, v=None, angv=None, active=None,
              tris_per_file=None):
        pd, md, p, keep = scene_inputs(x, r, m, k, etaconst, verts, params, v, angv, active, tris_per_file)
        sid = ctypes.c_int(-1)
        rc = lib().gdem_batch_admit(self.h, ctypes.byref(pd), ctypes.byref(md), ctypes.byref(p), int(nsteps),
                                    ctypes.byref(sid))
        if rc == GDEM_E_BUSY:
            return None
        self.ctx.check(rc, "gdem_batch_admit")
        self.n_of[sid.value] = pd.n
        return sid.value

    def step(self, n=1):
        self.ctx.check(lib().gdem_batch_step(self.h, int(n)), "gdem_batch_step")

    def poll(self):
        ids = (ctypes.c_int * 4096)()
        n = ctypes.c_int()
        self.ctx.check(lib().gdem_batch_poll(self.h, ids, 4096, ctypes.byref(n)), "gdem_batch_poll")
        return [ids[i] for i in range(min(n.value, 4096))]

    def download(self, sid):
        n = self.n_of[sid]
        x = np.empty((n, 3))
        v = np.empty((n, 3))
        w = np.empty((n, 3))
        a = np.empty(n, np.int32)
        out = ParticlesOut(_ptr(x), _ptr(v), _ptr(w), _ptr(a))
        self.ctx.check(lib().gdem_batch_download(self.h, int(sid), ctypes.byref(out)), "gdem_batch_download")
        return dict(x=x, v=v, angv=w, active=a)

    def status(self, sid, check=True):
        st = BatchSceneStats()
        rc = lib().gdem_batch_scene_status(self.h, int(sid), ctypes.byref(st))
        if check:
            self.ctx.check(rc, "gdem_batch_scene_status")
        return {f: getattr(st, f) for f, _ in BatchSceneStats._fields_}

    def history(self, sid):
        """Contact histories of a scene in its particles' own ids (pair keys
        partner ids, wall keys triangle ids of the scene's mesh)."""
        n, c = self.n_of[sid], self.caps
        d = dict(nhist=np.zeros(n, np.int32), hist_key=np.full((n, c["hist"]), -1, np.int32),
                 hist_delta=np.zeros((n, c["hist"], 3)), nwhist=np.zeros(n, np.int32),
                 whist_key=np.full((n, c["whist"]), -1, np.int32), whist_delta=np.zeros((n, c["whist"], 3)))
        out = HistoryOut(*[_ptr(d[f]) for f, _ in HistoryOut._fields_])
        self.ctx.check(lib().gdem_batch_history(self.h, int(sid), ctypes.byref(out)), "gdem_batch_history")
        return d

    def snapshot(self, sid):
        """State and contact histories of a scene at the current batch step."""
        d = self.download(sid)
        d.update(self.history(sid))
        return d

    def set_history(self, sid, hist):
        """Histories (as history() returns them) into a scene admitted and
        not stepped yet."""
        n, c = self.n_of[sid], self.caps
        arrs = dict(nhist=np.ascontiguousarray(hist["nhist"], np.int32),
                    hist_key=np.ascontiguousarray(hist["hist_key"], np.int32).reshape(n, c["hist"]),
                    hist_delta=np.ascontiguousarray(hist["hist_delta"], np.float64).reshape(n, c["hist"], 3),
                    nwhist=np.ascontiguousarray(hist["nwhist"], np.int32),
                    whist_key=np.ascontiguousarray(hist["whist_key"], np.int32).reshape(n, c["whist"]),
                    whist_delta=np.ascontiguousarray(hist["whist_delta"], np.float64).reshape(n, c["whist"], 3))
        view = HistoryView(*[_ptr(arrs[f]) for f, _ in HistoryView._fields_])
        self.ctx.check(lib().gdem_batch_set_history(self.h, int(sid), ctypes.byref(view)), "gdem_batch_set_history")

    def evict(self, sid):
        self.ctx.check(lib().gdem_batch_evict(self.h, int(sid)), "gdem_batch_evict")
        self.n_of.pop(sid, None)

    def info(self):
        step = ctypes.c_int64()
        live = ctypes.c_int()
        free = ctypes.c_int64()
        nbytes = ctypes.c_int64()
        self.ctx.check(lib().gdem_batch_info(self.h, ctypes.byref(step), ctypes.byref(live), ctypes.byref(free),
                                             ctypes.byref(nbytes)), "gdem_batch_info")
        return dict(step=step.value, live=live.value, free_particles=free.value, bytes=nbytes.value)

    def timing(self):
        """Device ms of the last step() call (waits for it)."""
        ms = ctypes.c_double()
        self.ctx.check(lib().gdem_batch_timing(self.h, ctypes.byref(ms)), "gdem_batch_timing")
        return ms.value

    def close(self):
        if self.h:
            if self.ctx.h:
                lib().gdem_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- demSettings.json and output files (host only) ---------------------------


def _b(path):
    return os.fsencode(path)


def load_settings(path):
    """demSettings.json as the C loader reads it (dict), plus the wall files."""
    st = Settings()
    rc = lib().gdem_settings_load(_b(path), ctypes.byref(st))
    if rc:
        raise GdemError(rc, f"gdem_settings_load({path})")
    out = {f: (list(getattr(st, f)) if isinstance(getattr(st, f), ctypes.Array) else getattr(st, f))
           for f, _ in Settings._fields_}
    walls = []
    buf = ctypes.create_string_buffer(4096)
    for i in range(st.nwalls):
        rc = lib().gdem_settings_wall(_b(path), i, buf, len(buf))
        if rc:
            raise GdemError(rc, f"gdem_settings_wall({path}, {i})")
        walls.append(os.fsdecode(buf.value))
    out["walls"] = walls
    return out


def settings_particles(path, n):
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    rc = lib().gdem_settings_particles(_b(path), _ptr(x), _ptr(v))
    if rc:
        raise GdemError(rc, f"gdem_settings_particles({path})")
    return x, v


def output_init(d):
    rc = lib().gdem_output_init(_b(d))
    if rc:
        raise GdemError(rc, f"gdem_output_init({d})")


def write_frame_bin(d, step, x, r, active, length_factor=1.0):
    x = _f64(x, (-1, 3))
    r = _f64(r).ravel()
    a = np.ascontiguousarray(active, np.int32)
    rc = lib().gdem_write_frame_bin(_b(d), int(step), len(r), _ptr(x), _ptr(r), _ptr(a), float(length_factor))
    if rc:
        raise GdemError(rc, "gdem_write_frame_bin")


def write_frame_compact(d, step, xyz, r):
    """result_%012d.bin from Scene.frame()'s arrays."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(r, np.float32).ravel()
    if len(r) != len(xyz):
        raise ValueError("write_frame_compact: xyz and r differ in length")
    rc = lib().gdem_write_frame_compact(_b(d), int(step), len(r), _ptr(xyz), _ptr(r))
    if rc:
        raise GdemError(rc, "gdem_write_frame_compact")


def write_pid_header(d, pid):
    rc = lib().gdem_write_pid_header(_b(d), int(pid))
    if rc:
        raise GdemError(rc, "gdem_write_pid_header")


def write_pid_row(d, step, dt, pid, x, v, a, angv, anga, active):
    arrs = [_f64(q).ravel()[:3].copy() for q in (x, v, a, angv, anga)]
    rc = lib().gdem_write_pid_row(_b(d), int(step), float(dt), int(pid), *[_ptr(q) for q in arrs], int(active))
    if rc:
        raise GdemError(rc, "gdem_write_pid_row")


def write_initial_pos(d, x, r):
    x = _f64(x, (-1, 3))
    r = _f64(r).ravel()
    rc = lib().gdem_write_initial_pos(_b(d), len(r), _ptr(x), _ptr(r))
    if rc:
        raise GdemError(rc, "gdem_write_initial_pos")
