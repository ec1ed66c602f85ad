"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a C++/OpenMP restatement of bi-rugaoishii/gpudem's per-step
contact path (cpu_dem.cu, BoundingBox.cu, BVH.cu, TriangleMesh.cu,
ParticleSystem.cu).  It is loaded only by tests/, by
``__graft_entry__.smoke()`` and by bench.py's ``cpu_baseline`` leg, always as
the checker and never as the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_dem.so")
MAX_NEI = 100  # hardCodedParameters.h:3

_lib = None


def build():
    """Compile the oracle with its Makefile (g++ -O3 -fopenmp -ffp-contract=off)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.orc_mesh_create.restype = vp
        L.orc_mesh_create.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int]
        L.orc_mesh_destroy.argtypes = [vp]
        L.orc_mesh_get.restype = ctypes.c_long
        L.orc_mesh_get.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [ctypes.c_int, dp, dp, dp, dp, ip, vp, dp]
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_sim_rebuild.argtypes = [vp]
        L.orc_sim_set_record.argtypes = [vp, ctypes.c_int]
        L.orc_sim_get.restype = ctypes.c_long
        L.orc_sim_get.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_sim_set.restype = ctypes.c_long
        L.orc_sim_set.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_sim_brute_lists.argtypes = [vp, ip, ip, ip, ip]
        L.orc_sim_morton.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_sim_morton_built.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_morton3.restype = ctypes.c_uint32
        L.orc_morton3.argtypes = [ctypes.c_uint32] * 3
        L.orc_radix_sort.argtypes = [ctypes.POINTER(ctypes.c_uint32), ip, ctypes.c_int]
        L.orc_tri_distance.restype = ctypes.c_int
        L.orc_tri_distance.argtypes = [vp, ctypes.c_int, dp, ctypes.c_double, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


_MESH_DOUBLE = ("mx my mz nx ny nz d e01x e01y e01z e02x e02y e02z e12x e12y e12z d00 d00inv "
                "d01 d11 d11inv d22 d22inv denom minx maxx miny maxy minz maxz cx cy cz "
                "bminx bminy bminz bmaxx bmaxy bmaxz gmin gmax").split()
_MESH_INT = "i0 i1 i2 te0 te1 te2 sortedIndex bleft bright btri".split()


class Mesh:
    """TriangleMesh precompute + BVH (TriangleMesh.cu:363-752, BVH.cu:712-787)."""

    def __init__(self, tri_verts, tris_per_file=None):
        v = np.ascontiguousarray(tri_verts, dtype=np.float64).reshape(-1, 9)
        self.ntri = v.shape[0]
        tpf = np.asarray([self.ntri] if tris_per_file is None else tris_per_file, dtype=np.int32)
        self._v = v
        self.h = lib().orc_mesh_create(_dp(v), self.ntri, _ip(tpf), len(tpf))

    def get(self, name):
        L = lib()
        key = name.encode()
        n = L.orc_mesh_get(self.h, key, None)
        if n < 0:
            raise KeyError(name)
        if name in ("nEdge", "nodeCount", "nTri"):
            return n
        dt = np.float64 if name in _MESH_DOUBLE else (np.uint32 if name == "mortonKey" else np.int32)
        out = np.empty(n, dtype=dt)
        L.orc_mesh_get(self.h, key, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def tri_distance(self, tri, p, rsq):
        out = np.zeros(4)
        p = np.ascontiguousarray(p, dtype=np.float64)
        hit = lib().orc_tri_distance(self.h, int(tri), _dp(p), float(rsq), _dp(out))
        return out[:3].copy(), float(out[3]), int(hit)

    def __del__(self):
        try:
            lib().orc_mesh_destroy(self.h)
        except Exception:
            pass


_SIM_DOUBLE = ("x v a f angv anga mom r rsq invr k m sqrtm invm moi invmoi etac refx refy refz "
               "dhx dhy dhz dwx dwy dwz recFn recFt recDel").split()
_SIM_SCALAR = ("steps", "rebuilds", "overflow", "ncell")


class Sim:
    """ParticleSys<HostMemory> + the reference step loop (cpu_dem.cu:1593-1668).

    ``params`` is a dict with dt, mu, g (3), bbox (minx,maxx,miny,maxy,minz,maxz),
    cell (grid spacing), skin, thresh, density, E, cor.
    """

    def __init__(self, x, r, mesh, params, v=None, angv=None, active=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.N = x.shape[0]
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.N)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        w = np.zeros_like(x) if angv is None else np.ascontiguousarray(angv, dtype=np.float64).reshape(-1, 3)
        act = np.ones(self.N, np.int32) if active is None else np.ascontiguousarray(active, dtype=np.int32)
        p = params
        P = np.array([p["dt"], p["mu"], *p["g"], *p["bbox"], p["cell"], p["skin"], p["thresh"],
                      p["density"], p["E"], p["cor"]], dtype=np.float64)
        self.mesh = mesh
        self.h = lib().orc_sim_create(self.N, _dp(x), _dp(v), _dp(w), _dp(r), _ip(act),
                                      mesh.h if mesh is not None else None, _dp(P))

    def step(self, n=1, gpu_order=True):
        lib().orc_sim_step(self.h, int(n), 1 if gpu_order else 0)

    def rebuild(self):
        lib().orc_sim_rebuild(self.h)

    def record(self, on=True):
        lib().orc_sim_set_record(self.h, 1 if on else 0)

    def get(self, name):
        L = lib()
        key = name.encode()
        n = L.orc_sim_get(self.h, key, None)
        if name in _SIM_SCALAR:
            return n
        if n < 0:
            raise KeyError(name)
        if name == "size":
            out = np.empty(3, np.int32)
        else:
            out = np.empty(n, dtype=np.float64 if name in _SIM_DOUBLE else np.int32)
        L.orc_sim_get(self.h, key, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def set(self, name, arr):
        dt = np.int32 if name == "isActive" else np.float64
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        if lib().orc_sim_set(self.h, name.encode(), a.ctypes.data_as(ctypes.c_void_p)) < 0:
            raise KeyError(name)

    def brute_lists(self):
        N = self.N
        pl = np.zeros(N * MAX_NEI, np.int32)
        pn = np.zeros(N, np.int32)
        wl = np.zeros(N * MAX_NEI, np.int32)
        wn = np.zeros(N, np.int32)
        lib().orc_sim_brute_lists(self.h, _ip(pl), _ip(pn), _ip(wl), _ip(wn))
        return pl.reshape(N, MAX_NEI), pn, wl.reshape(N, MAX_NEI), wn

    def lists(self):
        """Current pair and wall candidate lists as lists of sorted arrays."""
        pl = self.get("neiList").reshape(self.N, MAX_NEI)
        pn = self.get("numNei")
        wl = self.get("neiListWall").reshape(self.N, MAX_NEI)
        wn = self.get("numNeiWall")
        return ([pl[i, :pn[i]].copy() for i in range(self.N)],
                [wl[i, :wn[i]].copy() for i in range(self.N)])

    def morton(self):
        out = np.empty(self.N, np.uint32)
        lib().orc_sim_morton(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out

    def morton_built(self):
        """Morton keys of the cells stored at the last rebuild (0xffffffff inactive)."""
        out = np.empty(self.N, np.uint32)
        lib().orc_sim_morton_built(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out

    def history(self, wall=False):
        """Per particle list of (key, dx, dy, dz) in slot order."""
        sfx = "Wall" if wall else ""
        keys = self.get("indHis" + sfx).reshape(self.N, MAX_NEI)
        num = self.get("numCont" + sfx)
        names = ("dwx", "dwy", "dwz") if wall else ("dhx", "dhy", "dhz")
        d = np.stack([self.get(n).reshape(self.N, MAX_NEI) for n in names], axis=-1)
        return [(keys[i, :num[i]].copy(), d[i, :num[i]].copy()) for i in range(self.N)]

    def contacts(self):
        """Per-contact records of the last step (needs record(True) before it)."""
        n = self.get("recN")
        key = self.get("recKey").reshape(self.N, -1)
        kind = self.get("recKind").reshape(self.N, -1)
        fn = self.get("recFn").reshape(self.N, -1, 3)
        ft = self.get("recFt").reshape(self.N, -1, 3)
        de = self.get("recDel").reshape(self.N, -1)
        return [dict(key=key[i, :n[i]].copy(), kind=kind[i, :n[i]].copy(), fn=fn[i, :n[i]].copy(),
                     ft=ft[i, :n[i]].copy(), delta=de[i, :n[i]].copy()) for i in range(self.N)]

    def __del__(self):
        try:
            lib().orc_sim_destroy(self.h)
        except Exception:
            pass


def morton3(ix, iy, iz):
    return int(lib().orc_morton3(ix, iy, iz))


def radix_sort(keys, idx=None):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    i = np.arange(len(k), dtype=np.int32) if idx is None else np.ascontiguousarray(idx, dtype=np.int32).copy()
    lib().orc_radix_sort(k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _ip(i), len(k))
    return k, i


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())
