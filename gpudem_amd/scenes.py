"""Seeded scene producers: meshes, particle placements and step parameters.

Every scene is fully determined by its seed, so the same input is fed to the
CPU oracle and to the GPU engine.  Meshes are triangle soups in facet order
(one row of 9 coordinates per triangle), the form an ASCII STL file gives.
Material follows the reference's set-up (main.cu:67, ParticleSystem.cu:202,
301); the cell grid, skin and refresh rules follow BoundingBox.cu:364-478
unless a workload names its own.
"""
import math
import os

import numpy as np


def box_mesh(lo, hi):
    """Closed axis-aligned box as 12 right triangles (two per face)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    tris = []
    for a, b, cc, d in quads:
        tris.append(np.concatenate([c[a], c[b], c[cc]]))
        tris.append(np.concatenate([c[a], c[cc], c[d]]))
    return np.array(tris)


def bbox_rule(verts, inlet):
    """calc_BoundingBoxLimits (BoundingBox.cu:448-478): min/max over the mesh and
    the inlet box, each multiplied by 1.1 (so a positive minimum moves up)."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    out = np.empty(6)
    for a in range(3):
        mn = min(v[:, a].min(), inlet[2 * a]) if len(v) else inlet[2 * a]
        mx = max(v[:, a].max(), inlet[2 * a + 1]) if len(v) else inlet[2 * a + 1]
        out[2 * a] = mn * 1.1
        out[2 * a + 1] = mx * 1.1
    return out


class SceneSpec:
    """Inputs of one plain scene."""

    def __init__(self, name, x, r, verts, params, density, young, cor, v=None, steps=0):
        self.name = name
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.r = np.ascontiguousarray(r, dtype=np.float64)
        self.verts = np.ascontiguousarray(verts, dtype=np.float64)
        self.params = params
        self.density, self.young, self.cor = density, young, cor
        self.v = v
        self.steps = steps

    @property
    def n(self):
        return len(self.r)

    def oracle_params(self):
        p = self.params
        return dict(dt=p["dt"], mu=p["mu"], g=p["g"], bbox=p["bbox"], cell=p["cell"], skin=p["skin"],
                    thresh=p["refresh"], density=self.density, E=self.young, cor=self.cor)


def lattice_drop(seed=1, n_side=(5, 5, 4), r=0.005, box=0.1, density=2500., young=1e7, cor=0.5,
                 mu=0.5, dt=1e-5, steps=20000, jitter=0.002, verlet=False):
    """100 monodisperse spheres on a jittered 5x5x4 lattice in a closed 0.1 m
    cube (12 triangles), zero velocity, gravity -z.  By default a cell list of
    cell size 2r rebuilt every step (skin 0, refresh 0); verlet=True uses the
    reference's own rule (cell 4r, skin r, refresh r/2)."""
    rng = np.random.default_rng(seed)
    h = box / 2
    nx, ny, nz = n_side
    xs = (np.arange(nx) + 0.5) / nx * box - h
    ys = (np.arange(ny) + 0.5) / ny * box - h
    zs = (np.arange(nz) + 0.5) / nz * box - h
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    x = x + rng.uniform(-jitter, jitter, size=x.shape)
    n = len(x)
    radius = np.full(n, r)
    verts = box_mesh((-h, -h, -h), (h, h, h))
    bbox = bbox_rule(verts, [-h, h, -h, h, -h, h])
    if verlet:
        cell, skin, refresh = 4 * r, r, 0.5 * r
    else:
        cell, skin, refresh = 2 * r, 0.0, 0.0
    params = dict(dt=dt, mu=mu, g=(0., 0., -9.81), bbox=bbox, cell=cell, skin=skin, refresh=refresh)
    return SceneSpec("lattice_drop", x, radius, verts, params, density, young, cor, steps=steps)


def polydisperse_bed(seed=2, n=30000, r0=0.002, spread=0.1, size=(0.1, 0.1, 0.3), density=2500.,
                     young=1e7, cor=0.5, mu=0.5, dt=5e-6, steps=50000, spacing_factor=2.0225):
    """30,000 spheres, r ~ U[0.9, 1.1] r0, settling in a 0.1 x 0.1 x 0.3 m box.
    They start on a jittered simple-cubic lattice (spacing 2.0225 r_max,
    filled from the bottom) because 30,000 such spheres do not fit the upper
    half of the box at all.  Verlet skin 0.1 r_max, refresh at skin/2."""
    rng = np.random.default_rng(seed)
    rmax = r0 * (1 + spread)
    radius = rng.uniform(r0 * (1 - spread), rmax, size=n)
    hx, hy, hz = size[0] / 2, size[1] / 2, size[2] / 2
    sp = spacing_factor * rmax
    clear = 1.02 * rmax
    nx = int((size[0] - 2 * clear) / sp) + 1
    ny = int((size[1] - 2 * clear) / sp) + 1
    nz = int((size[2] - 2 * clear) / sp) + 1
    if nx * ny * nz < n:
        raise ValueError("box too small for the requested particle count")
    xs = -hx + clear + np.arange(nx) * sp
    ys = -hy + clear + np.arange(ny) * sp
    zs = -hz + clear + np.arange(nz) * sp
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)[:n]
    j = 0.5 * (sp - 2 * rmax) * 0.9
    x = x + rng.uniform(-j, j, size=x.shape)
    verts = box_mesh((-hx, -hy, -hz), (hx, hy, hz))
    bbox = bbox_rule(verts, [-hx, hx, -hy, hy, -hz, hz])
    skin = 0.1 * rmax
    params = dict(dt=dt, mu=mu, g=(0., 0., -9.81), bbox=bbox, cell=2 * rmax + skin, skin=skin,
                  refresh=0.5 * skin)
    return SceneSpec("polydisperse_bed", x, radius, verts, params, density, young, cor, steps=steps)


def helical_chute_mesh(seed=3, turns=4, radius=0.2, pitch=0.1, width=0.05, ntri=20000, rail=0.02):
    """Helical chute: a descending ramp of `turns` turns around the z axis with
    inner and outer rails and a closed top end.  The ramp's cross-slope tilt is
    drawn from the seed.  Returns (verts, tilt, z_of(theta, rho))."""
    rng = np.random.default_rng(seed)
    tilt = rng.uniform(math.radians(2.0), math.radians(6.0))
    seg = (ntri - 2) // 6
    rin, rout = radius - width / 2, radius + width / 2
    total = 2 * math.pi * turns
    ztop = pitch * turns / 2

    def z_floor(th, rho):
        return ztop - pitch * th / (2 * math.pi) + (rho - radius) * math.tan(tilt)

    th = np.linspace(0.0, total, seg + 1)
    tris = []

    def P(t, rho, dz=0.0):
        return np.array([rho * math.cos(t), rho * math.sin(t), z_floor(t, rho) + dz])

    for i in range(seg):
        t0, t1 = th[i], th[i + 1]
        a, b, c, d = P(t0, rin), P(t0, rout), P(t1, rout), P(t1, rin)
        tris += [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
        a, b, c, d = P(t0, rin), P(t1, rin), P(t1, rin, rail), P(t0, rin, rail)
        tris += [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
        a, b, c, d = P(t0, rout), P(t0, rout, rail), P(t1, rout, rail), P(t1, rout)
        tris += [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
    # close the top end
    a, b, c, d = P(0.0, rin), P(0.0, rin, rail), P(0.0, rout, rail), P(0.0, rout)
    tris += [np.concatenate([a, b, c]), np.concatenate([a, c, d])]
    verts = np.array(tris[:ntri])
    return verts, tilt, z_floor


def helical_chute(seed=3, n=1000000, r0=0.0005, spread=0.1, density=2500., young=1e7, cor=0.5,
                  mu=0.5, dt=2e-6, steps=10000, skin_factor=0.25, ntri=20000):
    """1,000,000 spheres (r ~ U[0.9, 1.1] x 0.5 mm) resting in layers along a
    4-turn helical chute of radius 0.2 m, pitch 0.1 m and width 0.05 m meshed as
    20,000 triangles.  Particles sit on a jittered face-centred lattice that
    follows the ramp, a few layers deep, so they settle and flow down the
    chute together.  The Verlet skin (0.25 r_max, refresh at skin/2) was picked
    by timing 0.1, 0.15, 0.2, 0.25 and 0.3 r_max on the settled bed; the skin
    never changes a trajectory, only how often the lists are rebuilt."""
    rng = np.random.default_rng(seed)
    verts, tilt, z_floor = helical_chute_mesh(seed=seed, ntri=ntri)
    rmax = r0 * (1 + spread)
    radius = rng.uniform(r0 * (1 - spread), rmax, size=n)
    sp = 2.05 * rmax
    rin, rout = 0.2 - 0.025, 0.2 + 0.025
    clear = 1.1 * rmax
    # square layers stacked face-centred (each layer sits in the hollows of
    # the one below: offset half a spacing, rise spacing/sqrt(2) plus a margin
    # for the ramp's slope and tilt across that offset)
    rhos = np.arange(rin + clear, rout - clear - 0.5 * sp + 1e-12, sp)
    dth = sp / rin
    total = 2 * math.pi * 4
    ths = np.arange(dth, total - 2 * dth, dth)
    nlayer = int(math.ceil(n / (len(rhos) * len(ths))))
    L, T, Rr = np.meshgrid(np.arange(nlayer), ths, rhos, indexing="ij")
    L, T, Rr = L.ravel()[:n], T.ravel()[:n], Rr.ravel()[:n]
    odd = (L % 2).astype(np.float64)
    T = T + odd * 0.5 * dth
    Rr = Rr + odd * 0.5 * sp
    zf = z_floor(T, Rr)
    z = zf + clear + L * sp * (1.0 / math.sqrt(2.0) + 0.15)
    j = 0.5 * (sp - 2 * rmax) * 0.8
    x = np.stack([Rr * np.cos(T), Rr * np.sin(T), z], axis=1)
    x = x + rng.uniform(-j, j, size=x.shape)
    lo = x.min(axis=0)
    hi = x.max(axis=0)
    bbox = bbox_rule(verts, [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    skin = skin_factor * rmax
    params = dict(dt=dt, mu=mu, g=(0., 0., -9.81), bbox=bbox, cell=2 * rmax + skin, skin=skin,
                  refresh=0.5 * skin, nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)
    spec = SceneSpec("helical_chute", x, radius, verts, params, density, young, cor, steps=steps)
    spec.tilt = tilt
    return spec


def feature_impact(kind, speed=0.5, r=0.005, side=0.05, dt=1e-6, density=2500., young=1e7, cor=0.5,
                   mu=0.5, direction=None, gap=1e-5):
    """One sphere driven head-on, from outside, into a convex mesh feature:
    kind 'vertex' -- the corner O of three right triangles that meet there at
    their right angles (faces z=0, x=0 and y=0 of an open corner of side
    `side`); kind 'edge' -- the edge from O to (0, side, 0) shared by two right
    triangles (faces z=0 and x=0).  Every triangle at the feature reports it
    as its closest feature (two barycentric weights negative for the vertex,
    one for the edge), so the wall pass sees the same vertex or edge id once
    per triangle and must keep one contact.  A far triangle comes first in the
    mesh so the feature's triangles have ids 1, 2 (, 3).  The sphere moves
    along `direction` (a unit vector into the feature), its centre on the
    line through the contact point, so the contact normal is -direction and
    no tangential motion arises.  No gravity."""
    L = side
    O = np.zeros(3)
    X, Y, Z = np.array([L, 0., 0.]), np.array([0., L, 0.]), np.array([0., 0., L])
    far = np.concatenate([[3 * L, 3 * L, 3 * L], [3.5 * L, 3 * L, 3 * L], [3 * L, 3.5 * L, 3 * L]])
    if kind == "vertex":
        tris = [far, np.concatenate([O, X, Y]), np.concatenate([O, Y, Z]), np.concatenate([O, Z, X])]
        d = np.array([1.0, 1.1, 1.25]) if direction is None else np.asarray(direction, np.float64)
        hit = O
    elif kind == "edge":
        tris = [far, np.concatenate([O, Y, X]), np.concatenate([O, Y, Z])]
        d = np.array([math.cos(0.6), 0.0, math.sin(0.6)]) if direction is None else \
            np.asarray(direction, np.float64)
        hit = np.array([0.0, 0.4 * L, 0.0])
    else:
        raise ValueError(kind)
    d = d / np.linalg.norm(d)
    x = (hit - (r + gap) * d)[None, :]
    v = (speed * d)[None, :]
    verts = np.array(tris)
    lo = np.minimum(x[0] - 2 * r, -2 * r)
    hi = np.full(3, L)
    bbox = bbox_rule(verts, [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    params = dict(dt=dt, mu=mu, g=(0., 0., 0.), bbox=bbox, cell=4 * r, skin=r, refresh=0.5 * r)
    spec = SceneSpec(f"{kind}_impact", x, np.array([r]), verts, params, density, young, cor, v=v)
    spec.direction = d
    spec.feature_tris = list(range(1, len(tris)))
    return spec


def out_of_bounds_flight(dt=1e-5, g=-9.81):
    """Six spheres far from one another, no mesh, each flying out of the
    bounding box [-0.1, 0.1]^3 through a different face (+x, -x, +y, -y under
    no acceleration along their path, +z against gravity, -z with it).
    Returns the spec; the exact step at which each one crosses follows from
    the discrete ballistic sequence x_n = x0 + n v0 dt + g dt^2 n (n+1) / 2."""
    x = np.array([[0.05, 0.0, 0.0], [-0.05, 0.05, 0.05], [0.0, 0.07, -0.05],
                  [0.05, -0.06, 0.05], [-0.05, -0.05, 0.06], [0.05, 0.05, -0.04]])
    v = np.array([[0.71, 0.0, 0.0], [-0.53, 0.0, 0.0], [0.0, 0.37, 0.0],
                  [0.0, -0.61, 0.0], [0.0, 0.0, 0.9], [0.0, 0.0, -0.2]])
    r = np.full(len(x), 0.004)
    bbox = [-0.1, 0.1, -0.1, 0.1, -0.1, 0.1]
    params = dict(dt=dt, mu=0.5, g=(0., 0., g), bbox=bbox, cell=0.02, skin=0.004, refresh=0.002)
    return SceneSpec("out_of_bounds", x, r, np.zeros((0, 9)), params, 2500., 1e7, 0.5, v=v)


def read_ascii_stl(paths):
    """ASCII STL files -> (verts [ntri,9], tris_per_file), reading facets the
    way load_ascii_stl_double does (TriangleMesh.cu:363-461)."""
    if isinstance(paths, str):
        paths = [paths]
    tris, counts = [], []
    for p in paths:
        cnt = 0
        with open(p) as f:
            lines = f.readlines()
        i = 0
        while i < len(lines):
            if lines[i].strip().startswith("facet normal"):
                vs = []
                for k in range(2, 5):
                    parts = lines[i + k].split()
                    vs += [float(parts[1]), float(parts[2]), float(parts[3])]
                tris.append(vs)
                cnt += 1
                i += 7
            else:
                i += 1
        counts.append(cnt)
    return np.array(tris, dtype=np.float64).reshape(-1, 9), counts


def _seed(master, idx):
    """Per-scene seed derived from (master seed, scene index)."""
    return np.random.SeedSequence([int(master), int(idx)])


def box_scene(seed, n, r0=0.001, r_lo=0.9, r_hi=1.1, phi=0.3, mu=0.5, cor=0.5, density=2500., young=1e7,
              dt=5e-6, skin_factor=0.1, steps=5000, name="box"):
    """n spheres, r ~ U[r_lo, r_hi] r0, at rest in a closed cube (12
    triangles) whose volume makes the spheres a fraction `phi` of it; gravity
    -z.  Spheres sit on jittered square layers filled from the floor, each
    layer in the hollows of the one below (spacing 2.05 r_max, rise
    spacing/sqrt(2)), so none overlap.  Verlet skin skin_factor * r_max,
    refresh at skin/2, cell 2 r_max + skin."""
    rng = np.random.default_rng(seed)
    radius = rng.uniform(r_lo * r0, r_hi * r0, size=n)
    rmax = r_hi * r0
    vol = float(np.sum(4.0 / 3.0 * math.pi * radius ** 3))
    side = (vol / phi) ** (1.0 / 3.0)
    h = side / 2
    clear = 1.03 * rmax
    sp = 2.05 * rmax
    span = side - 2 * clear
    dz = sp / math.sqrt(2.0)
    pts = []
    layer = 0
    while len(pts) < n:
        z = -h + clear + layer * dz
        if z > h - clear:
            raise ValueError("box too small for the requested solid fraction")
        off = 0.5 * sp if layer % 2 else 0.0
        m = int((span - off) / sp) + 1
        c = -h + clear + off + np.arange(m) * sp
        Y, X = np.meshgrid(c, c, indexing="ij")
        pts.extend(zip(X.ravel(), Y.ravel(), np.full(X.size, z)))
        layer += 1
    x = np.array(pts[:n], dtype=np.float64)
    # jitter small enough that neither a diagonal neighbour nor a wall is reached
    j = 0.9 * min(sp - 2 * rmax, 2 * (clear - rmax)) / (2 * math.sqrt(3.0))
    x = x + rng.uniform(-j, j, size=x.shape)
    verts = box_mesh((-h, -h, -h), (h, h, h))
    bbox = bbox_rule(verts, [-h, h, -h, h, -h, h])
    skin = skin_factor * rmax
    params = dict(dt=dt, mu=mu, g=(0., 0., -9.81), bbox=bbox, cell=2 * rmax + skin, skin=skin,
                  refresh=0.5 * skin, nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)
    return SceneSpec(name, x, radius, verts, params, density, young, cor, steps=steps)


def batch_mix(seed=4, nscenes=256, nmin=1000, nmax=50000, steps=5000, **kw):
    """The batched workload: independent box scenes, sphere counts log-uniform
    over [nmin, nmax], friction ~ U[0.2, 0.8] and restitution ~ U[0.3, 0.9]
    per scene, cubes sized to 30 % solid fraction."""
    out = []
    for i in range(nscenes):
        rng = np.random.default_rng(_seed(seed, i))
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        mu = float(rng.uniform(0.2, 0.8))
        cor = float(rng.uniform(0.3, 0.9))
        out.append(box_scene(_seed(seed, 100000 + i), n, mu=mu, cor=cor, steps=steps, name=f"box{i}", **kw))
    return out


def served_stream(seed=5, nreq=100, rate=20.0, nmin=500, nmax=200000, smin=1000, smax=20000, **kw):
    """The served workload: a Poisson stream (mean `rate` requests/s) of
    drop-into-box scenes with sphere counts log-uniform over [nmin, nmax],
    r ~ U[0.8, 1.2] r0 and U[smin, smax] steps each.  Returns a list of
    (arrival_time_s, spec, nsteps)."""
    rng = np.random.default_rng(_seed(seed, 0))
    t = 0.0
    out = []
    for i in range(nreq):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        nsteps = int(rng.integers(smin, smax + 1))
        spec = box_scene(_seed(seed, 100000 + i), n, r_lo=0.8, r_hi=1.2, phi=0.2, steps=nsteps,
                         name=f"req{i}", **kw)
        out.append((t, spec, nsteps))
    return out


def from_settings(path):
    """The scene a demSettings.json describes, as main.cu builds it
    (main.cu:57-227): material from particleTypes.particle0, walls from the
    listed ASCII STL files, the inlet's particles, the bounding box of
    calc_BoundingBoxLimits, cell 4 r, skin r, refresh r/2
    (BoundingBox.cu:381-388) and dt from dtFactor/outputTiming.  Returns
    (spec, settings dict)."""
    from . import engine
    st = engine.load_settings(path)
    x, v = engine.settings_particles(path, st["n"])
    verts, counts = read_ascii_stl(st["walls"]) if st["walls"] else (np.zeros((0, 9)), [])
    r = np.full(st["n"], st["radius"])
    bbox = bbox_rule(verts if len(verts) else None, st["inlet_box"])
    params = dict(dt=st["dt"], mu=st["mu"], g=tuple(st["g"]), bbox=bbox, cell=4 * st["radius"],
                  skin=st["radius"], refresh=0.5 * st["radius"])
    spec = SceneSpec(os.path.basename(os.path.dirname(os.path.abspath(path))), x, r, verts, params,
                     st["density"], 0.0, st["cor"], v=v, steps=st["steps"])
    spec.tris_per_file = counts
    spec.material = (np.full(st["n"], st["m"]), np.full(st["n"], st["k"]), np.full(st["n"], st["etaconst"]))
    return spec, st
