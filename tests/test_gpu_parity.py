"""GPU parity: the HIP engine against the CPU oracle through the C ABI.

Integer and index results (cell coordinates, Morton keys, sorted order,
candidate lists, history keys) must be identical.  FP64 forces, histories and
trajectories are compared bit for bit; the stated bound is 1e-12 relative.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from conftest import golden
from gpudem_amd import GDEM_FP64, GDEM_MIXED, Scene, engine, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(ctx, spec, precision=GDEM_FP64, **extra):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision, **extra)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params, v=spec.v)
    orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params(), v=spec.v)
    return sc, orc


def assert_state_equal(sc, orc, what=""):
    d = sc.download()
    x = orc.get("x").reshape(-1, 3)
    v = orc.get("v").reshape(-1, 3)
    w = orc.get("angv").reshape(-1, 3)
    a = orc.get("isActive")
    bad = np.where(np.any(d["x"] != x, axis=1))[0]
    assert len(bad) == 0, f"{what}: x differs for {len(bad)} particles, first {bad[:5]}, " \
                          f"max |dx| {np.abs(d['x'] - x).max()}"
    assert np.array_equal(d["v"], v), what
    assert np.array_equal(d["angv"], w), what
    assert np.array_equal(d["active"], a), what


def _rows_equal(n_eng, rows_eng, n_orc, rows_orc, live, what):
    """Row counts equal wherever `live`, and the first count entries of every
    such row equal (vectorised over all particles)."""
    bad = np.flatnonzero(live & (n_eng != n_orc))
    assert len(bad) == 0, f"{what}: row lengths differ for {len(bad)} particles, first {bad[:5]}"
    cap = rows_eng.shape[1]
    mask = (np.arange(cap)[None, :] < n_eng[:, None]) & live[:, None]
    diff = (rows_eng != rows_orc[:, :cap])
    if diff.ndim == 3:
        diff = np.any(diff, axis=2)
    bad = np.flatnonzero(np.any(diff & mask, axis=1))
    assert len(bad) == 0, f"{what}: rows differ for {len(bad)} particles, first {bad[:5]}"


def assert_lists_equal(sc, orc):
    dbg = sc.debug()
    live = orc.get("isActive") == 1
    N = orc.N
    _rows_equal(dbg["nnb"], dbg["nb"], orc.get("numNei"), orc.get("neiList").reshape(N, -1), live, "pair list")
    _rows_equal(dbg["nwall"], dbg["wall"], orc.get("numNeiWall"), orc.get("neiListWall").reshape(N, -1), live,
                "wall list")
    return dbg


def assert_history_equal(sc, orc):
    dbg = sc.debug()
    N = orc.N
    live = np.ones(N, bool)
    for sfx, kk, dd, nn, names in (("", "hist_key", "hist_delta", "nhist", ("dhx", "dhy", "dhz")),
                                   ("Wall", "whist_key", "whist_delta", "nwhist", ("dwx", "dwy", "dwz"))):
        num = orc.get("numCont" + sfx)
        _rows_equal(dbg[nn], dbg[kk], num, orc.get("indHis" + sfx).reshape(N, -1), live, "history keys" + sfx)
        d = np.stack([orc.get(n).reshape(N, -1) for n in names], axis=-1)
        _rows_equal(dbg[nn], dbg[dd], num, d, live, "history deltas" + sfx)


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------


def test_fp64_sqrt_and_division_are_correctly_rounded(ctx):
    rng = np.random.default_rng(0)
    a = np.exp(rng.uniform(-700, 700, 200000)) * rng.uniform(0.5, 2.0, 200000)
    b = np.exp(rng.uniform(-300, 300, 200000))
    a[:5] = [0.0, 1e-310, 2.0, 4.0, 1e-20]
    s, d = ctx.fp64_probe(a, b)
    assert np.array_equal(s, np.sqrt(a))
    with np.errstate(over="ignore", under="ignore"):
        assert np.array_equal(d, a / b)


@pytest.mark.parametrize("n", [1, 255, 257, 4099, 100000, 1000003])
def test_radix_sort_is_stable_and_exact(ctx, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 30, size=n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # many duplicates
    if n > 10:
        keys[-5:] = 0xFFFFFFFF
    k, idx = ctx.radix_sort(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(idx, ref)
    assert np.array_equal(k, keys[ref])
    ko, io = oracle.radix_sort(keys)
    assert np.array_equal(io, idx)


def test_tri_distance_matches_oracle(ctx):
    v, counts = scenes.read_ascii_stl(golden("spiral.stl"))
    m = oracle.Mesh(v, counts)
    rng = np.random.default_rng(7)
    for tri in rng.choice(len(v), 40, replace=False):
        c = v[tri].reshape(3, 3)
        span = np.linalg.norm(c[1] - c[0]) + np.linalg.norm(c[2] - c[0])
        pts = c.mean(0) + rng.normal(size=(500, 3)) * span
        rsq = rng.uniform(0.1, 2.0, 500) * span * span
        out, hit = ctx.tri_distance(v, int(tri), pts, rsq)
        for i in range(0, 500, 7):
            n, d, h = m.tri_distance(int(tri), pts[i], rsq[i])
            assert h == hit[i] or d == 1e10
            assert d == out[i, 3]
            assert np.array_equal(n, out[i, :3])


# ---------------------------------------------------------------------------
# plain FP64 scenes: keys, lists, forces, histories, trajectories
# ---------------------------------------------------------------------------


def assert_rebuild_store(sc, orc, prev_order):
    """The keys, slot order and cell table the engine's last rebuild stored
    equal the oracle's Morton keys (dk_morton_key, BoundingBox.cu:1007-1023)
    and cells (d_calcCellId, :1089-1104) at the rebuild positions, the stable
    sort of the previous slot order by those keys (SortPairs, :1235-1242), and
    the oracle's per-cell counts (CSR, :1062-1087)."""
    o = sc.debug_order()
    key_of = orc.morton_built()
    assert np.array_equal(o["key"], key_of[o["order"]])
    expect = prev_order[np.argsort(key_of[prev_order], kind="stable")]
    assert np.array_equal(o["order"], expect)
    nact = int(np.sum(orc.get("isActive") == 1))
    assert o["n_sorted"] == nact
    held = o["order"][:nact]
    assert np.array_equal(o["cell"][:nact], orc.get("cellId")[held])
    assert np.all(o["cell"][nact:] == -1) and np.all(o["key"][nact:] == 0xFFFFFFFF)
    k = o["key"][:nact]
    assert np.all(np.diff(k.astype(np.int64)) >= 0)
    starts = np.r_[0, np.flatnonzero(np.diff(k) != 0) + 1]
    ends = np.r_[starts[1:], nact]
    run = np.repeat(np.arange(len(starts)), ends - starts)
    assert np.array_equal(o["cell_first"][:nact], starts[run])
    assert np.array_equal(o["cell_end"][:nact], ends[run])
    assert np.array_equal(ends - starts, orc.get("pNum")[o["cell"][starts]])
    return o["order"].copy()


def test_cells_keys_and_lists_at_setup(ctx):
    spec = scenes.polydisperse_bed(n=5000)
    say("scene polydisperse_bed n=5000")
    sc, orc = make_pair(ctx, spec)
    dbg = assert_lists_equal(sc, orc)
    assert np.array_equal(dbg["cellx"], orc.get("cellx").reshape(-1, 3))
    assert np.array_equal(dbg["cell_id"], orc.get("cellId"))
    assert np.array_equal(dbg["morton"], orc.morton())
    # what the production rebuild stored: the stable sort of the ids by key
    order = assert_rebuild_store(sc, orc, np.arange(spec.n, dtype=np.int32))
    k, ref = oracle.radix_sort(orc.morton_built())
    assert np.array_equal(order, ref)


def test_per_contact_forces_bitwise(ctx):
    spec = scenes.lattice_drop(verlet=True)
    say("scene lattice_drop verlet")
    sc, orc = make_pair(ctx, spec)
    orc.record(True)
    seen = 0
    for block in range(20):
        sc.step(999)
        orc.step(999)
        rec = sc.step_record(cap=64)
        orc.step(1)
        oc = orc.contacts()
        for i in range(orc.N):
            n = rec["count"][i]
            assert n == len(oc[i]["key"]), (block, i)
            assert np.array_equal(rec["key"][i, :n], oc[i]["key"])
            assert np.array_equal(rec["kind"][i, :n], oc[i]["kind"])
            assert np.array_equal(rec["fn"][i, :n], oc[i]["fn"]), (block, i)
            assert np.array_equal(rec["ft"][i, :n], oc[i]["ft"]), (block, i)
            assert np.array_equal(rec["delta"][i, :n], oc[i]["delta"])
            seen += n
    assert seen > 300  # the bed settles into many pair and wall contacts
    assert_state_equal(sc, orc, "after recorded steps")


@pytest.mark.parametrize("verlet", [False, True])
def test_lattice_drop_trajectory_bitwise(ctx, verlet):
    # 100 spheres, 20,000 steps (the whole drop-and-settle run)
    spec = scenes.lattice_drop(verlet=verlet)
    say(f"scene lattice_drop verlet={verlet} 20000 steps")
    sc, orc = make_pair(ctx, spec)
    for chunk in range(4):
        sc.step(5000)
        orc.step(5000)
        assert_state_equal(sc, orc, f"step {(chunk + 1) * 5000}")
    assert_history_equal(sc, orc)
    assert_lists_equal(sc, orc)
    st = sc.status()
    assert st["errors"] == 0 and st["steps"] == 20000
    assert st["mean_contacts"] > 0.3


def test_polydisperse_bed_trajectory_bitwise(ctx):
    # C2: 30,000 polydisperse spheres settling in a box, 5,000 steps bit for
    # bit; after every rebuild the stored keys, order and cells are checked
    spec = scenes.polydisperse_bed(n=30000)
    steps, chunk = 5000, 10
    say(f"scene polydisperse_bed n=30000 {steps} steps")
    sc, orc = make_pair(ctx, spec)
    order = assert_rebuild_store(sc, orc, np.arange(spec.n, dtype=np.int32))
    checked = 0
    for c in range(steps // chunk):
        rb = orc.get("rebuilds")
        sc.step(chunk)
        orc.step(chunk)
        if orc.get("rebuilds") == rb + 1:
            order = assert_rebuild_store(sc, orc, order)
            checked += 1
        elif orc.get("rebuilds") > rb + 1:
            order = sc.debug_order()["order"].copy()
        if (c + 1) * chunk % 500 == 0:
            assert_state_equal(sc, orc, f"step {(c + 1) * chunk}")
    say(f"  rebuilds {orc.get('rebuilds')}, {checked} checked")
    assert checked >= 3
    assert_lists_equal(sc, orc)
    assert_history_equal(sc, orc)
    st = sc.status()
    assert st["errors"] == 0 and st["rebuilds"] == orc.get("rebuilds")
    assert st["mean_contacts"] > 0.1  # the lower layers have landed and touch


@pytest.mark.parametrize("broadphase", [1, 2])
def test_spiral_mesh_walls_bitwise(ctx, broadphase):
    # 2,498-triangle reference mesh with obtuse facets: wall lists, vertex/edge
    # dedup and wall histories.  Spheres start just above facet centroids.
    v, counts = scenes.read_ascii_stl(golden("spiral.stl"))
    rng = np.random.default_rng(11)
    tri = v.reshape(-1, 3, 3)
    cen = tri.mean(axis=1)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    r0 = 0.012
    pick = rng.choice(len(tri), 1500, replace=False)
    x = cen[pick] + nrm[pick] * r0 * 1.02
    keep = np.ones(len(x), bool)
    for i in range(len(x)):  # drop overlapping starts
        if keep[i]:
            d = np.linalg.norm(x[i + 1:] - x[i], axis=1)
            keep[i + 1:][d < 2.05 * r0] = False
    x = x[keep]
    r = np.full(len(x), r0)
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    bbox = scenes.bbox_rule(v, [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]])
    params = dict(dt=1e-5, mu=0.3, g=(0., -9.81, 0.), bbox=bbox, cell=0.048, skin=0.012, refresh=0.006,
                  wall_cap=48, whist_cap=32, wall_broadphase=broadphase)
    spec = scenes.SceneSpec("spiral", x, r, v, params, 1000., 1e7, 0.3)
    say(f"scene spiral n={len(x)}")
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    sc = Scene(ctx, x, r, m, k, e, v, params, tris_per_file=counts)
    orc = oracle.Sim(x, r, oracle.Mesh(v, counts), spec.oracle_params())
    assert_lists_equal(sc, orc)
    for chunk in range(4):
        sc.step(500)
        orc.step(500)
        st = sc.status(check=False)
        assert st["errors"] == 0 and orc.get("overflow") == 0, st
        assert_state_equal(sc, orc, f"step {(chunk + 1) * 500}")
    assert_history_equal(sc, orc)
    assert sc.status()["mean_contacts"] > 0.2


def test_inactive_particles_stay_put(ctx):
    # particles flagged inactive at creation take no part in contacts and are
    # never integrated (cpu_dem.cu:1084-1089, device_dem.cu:1065-1068)
    spec = scenes.lattice_drop(verlet=True)
    act = np.ones(spec.n, np.int32)
    act[::7] = 0
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, spec.params, active=act)
    orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params(), active=act)
    sc.step(3000)
    orc.step(3000)
    assert_state_equal(sc, orc, "inactive")
    d = sc.download()
    assert np.array_equal(d["active"], act)
    assert np.array_equal(d["x"][act == 0], spec.x[act == 0])


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_out_of_bounds_on_closed_form_step(ctx, precision):
    # six spheres leave the box through its six faces; each must turn inactive
    # on the step the discrete ballistic sequence crosses the face
    # (dk_checkOoB, device_dem.cu:1116-1125; checkOoB, cpu_dem.cu:1575-1589)
    from test_oracle_features import crossing_steps
    spec = scenes.out_of_bounds_flight()
    steps, margin = crossing_steps(spec)
    assert np.all(margin > 1e-9)
    say(f"scene out_of_bounds precision {precision}, crossings at {sorted(steps.tolist())}")
    sc, orc = make_pair(ctx, spec, precision)
    done = 0
    for k in np.sort(steps):
        for n in (int(k) - 1 - done, 1):
            sc.step(n)
            orc.step(n)
            done += n
            d = sc.download()
            assert np.array_equal(d["active"], (steps > done).astype(np.int32)), (done, d["active"])
            assert np.array_equal(d["active"], orc.get("isActive"))
            if precision == GDEM_FP64:
                assert_state_equal(sc, orc, f"step {done}")
    assert sc.status()["n_active"] == 0
    x = sc.download()["x"]
    sc.step(50)
    assert np.array_equal(sc.download()["x"], x)


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
@pytest.mark.parametrize("kind", ["vertex", "edge"])
def test_feature_contact_counted_once(ctx, kind, precision):
    # A sphere driven into a convex vertex (three triangles) or edge (two):
    # every triangle reports the same feature id, the wall pass keeps the first
    # (d_wall_collision_verlet, device_dem.cu:288-344), so one contact acts,
    # its history key is the lowest triangle id, and the rebound follows the
    # single-contact closed form e^(pi/3.1415).
    spec = scenes.feature_impact(kind)
    say(f"scene {kind}_impact precision {precision}")
    sc, orc = make_pair(ctx, spec, precision)
    first = min(spec.feature_tris)
    seen = 0
    for _ in range(40):
        sc.step(100)
        orc.step(100)
        if precision == GDEM_FP64:
            assert_state_equal(sc, orc, kind)
            assert_history_equal(sc, orc)
        dbg = sc.debug()
        if dbg["nwhist"][0]:
            assert dbg["nwhist"][0] == 1 and dbg["whist_key"][0, 0] == first
            assert sorted(dbg["wall"][0, :dbg["nwall"][0]].tolist()) == spec.feature_tris
            seen += 1
    assert seen >= 10
    v = sc.download()["v"][0]
    e = np.linalg.norm(v) / np.linalg.norm(spec.v[0])
    assert e == pytest.approx(0.5 ** (math.pi / 3.1415), rel=2e-3)
    assert np.allclose(v / np.linalg.norm(v), -spec.direction, atol=1e-6)


def test_out_of_bounds_is_strict(ctx):
    # on the face is inside; the first position beyond it is outside
    hi = 0.1
    x = np.array([[hi, 0.0, 0.0], [0.0, hi - 2e-9, 0.0], [0.0, 0.0, 0.05]])
    v = np.zeros((3, 3))
    v[1, 1] = 1e-3  # one step of 1e-5 s moves it 1e-8 m, past the face
    params = dict(dt=1e-5, mu=0.5, g=(0., 0., 0.), bbox=[-0.1, hi, -0.1, hi, -0.1, hi], cell=0.02,
                  skin=0.004, refresh=0.002)
    spec = scenes.SceneSpec("strict_oob", x, np.full(3, 0.004), np.zeros((0, 9)), params, 2500., 1e7, 0.5, v=v)
    sc, orc = make_pair(ctx, spec)
    sc.step(1)
    orc.step(1)
    assert list(sc.download()["active"]) == [1, 0, 1] == list(orc.get("isActive"))
    sc.step(10)
    orc.step(10)
    assert_state_equal(sc, orc, "strict")


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_refresh_steps_match_oracle_near_threshold(ctx, precision):
    # 64 spheres in free flight (no gravity, never within reach of each other
    # or a wall) cross the refresh distance again and again.  Each step moves
    # them 0.1-0.2 um, so the refresh test often falls within the error of the
    # packed reference position (about 0.05 um on this 0.11 m box), where the
    # exact position decides.  With no forces the mixed form integrates
    # exactly like FP64, so both must rebuild on the oracle's steps.
    rng = np.random.default_rng(21)
    g = (np.arange(4) - 1.5) * 0.012
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    dirs = rng.normal(size=x.shape)
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    v = dirs * rng.uniform(0.01, 0.02, len(x))[:, None]
    r0, dt, steps = 1e-3, 1e-5, 20000
    verts = scenes.box_mesh((-0.05,) * 3, (0.05,) * 3)
    bbox = scenes.bbox_rule(verts, [-0.05, 0.05] * 3)
    params = dict(dt=dt, mu=0.5, g=(0., 0., 0.), bbox=bbox, cell=2.5 * r0, skin=0.5 * r0, refresh=0.25 * r0)
    spec = scenes.SceneSpec("free_flight", x, np.full(len(x), r0), verts, params, 2500., 1e7, 0.5, v=v)
    say(f"scene free_flight precision {precision} {steps} steps")
    sc, orc = make_pair(ctx, spec, precision)
    sc.step(steps)
    orc.step(steps)
    assert_state_equal(sc, orc, "free flight")
    st = sc.status()
    assert st["errors"] == 0 and st["rebuilds"] == orc.get("rebuilds"), (st, orc.get("rebuilds"))
    # the same flight on the host: the oracle's rebuild rule, and how many
    # tests fell within the packed position's error of the threshold
    lo, hi = np.array(bbox[0::2]), np.array(bbox[1::2])
    h = (hi - lo) * (1.0 / 2097151.0)
    eps = 0.5 * np.sqrt(np.sum(h * h))
    t = 0.25 * r0
    p, ref, rebuilds, near = x.copy(), x.copy(), 0, 0
    for _ in range(steps):
        p = p + v * dt
        q = lo + np.rint((ref - lo) / h) * h
        near += int(np.sum(np.abs(np.linalg.norm(p - q, axis=1) - t) < eps))
        if np.any(np.sum((p - ref) ** 2, axis=1) > t * t):
            ref = p.copy()
            rebuilds += 1
    say(f"  rebuilds {st['rebuilds']} (host {rebuilds}), tests near the threshold {near}")
    assert rebuilds == st["rebuilds"] >= 5 and near >= 5


GDEM_ERR_NB_OVERFLOW, GDEM_ERR_WALL_OVERFLOW, GDEM_ERR_HIST_OVERFLOW, GDEM_ERR_WHIST_OVERFLOW = 1, 2, 4, 8
GDEM_E_SLOT_OVERFLOW = -4


def corner_cluster():
    """A 4x4x4 cubic lattice of r = 1 mm spheres, spacing 1.98 mm (every face
    neighbour overlaps by 0.02 mm), in the corner of a 20 mm box: the lattice's
    corner sphere overlaps the three walls there by 0.01 mm."""
    r0, s = 1e-3, 1.98e-3
    g = 0.99 * r0 + s * np.arange(4)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    verts = scenes.box_mesh((0., 0., 0.), (0.02, 0.02, 0.02))
    bbox = scenes.bbox_rule(verts, [0., 0.02, 0., 0.02, 0., 0.02])
    # skin 0.2 r: the face neighbours are candidates, the diagonal ones are not
    params = dict(dt=1e-6, mu=0.5, g=(0., 0., -9.81), bbox=bbox, cell=2.2 * r0, skin=0.2 * r0, refresh=0.1 * r0)
    return scenes.SceneSpec("corner_cluster", x, np.full(len(x), r0), verts, params, 2500., 1e7, 0.5)


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_capacity_overflow_is_reported(ctx, precision):
    # Rows shorter than the packing needs: the list builders and the step keep
    # to their rows and report the overflow through the status bits (the
    # reference prints a warning, BoundingBox.cu:1146-1148, BVH.cu:526-528).
    spec = corner_cluster()
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    cases = [(dict(), 0),
             (dict(nb_cap=2, hist_cap=1), GDEM_ERR_NB_OVERFLOW | GDEM_ERR_HIST_OVERFLOW),
             (dict(wall_cap=1), GDEM_ERR_WALL_OVERFLOW),
             (dict(whist_cap=1), GDEM_ERR_WHIST_OVERFLOW)]
    for caps, bits in cases:
        say(f"scene corner_cluster precision {precision} caps {caps}")
        sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision, **caps))
        sc.step(5)
        st = sc.status(check=False)
        assert st["errors"] == bits, (caps, st)
        assert st["rc"] == (GDEM_E_SLOT_OVERFLOW if bits else 0), (caps, st)
        d = sc.download()
        assert np.all(np.isfinite(d["x"])) and np.all(d["active"] == 1), caps
        dbg = sc.debug()
        assert np.all(dbg["nnb"] <= sc.caps["nb"]) and np.all(dbg["nwall"] <= sc.caps["wall"]), caps
        assert np.all(dbg["nhist"] <= sc.caps["hist"]) and np.all(dbg["nwhist"] <= sc.caps["whist"]), caps
        sc.close()


# ---------------------------------------------------------------------------
# mixed precision against FP64
# ---------------------------------------------------------------------------


def test_mixed_precision_force_error(ctx):
    spec = scenes.polydisperse_bed(n=30000)
    say("scene polydisperse_bed mixed vs fp64")
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    s64 = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, spec.params)
    s64.step(20000)
    st = s64.download()
    # same state into both precisions, one recorded step each
    a = Scene(ctx, st["x"], spec.r, m, k, e, spec.verts, spec.params, v=st["v"], angv=st["angv"],
              active=st["active"])
    b = Scene(ctx, st["x"], spec.r, m, k, e, spec.verts, dict(spec.params, precision=GDEM_MIXED),
              v=st["v"], angv=st["angv"], active=st["active"])
    ra = a.step_record()
    rb = b.step_record()
    assert np.array_equal(ra["count"], rb["count"])
    assert ra["count"].sum() > 30000
    f64 = np.concatenate([ra["fn"][i, :c] + ra["ft"][i, :c] for i, c in enumerate(ra["count"])])
    f32 = np.concatenate([rb["fn"][i, :c] + rb["ft"][i, :c] for i, c in enumerate(rb["count"])])
    rel = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    say(f"mixed aggregate force rel err {rel:.3e}")
    assert rel <= 1e-4


def test_mixed_precision_restitution(ctx):
    r = 0.005
    verts = scenes.box_mesh((-0.05, -0.05, -0.05), (0.05, 0.05, 0.05))
    params = dict(dt=2e-6, mu=0.5, g=(0., 0., 0.), bbox=scenes.bbox_rule(verts, [-0.05, 0.05] * 3),
                  cell=4 * r, skin=r, refresh=r / 2, precision=GDEM_MIXED)
    x = np.array([[0.0, 0.0, -0.05 + r + 1e-5]])
    m, k, e = engine.material([r], 2500., 1e7, 0.5)
    sc = Scene(ctx, x, [r], m, k, e, verts, params, v=np.array([[0, 0, -1.0]]))
    sc.step(3000)
    got = sc.download()["v"][0, 2]
    expect = 0.5 ** (math.pi / 3.1415)
    assert got == pytest.approx(expect, rel=3e-3)


# ---------------------------------------------------------------------------
# full-size chute (1,000,000 spheres, 20,000 triangles): size-independent checks
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def chute():
    return scenes.helical_chute()


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_chute_full_size_properties(ctx, chute, precision):
    spec = chute
    say(f"scene helical_chute n={spec.n} precision={precision}")
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision)
    a = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    a.step(300)
    st = a.status()
    assert st["errors"] == 0 and st["n_active"] > 0.99 * spec.n
    dbg = a.debug()
    da = a.download()
    # candidate lists: sorted, symmetric, and exact for a sample against brute force
    rng = np.random.default_rng(0)
    sample = rng.choice(np.where(da["active"] == 1)[0], 200, replace=False)
    x, r = da["x"], spec.r
    act = da["active"] == 1
    skin = params["skin"]
    for i in sample:
        dx = x[i, 0] - x[:, 0]
        dy = x[i, 1] - x[:, 1]
        dz = x[i, 2] - x[:, 2]
        # lists were built at the last refresh, so compare at the built positions:
        # every listed partner must be within reach (lists are supersets of contacts)
        lst = dbg["nb"][i, :dbg["nnb"][i]]
        assert np.all(np.diff(lst) > 0)
        dd = dx * dx + dy * dy + dz * dz
        R = r[i] + r + skin
        near = np.where((dd < (r[i] + r) ** 2) & act)[0]
        near = near[near != i]
        assert set(near.tolist()) <= set(lst.tolist())
        for j in lst[:5]:
            assert i in set(dbg["nb"][j, :dbg["nnb"][j]].tolist())
        assert np.all(R[lst] > 0)
    # determinism: a second scene reproduces the trajectory bit for bit
    b = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    b.step(300)
    db = b.download()
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["angv"], db["angv"])
    a.close()
    b.close()


def test_chute_wall_table_matches_bvh(ctx, chute):
    # static per-cell table and BVH traversal give identical candidate lists
    spec = chute
    say("scene helical_chute table vs bvh")
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    a = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, wall_broadphase=2,
                                                             precision=GDEM_MIXED))
    b = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, wall_broadphase=1,
                                                             precision=GDEM_MIXED))
    a.step(2000)
    b.step(2000)
    da, db = a.debug(), b.debug()
    assert np.array_equal(da["nwall"], db["nwall"]) and np.array_equal(da["wall"], db["wall"])
    assert da["nwall"].sum() > 10000
    assert np.array_equal(a.download()["x"], b.download()["x"])
    a.close()
    b.close()


def test_graph_replay_matches_direct_launches(ctx):
    spec = scenes.polydisperse_bed(n=8000)
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    a = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, spec.params)
    b = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, graph_steps=25))
    for _ in range(4):
        a.step(250)
        b.step(250)
        assert np.array_equal(a.download()["x"], b.download()["x"])
