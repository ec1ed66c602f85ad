"""The benchmark configurations at their full size, against the oracle.

* C3: the 1,000,000-sphere, 20,000-triangle helical chute in FP64 from a
  flowing state: pair and wall lists and the stored keys, order and cells
  exact at set-up and after every rebuild, 300 steps bit for bit.
* C4: the 256-scene box mix (about 3.4 M spheres, mixed precision, hipGraph,
  the tuner's settings) for 500 steps: no scene raises an error, the
  smallest, median and largest scenes equal the same scenes run alone, and
  in an FP64 batch of the whole mix the median scene equals the oracle.
* C5: a served stream whose scenes span 500-200,000 spheres under an 8 GiB
  HBM budget in FP64: every request completes, three of them (500, 30,000
  and 200,000 spheres) equal the oracle bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle
from gpudem_amd import GDEM_FP64, GDEM_MIXED, Batch, Scene, autotune, engine, scenes
from test_gpu_parity import assert_history_equal, assert_lists_equal, assert_rebuild_store, assert_state_equal

pytestmark = pytest.mark.gpu

CAPS = dict(nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_state(orc):
    return dict(x=orc.get("x").reshape(-1, 3), v=orc.get("v").reshape(-1, 3), angv=orc.get("angv").reshape(-1, 3),
                active=orc.get("isActive"))


def assert_same(a, b, what):
    for key in ("x", "v", "angv", "active"):
        bad = np.where(np.any((a[key] != b[key]).reshape(len(a[key]), -1), axis=1))[0]
        assert len(bad) == 0, f"{what}: {key} differs for {len(bad)} particles, first {bad[:5]}"


def test_chute_1m_fp64_matches_oracle(ctx):
    spec = scenes.helical_chute()
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=GDEM_FP64)
    say(f"helical chute n={spec.n} ntri={len(spec.verts)}: 10000 FP64 steps to a flowing state")
    s0 = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    s0.step(10000)
    st = s0.download()
    s0.close()
    sc = Scene(ctx, st["x"], spec.r, m, k, e, spec.verts, params, v=st["v"], angv=st["angv"], active=st["active"])
    orc = oracle.Sim(st["x"], spec.r, oracle.Mesh(spec.verts), spec.oracle_params(), v=st["v"], angv=st["angv"],
                     active=st["active"])
    say("  lists and rebuild store at set-up")
    dbg = assert_lists_equal(sc, orc)
    assert dbg["nnb"].sum() > 5 * spec.n and dbg["nwall"].sum() > 0.1 * spec.n
    order = assert_rebuild_store(sc, orc, np.arange(spec.n, dtype=np.int32))
    checked = 0
    for c in range(6):
        rb = orc.get("rebuilds")
        sc.step(50)
        orc.step(50)
        if orc.get("rebuilds") == rb + 1:
            order = assert_rebuild_store(sc, orc, order)
            checked += 1
        elif orc.get("rebuilds") > rb + 1:
            order = sc.debug_order()["order"].copy()
        assert_state_equal(sc, orc, f"step {(c + 1) * 50}")
        say(f"  step {(c + 1) * 50}: bitwise, rebuilds {orc.get('rebuilds')}")
    assert checked >= 1
    assert_lists_equal(sc, orc)
    assert_history_equal(sc, orc)
    s = sc.status()
    assert s["errors"] == 0 and s["rebuilds"] == orc.get("rebuilds") and s["mean_contacts"] > 1.0
    sc.close()


def test_batch_mix_256_at_size(ctx):
    mix = scenes.batch_mix(nscenes=256, steps=500)
    tuned = dict(autotune.DEFAULT)
    nsteps = 500
    total = sum(s.n for s in mix)
    say(f"batch mix: {len(mix)} scenes, {total} spheres, {tuned}")
    by_n = sorted(range(len(mix)), key=lambda i: mix[i].n)
    pick = [by_n[0], by_n[len(by_n) // 2], by_n[-1]]

    def run(precision):
        b = Batch(ctx, precision=precision, max_scenes=len(mix), particle_capacity=total + 64 * len(mix),
                  graph_steps=tuned["graph_steps"], **CAPS)
        ids = []
        for s in mix:
            m, k, e = engine.material(s.r, s.density, s.young, s.cor)
            ids.append(b.admit(s.x, s.r, m, k, e, s.verts,
                               autotune.with_tuning(s, tuned["skin_factor"], tuned["cell_factor"]), nsteps, v=s.v))
        assert all(i is not None for i in ids)
        b.step(nsteps)
        assert sorted(b.poll()) == sorted(ids)
        stats = [b.status(i, check=False) for i in ids]
        out = {i: b.download(ids[i]) for i in pick}
        b.close()
        return stats, out

    stats, got = run(GDEM_MIXED)
    assert all(s["errors"] == 0 and s["steps_done"] == nsteps and s["rebuilds"] >= 1 for s in stats)
    for i in pick:
        s = mix[i]
        m, k, e = engine.material(s.r, s.density, s.young, s.cor)
        sc = Scene(ctx, s.x, s.r, m, k, e, s.verts,
                   dict(autotune.with_tuning(s, tuned["skin_factor"], tuned["cell_factor"]), precision=GDEM_MIXED,
                        **CAPS), v=s.v)
        sc.step(nsteps)
        say(f"  scene {i} n={s.n}: batch vs alone")
        assert_same(got[i], sc.download(), f"scene {i}")
        sc.close()
    stats, got = run(GDEM_FP64)
    assert all(s["errors"] == 0 for s in stats)
    i = pick[1]
    s = mix[i]
    say(f"  FP64 batch: scene {i} n={s.n} vs oracle")
    orc = oracle.Sim(s.x, s.r, oracle.Mesh(s.verts), s.oracle_params(), v=s.v)
    orc.step(nsteps)
    assert_same(got[i], oracle_state(orc), f"scene {i} vs oracle")


def test_served_stream_full_range_fp64(ctx):
    from gpudem_amd import serve
    rng = np.random.default_rng(17)
    sizes = [500, 200000, 30000, 2000, 90000, 8000, 700, 50000]
    stream, t = [], 0.0
    for i, n in enumerate(sizes):
        t += float(rng.exponential(1.0 / 20.0))
        nsteps = int(rng.integers(300, 700))
        spec = scenes.box_scene(scenes._seed(77, i), n, r_lo=0.8, r_hi=1.2, phi=0.2, steps=nsteps,
                                name=f"req{i}", mu=float(rng.uniform(0.2, 0.8)), cor=float(rng.uniform(0.3, 0.9)))
        stream.append((t, spec, nsteps))
    say(f"served FP64 stream: {len(stream)} requests, N {min(sizes)}-{max(sizes)}, 8 GiB budget")
    srv = serve.Server(ctx, GDEM_FP64, hbm_budget=8 << 30, chunk=100, graph_steps=50, **CAPS)
    done = srv.run(stream)
    srv.close()
    summ = serve.summarize(done, srv.makespan)
    say(f"  {summ['value']:.3e} particle-steps/s, p50 {summ['p50_s']:.3f} s, p99 {summ['p99_s']:.3f} s")
    assert len(done) == len(stream) and summ["failed"] == 0
    for req in done:
        assert req.stats["errors"] == 0 and req.stats["steps_done"] == req.nsteps
    by = {req.rid: req for req in done}
    for rid in (0, 1, 2):  # 500, 200,000 and 30,000 spheres
        req = by[rid]
        s = req.spec
        say(f"  request {rid} n={s.n} steps={req.nsteps} vs oracle")
        orc = oracle.Sim(s.x, s.r, oracle.Mesh(s.verts), s.oracle_params(), v=s.v)
        orc.step(req.nsteps)
        assert_same(req.result, oracle_state(orc), f"request {rid}")
