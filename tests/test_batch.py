"""Batched and served scenes: every scene stepped inside a batch must end in
exactly the state it reaches when stepped alone (gdem_scene, itself
bit-identical to the oracle in FP64).  Results are compared bit for bit, in
both precisions, for scenes admitted together, admitted while others run,
and admitted into the ranges of evicted scenes.
"""
import sys

import numpy as np
import pytest

import oracle
from gpudem_amd import GDEM_FP64, GDEM_MIXED, Batch, Scene, engine, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CAPS = dict(nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)


def mk(seed, n, **kw):
    return scenes.box_scene(seed, n, **kw)


def admit(batch, spec, nsteps):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    return batch.admit(spec.x, spec.r, m, k, e, spec.verts, spec.params, nsteps, v=spec.v)


def alone(ctx, spec, precision, nsteps):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision), v=spec.v)
    sc.step(nsteps)
    st = sc.status()
    d = sc.download()
    sc.close()
    return d, st


def assert_same(a, b, what):
    for key in ("x", "v", "angv", "active"):
        bad = np.where(np.any((a[key] != b[key]).reshape(len(a[key]), -1), axis=1))[0]
        assert len(bad) == 0, f"{what}: {key} differs for {len(bad)} particles, first {bad[:5]}"


def mixed_specs():
    # uneven sizes, per-scene friction, restitution and time step
    return [mk(1, 300, mu=0.2, cor=0.3), mk(2, 1500, mu=0.8, cor=0.9, dt=4e-6), mk(3, 2600, mu=0.5, cor=0.6),
            mk(4, 700, mu=0.35, cor=0.45, dt=6e-6), mk(5, 4100, mu=0.65, cor=0.75), mk(6, 130, mu=0.5, cor=0.5)]


@pytest.mark.parametrize("precision,graph", [(GDEM_FP64, 0), (GDEM_MIXED, 100)])
def test_batch_scenes_match_alone(ctx, precision, graph):
    specs = mixed_specs()
    nsteps = 1500
    say(f"batch of {len(specs)} box scenes, precision {precision}, graph_steps {graph}, {nsteps} steps")
    b = Batch(ctx, precision=precision, max_scenes=16, particle_capacity=16384, graph_steps=graph, **CAPS)
    ids = [admit(b, s, nsteps) for s in specs]
    assert all(i is not None for i in ids)
    b.step(nsteps)
    assert sorted(b.poll()) == sorted(ids)
    for sid, spec in zip(ids, specs):
        st = b.status(sid)
        assert st["errors"] == 0 and st["steps_done"] == nsteps and st["rebuilds"] > 0
        d_alone, st_alone = alone(ctx, spec, precision, nsteps)
        say(f"  scene {spec.name} n={spec.n} rebuilds {st['rebuilds']} (alone {st_alone['rebuilds']}) "
            f"contacts {st['mean_contacts']:.2f}")
        assert_same(b.download(sid), d_alone, spec.name)
    b.close()


def test_batch_fp64_matches_oracle(ctx):
    specs = [mk(7, 400, mu=0.3, cor=0.4), mk(8, 900, mu=0.7, cor=0.8, dt=4e-6)]
    nsteps = 2000
    say(f"batch vs oracle, 2 scenes, {nsteps} steps")
    b = Batch(ctx, precision=GDEM_FP64, max_scenes=4, particle_capacity=4096, **CAPS)
    ids = [admit(b, s, nsteps) for s in specs]
    b.step(nsteps)
    for sid, spec in zip(ids, specs):
        orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params())
        orc.step(nsteps)
        ref = dict(x=orc.get("x").reshape(-1, 3), v=orc.get("v").reshape(-1, 3),
                   angv=orc.get("angv").reshape(-1, 3), active=orc.get("isActive"))
        assert_same(b.download(sid), ref, spec.name)
    b.close()


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_continuous_admission_and_eviction(ctx, precision):
    say(f"continuous batching, precision {precision}")
    b = Batch(ctx, precision=precision, max_scenes=3, particle_capacity=8192, graph_steps=50, **CAPS)
    a, bb, c, d = mk(11, 3000), mk(12, 2000, mu=0.7), mk(13, 4000), mk(14, 2500, cor=0.8)
    ia = admit(b, a, 800)
    b.step(300)
    ib = admit(b, bb, 600)  # joins while A runs
    b.step(500)
    assert b.poll() == [ia]
    assert_same(b.download(ia), alone(ctx, a, precision, 800)[0], "A")
    b.evict(ia)
    assert admit(b, c, 500) is None  # 4032 slots do not fit the fragments left (3008 + 3136)
    id_ = admit(b, d, 700)  # fits where A was
    assert id_ is not None
    e = mk(15, 500)
    ie = admit(b, e, 100)
    assert admit(b, mk(16, 300), 10) is None  # all three scene slots taken
    b.step(100)
    assert sorted(b.poll()) == sorted([ib, ie])
    assert_same(b.download(ie), alone(ctx, e, precision, 100)[0], "E")
    b.step(600)
    assert sorted(b.poll()) == sorted([ib, ie, id_])
    assert_same(b.download(ib), alone(ctx, bb, precision, 600)[0], "B")
    assert_same(b.download(id_), alone(ctx, d, precision, 700)[0], "D")
    for s in (ib, ie, id_):
        assert b.status(s)["errors"] == 0
        b.evict(s)
    assert b.info()["live"] == 0
    b.close()


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_readmission_while_a_rebuild_is_pending(ctx, precision):
    # The step kernel lists every scene that requests a rebuild, for the
    # rebuild of the next batch step.  A scene evicted right after such a
    # request must not be rebuilt (eviction scrubs its entry); a scene
    # admitted into the same scene slot before the next step must be rebuilt
    # exactly once and end as when stepped alone, and the scene running
    # beside them must be unaffected.  With velocities of 0.3 m/s per component and a refresh
    # distance of 1.1e-6 m, some sphere of the evicted scene passes it in
    # every 5e-6 s step, so the scene requests a rebuild after every step.
    fast = mk(21, 1500, skin_factor=0.002)
    fast.v = np.random.default_rng(21).normal(0.0, 0.3, size=(fast.n, 3))
    other, newc = mk(22, 2000, mu=0.7), mk(23, 1200, cor=0.8)
    say(f"re-admission with a pending rebuild, precision {precision}")
    b = Batch(ctx, precision=precision, max_scenes=2, particle_capacity=8192, graph_steps=0, **CAPS)
    ia = admit(b, fast, 400)
    ib = admit(b, other, 700)
    b.step(300)
    st = b.status(ia)
    # the admission's build and one for each of steps 0..298; step 299's
    # request is the pending one
    assert st["errors"] == 0 and st["rebuilds"] == 300, st
    b.evict(ia)
    ic = admit(b, newc, 400)
    assert ic == ia  # the only free scene slot
    b.step(400)
    assert sorted(b.poll()) == sorted([ib, ic])
    assert b.status(ic)["errors"] == 0 and b.status(ib)["errors"] == 0
    assert_same(b.download(ic), alone(ctx, newc, precision, 400)[0], "re-admitted")
    assert_same(b.download(ib), alone(ctx, other, precision, 700)[0], "neighbour")
    b.close()


def test_skin_and_cell_do_not_change_results(ctx):
    # lists are supersets of the contacts whenever refresh <= skin/2, so the
    # tuner's skin/cell choice never changes a trajectory
    base = mk(21, 2000, mu=0.4, cor=0.5)
    nsteps = 1200
    say("skin invariance")
    b = Batch(ctx, precision=GDEM_FP64, max_scenes=4, particle_capacity=8192, **CAPS)
    ids = []
    for f, cellf in ((0.1, 1.0), (0.3, 1.0), (0.3, 1.5)):
        spec = mk(21, 2000, mu=0.4, cor=0.5, skin_factor=f)
        spec.params["cell"] *= cellf
        ids.append(admit(b, spec, nsteps))
    b.step(nsteps)
    ref = b.download(ids[0])
    reb = [b.status(i)["rebuilds"] for i in ids]
    say(f"  rebuilds per skin choice {reb}")
    assert reb[0] > reb[1]
    for i in ids[1:]:
        assert_same(b.download(i), ref, f"scene {i}")
    assert_same(ref, alone(ctx, base, GDEM_FP64, nsteps)[0], "alone")
    b.close()


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_served_results_equal_alone(ctx, precision):
    # a short Poisson-like stream through the continuous-batching server: every
    # request's result equals the same scene stepped alone
    from gpudem_amd import serve
    rng = np.random.default_rng(3)
    stream, t = [], 0.0
    for i in range(6):
        t += float(rng.exponential(0.05))
        n = int(rng.integers(200, 3000))
        nsteps = int(rng.integers(200, 900))
        stream.append((t, mk(40 + i, n, r_lo=0.8, r_hi=1.2, phi=0.2, mu=float(rng.uniform(0.2, 0.8))), nsteps))
    say(f"served stream of {len(stream)} requests, precision {precision}")
    srv = serve.Server(ctx, precision, hbm_budget=64 << 20, max_scenes=4, chunk=64, graph_steps=32, **CAPS)
    done = srv.run(stream)
    srv.close()
    assert len(done) == len(stream)
    summ = serve.summarize(done, srv.makespan)
    say(f"  p50 {summ['p50_s']:.3f} s p99 {summ['p99_s']:.3f} s, {summ['value']:.3e} particle-steps/s")
    for req in done:
        assert req.stats["errors"] == 0
        assert req.stats["steps_done"] == req.nsteps
        assert_same(req.result, alone(ctx, req.spec, precision, req.nsteps)[0], f"request {req.rid}")


def test_autotuner_returns_a_timed_candidate(ctx):
    from gpudem_amd import autotune
    specs = [mk(50 + i, 500 + 300 * i) for i in range(4)]
    best, table = autotune.tune_batch(ctx, specs, GDEM_MIXED, sample=4, warm=20, steps=40,
                                      candidates=[(0.1, 1.0, 20), (0.3, 1.25, 40)], log=say)
    assert len(table) == 2 and all(t[1] > 0 for t in table)
    assert best in [t[0] for t in table]


def test_oversized_request_is_rejected(ctx):
    # a request larger than the whole batch can never be admitted: the batch
    # answers GDEM_E_OOM (not the retry-later GDEM_E_BUSY) and the server fails
    # that request and serves the rest
    from gpudem_amd import GdemError, serve
    small = mk(61, 600)
    big = mk(62, 9000)
    b = Batch(ctx, precision=GDEM_MIXED, max_scenes=4, particle_capacity=4096, **CAPS)
    with pytest.raises(GdemError) as err:
        admit(b, big, 100)
    assert err.value.code == -3
    b.close()
    srv = serve.Server(ctx, GDEM_MIXED, hbm_budget=24 << 20, max_scenes=4, chunk=64, graph_steps=32, **CAPS)
    cap = srv.batch.info()["free_particles"]
    big = mk(63, cap + 500)
    say(f"server of {cap} particle slots with a {big.n}-sphere request")
    done = srv.run([(0.0, big, 100), (0.01, small, 200)])
    srv.close()
    by = {r.rid: r for r in done}
    assert len(done) == 2
    assert by[0].error is not None and by[0].result is None
    assert by[1].error is None and by[1].stats["steps_done"] == 200
    assert serve.summarize(done, 1.0)["failed"] == 1


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_preempted_request_resumes_bitwise(ctx, precision):
    # One scene slot: a long request runs, a short one arrives and waits, so
    # the server evicts the long one with a snapshot of its state and contact
    # histories, serves the short one, then resumes the long one from the
    # snapshot.  Both end exactly as when stepped alone.
    from gpudem_amd import serve
    long_req, short_req = mk(81, 2500, mu=0.6, cor=0.4), mk(82, 800)
    stream = [(0.0, long_req, 3000), (0.02, short_req, 400)]
    say(f"preemption, precision {precision}")
    srv = serve.Server(ctx, precision, hbm_budget=64 << 20, max_scenes=1, chunk=100, graph_steps=50,
                       preempt_after=0.0, preempt_min_steps=100, **CAPS)
    done = srv.run(stream)
    srv.close()
    by = {r.rid: r for r in done}
    assert len(done) == 2 and by[0].preemptions >= 1 and by[1].preemptions == 0
    assert by[1].done < by[0].done
    for rid, (_, spec, nsteps) in enumerate(stream):
        req = by[rid]
        assert req.error is None and req.stats["errors"] == 0 and req.stats["steps_done"] == nsteps
        assert_same(req.result, alone(ctx, spec, precision, nsteps)[0], f"request {rid}")
