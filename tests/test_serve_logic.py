"""The serving loop's scheduling on the host (no GPU): admission in arrival
order, failing what can never fit, preemption with snapshots, and no
livelock.  The batch is a host-only stand-in with Batch's interface whose
scenes advance one step per batch step; a scene's 'state' records the steps
it has run, so a resumed scene must carry its count across the eviction."""
import numpy as np

from gpudem_amd import GdemError, scenes, serve


class FakeCtx:
    def sync(self):
        pass


class FakeBatch:
    def __init__(self, max_scenes, capacity):
        self.max_scenes, self.capacity = max_scenes, capacity
        self.step_no = 0
        self.scenes = {}  # sid -> dict(n, start, end, base)
        self.admitted = []

    def admit(self, x, r, m, k, e, verts, params, nsteps, v=None, angv=None, active=None):
        n = len(r)
        if n > self.capacity:
            raise GdemError(-3, "larger than the batch")
        if len(self.scenes) >= self.max_scenes or n + sum(s["n"] for s in self.scenes.values()) > self.capacity:
            return None
        sid = min(set(range(self.max_scenes)) - set(self.scenes))
        base = 0 if v is None else int(v[0, 0])  # steps run before (carried in the snapshot)
        self.scenes[sid] = dict(n=n, start=self.step_no, end=self.step_no + nsteps, base=base, hist=None)
        self.admitted.append((sid, n))
        return sid

    def set_history(self, sid, hist):
        self.scenes[sid]["hist"] = hist

    def step(self, n):
        self.step_no += n

    def info(self):
        return dict(step=self.step_no, live=len(self.scenes), free_particles=0, bytes=0)

    def poll(self):
        return [sid for sid, s in self.scenes.items() if s["end"] <= self.step_no]

    def _done(self, s):
        return s["base"] + min(self.step_no, s["end"]) - s["start"]

    def status(self, sid, check=True):
        s = self.scenes[sid]
        return dict(errors=0, steps_done=min(self.step_no, s["end"]) - s["start"])

    def download(self, sid):
        s = self.scenes[sid]
        v = np.full((s["n"], 3), float(self._done(s)))
        return dict(x=np.zeros((s["n"], 3)), v=v, angv=np.zeros((s["n"], 3)), active=np.ones(s["n"], np.int32))

    def snapshot(self, sid):
        d = self.download(sid)
        d["nhist"] = np.zeros(self.scenes[sid]["n"], np.int32)
        return d

    def evict(self, sid):
        del self.scenes[sid]

    def close(self):
        pass


def spec(n):
    return scenes.SceneSpec("s", np.zeros((n, 3)), np.full(n, 1e-3), np.zeros((0, 9)),
                            dict(dt=1e-5, mu=0.5, g=(0, 0, 0), bbox=[0, 1] * 3, cell=1, skin=0, refresh=0),
                            2500., 1e7, 0.5)


def run(stream, **kw):
    b = FakeBatch(kw.pop("max_scenes", 1), kw.pop("capacity", 10000))
    srv = serve.Server(FakeCtx(), 0, chunk=50, batch=b, **kw)
    done = srv.run(stream)
    return {r.rid: r for r in done}, b


def test_preemption_hands_over_and_resumes():
    by, b = run([(0.0, spec(100), 3000), (0.0001, spec(50), 400)], preempt_after=0.0, preempt_min_steps=100)
    assert by[0].preemptions == 1 and by[1].preemptions == 0
    assert by[1].done <= by[0].done
    # the long request ran all its steps across both residencies
    assert by[0].stats["steps_done"] == 3000 and by[0].result["v"][0, 0] == 3000
    assert by[1].stats["steps_done"] == 400


def test_no_preemption_for_a_longer_head():
    by, b = run([(0.0, spec(100), 500), (0.0001, spec(50), 4000)], preempt_after=0.0, preempt_min_steps=100)
    assert by[0].preemptions == 0 and by[1].preemptions == 0
    assert [s for s, _ in b.admitted] == [0, 0]


def test_many_requests_no_livelock():
    rng = np.random.default_rng(3)
    stream = [(float(t), spec(int(n)), int(st)) for t, n, st in
              zip(np.cumsum(rng.exponential(0.0002, 30)), rng.integers(10, 500, 30), rng.integers(100, 3000, 30))]
    by, b = run(stream, preempt_after=0.0, preempt_min_steps=50, max_scenes=3, capacity=1000)
    assert len(by) == 30
    for rid, (_, sp, nsteps) in enumerate(stream):
        assert by[rid].stats["steps_done"] == nsteps and by[rid].result["v"][0, 0] == nsteps


def test_oversized_request_fails_others_run():
    by, b = run([(0.0, spec(20000), 100), (0.0, spec(10), 100)])
    assert by[0].error is not None and by[1].error is None
