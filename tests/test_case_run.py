"""Whole-case pipeline and snapshot/restore on the GPU.

* a case read from demSettings.json and stepped with its outputs written
  (frames, pid rows, initial positions) gives the same frames as the oracle
  stepped from the same inputs;
* a snapshot of the SAVE_ON state (positions, velocities, spins, activity and
  contact histories) restored into a new scene continues bit for bit;
* a recorded step reports each particle's accelerations (f/m + g, mom/I).
"""
import os
import shutil
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from gpudem_amd import GDEM_FP64, GDEM_MIXED, Scene, engine, run_case, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def young_case(tmp_path):
    """The reference's 100SimpleBox case with a YoungModulus (the oracle's
    material input) in place of the legacy stiffness."""
    d = tmp_path / "case"
    (d / "geometry").mkdir(parents=True)
    js = open(os.path.join(GOLDEN, "cases", "100SimpleBox.json")).read()
    (d / "demSettings.json").write_text(js.replace('"k":1.0e6', '"YoungModulus":5.0e7'))
    shutil.copy(os.path.join(GOLDEN, "simpleBox.stl"), d / "geometry" / "simpleBox.stl")
    return str(d / "demSettings.json")


def test_case_outputs_match_oracle(ctx, tmp_path):
    path = young_case(tmp_path)
    out = str(tmp_path / "gpu")
    nsteps = 1650
    say(f"case 100SimpleBox (YoungModulus), {nsteps} steps with outputs")
    spec, st, frames = run_case.run_case(ctx, path, out, GDEM_FP64, steps=nsteps)
    assert frames == nsteps // st["out_step"]
    orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts, spec.tris_per_file),
                     dict(spec.params, thresh=spec.params["refresh"], density=st["density"], E=st["young"],
                          cor=st["cor"]), v=spec.v)
    ref = str(tmp_path / "orc")
    engine.output_init(ref)
    done = 0
    while done + st["out_step"] <= nsteps:
        orc.step(st["out_step"])
        done += st["out_step"]
        engine.write_frame_bin(ref, done, orc.get("x").reshape(-1, 3), spec.r, orc.get("isActive"))
        name = f"result_{done:012d}.bin"
        assert open(os.path.join(out, name), "rb").read() == open(os.path.join(ref, name), "rb").read(), name
    rows = open(os.path.join(out, "pid1.txt")).read().splitlines()
    assert len(rows) == frames + 1
    x1 = orc.get("x").reshape(-1, 3)[1]
    last = rows[-1].split()
    assert [float(q) for q in last[2:5]] == [float(f"{q:f}") for q in x1]
    assert open(os.path.join(out, "initialPos.csv")).read().count("\n") == spec.n


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_snapshot_restore_continues_bitwise(ctx, precision):
    spec = scenes.box_scene(9, 1000)
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision)
    say(f"snapshot/restore, precision {precision}")
    a = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    a.step(24000)
    snap = a.snapshot()
    assert snap["nhist"].sum() > 1000 and snap["nwhist"].sum() > 50  # a settled bed with live histories
    a.step(3000)
    want = a.download()
    b = Scene(ctx, snap["x"], spec.r, m, k, e, spec.verts, params, v=snap["v"], angv=snap["angv"],
              active=snap["active"])
    b.set_history(snap)
    b.step(3000)
    got = b.download()
    for key in ("x", "v", "angv", "active"):
        assert np.array_equal(got[key], want[key]), key
    # without the histories the continuation differs (the restore matters)
    c = Scene(ctx, snap["x"], spec.r, m, k, e, spec.verts, params, v=snap["v"], angv=snap["angv"],
              active=snap["active"])
    c.step(3000)
    assert not np.array_equal(c.download()["x"], want["x"])
    for s in (a, b, c):
        s.close()


def test_recorded_accelerations(ctx):
    spec = scenes.lattice_drop(verlet=True)
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, spec.params)
    rec = sc.step_record(cap=1)  # first step: no contacts, free fall
    assert np.array_equal(rec["accel"], np.tile(np.array(spec.params["g"]), (spec.n, 1)))
    assert np.all(rec["ang_accel"] == 0.0)
    sc.step(20000)  # the settled lattice: spheres resting on the floor and on each other
    v0 = sc.download()["v"]
    rec = sc.step_record(cap=64)
    v1 = sc.download()["v"]
    dt = spec.params["dt"]
    # v1 = v0 + a dt, in FP64
    assert np.array_equal(v1, v0 + rec["accel"] * dt)
    assert np.any(rec["accel"][:, 2] != spec.params["g"][2])
    sc.close()
