"""The list rebuild's grid barrier has a bounded wait.  With the poll limit
forced to zero from the host (GDEM_GBAR_SPIN=0), the first workgroup that
finds a barrier incomplete cancels it: every kernel still ends normally, no
phase after the cancelled barrier runs, the scene switches to the per-phase
launch chain and rebuilds through it, and gdem_scene_status reports
GDEM_ERR_BARRIER once (GDEM_E_HIP).  The scene's lists and trajectory stay
equal to the oracle's.
"""
import numpy as np
import pytest

import oracle
from gpudem_amd import GDEM_FP64, GDEM_MIXED, Scene, engine, scenes
from test_gpu_parity import assert_lists_equal, assert_state_equal

pytestmark = pytest.mark.gpu

GDEM_E_HIP = -2
GDEM_ERR_BARRIER = 64


def make(ctx, spec, precision):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision))


def test_barrier_timeout_falls_back_to_chain(ctx, monkeypatch):
    spec = scenes.box_scene(3, 40000)  # about 160 workgroups in the rebuild grid
    print("box scene n=40000 FP64, barrier poll limit 0", flush=True)
    monkeypatch.setenv("GDEM_GBAR_SPIN", "0")
    sc = make(ctx, spec, GDEM_FP64)  # its first rebuild runs at creation
    monkeypatch.delenv("GDEM_GBAR_SPIN")
    orc = oracle.Sim(spec.x, spec.r, oracle.Mesh(spec.verts), spec.oracle_params())
    st = sc.status(check=False)
    assert st["errors"] & GDEM_ERR_BARRIER and st["rc"] == GDEM_E_HIP
    st = sc.status(check=False)  # reported once
    assert st["errors"] == 0 and st["rc"] == 0
    # the chain rebuilt the lists at creation
    assert_lists_equal(sc, orc)
    for chunk in range(3):
        sc.step(400)
        orc.step(400)
        assert_state_equal(sc, orc, f"step {(chunk + 1) * 400}")
    assert_lists_equal(sc, orc)
    st = sc.status()
    assert st["errors"] == 0 and st["rebuilds"] == orc.get("rebuilds") and st["rebuilds"] > 0
    sc.close()


def test_barrier_timeout_mixed_then_default_unaffected(ctx, monkeypatch):
    spec = scenes.box_scene(3, 40000)
    print("box scene n=40000 mixed, barrier poll limit 0", flush=True)
    monkeypatch.setenv("GDEM_GBAR_SPIN", "0")
    bad = make(ctx, spec, GDEM_MIXED)
    monkeypatch.delenv("GDEM_GBAR_SPIN")
    bad.step(1500)
    st = bad.status(check=False)
    assert st["errors"] & GDEM_ERR_BARRIER and st["rc"] == GDEM_E_HIP
    print("same scene, default poll limit", flush=True)
    good = make(ctx, spec, GDEM_MIXED)
    good.step(1500)
    st = good.status()
    assert st["errors"] == 0 and st["rebuilds"] > 0
    # the chain and the grid-barrier kernel build the same lists and states
    assert np.array_equal(bad.download()["x"], good.download()["x"])
    db, dg = bad.debug(), good.debug()
    assert np.array_equal(db["nnb"], dg["nnb"]) and np.array_equal(db["nb"], dg["nb"])
    bad.close()
    good.close()
