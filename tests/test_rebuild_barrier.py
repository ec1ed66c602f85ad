"""The list rebuild's grid barrier has a bounded wait.  With the poll limit
forced to zero from the host (GDEM_GBAR_SPIN=0), the workgroups that arrive
early give up at once: the kernels still end normally, the scene reports
GDEM_ERR_BARRIER through gdem_scene_status (GDEM_E_HIP), and a scene created
afterwards with the default limit is unaffected.
"""
import numpy as np
import pytest

from gpudem_amd import GDEM_MIXED, Scene, engine, scenes

pytestmark = pytest.mark.gpu

GDEM_E_HIP = -2
GDEM_ERR_BARRIER = 64


def make(ctx, spec):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=GDEM_MIXED))


def test_barrier_timeout_is_reported(ctx, monkeypatch):
    spec = scenes.box_scene(3, 40000)  # about 160 workgroups in the rebuild grid
    print("box scene n=40000, barrier poll limit 0", flush=True)
    monkeypatch.setenv("GDEM_GBAR_SPIN", "0")
    bad = make(ctx, spec)  # its first rebuild runs at creation
    bad.step(50)
    st = bad.status(check=False)
    assert st["errors"] & GDEM_ERR_BARRIER
    assert st["rc"] == GDEM_E_HIP
    bad.close()
    monkeypatch.delenv("GDEM_GBAR_SPIN")
    print("same scene, default poll limit", flush=True)
    good = make(ctx, spec)
    good.step(50)
    st = good.status()
    assert st["errors"] == 0
    assert np.all(np.isfinite(good.download()["x"]))
    good.close()
