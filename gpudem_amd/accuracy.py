"""Error of the mixed-precision form against the plain FP64 form.

The FP64 form is bit-identical to the CPU restatement of the reference (the
GPU parity tests check trajectories, lists, histories and per-contact forces
bit for bit), so it stands in for the FP64 oracle here.  Three measures:

* force: every contact's force (normal + tangential) in one recorded step
  from the same settled state, aggregate relative error;
* restitution: rebound speed of a sphere hitting a wall head-on over its
  impact speed (the spring-dashpot closed form is e^(pi/3.1415));
* packing fraction: solid fraction of a bed settled under gravity in a box.
"""
import math

import numpy as np

from . import engine, scenes
from .engine import GDEM_FP64, GDEM_MIXED, Scene


def _scene(ctx, spec, precision, state=None):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision)
    if state is None:
        return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params, v=spec.v)
    return Scene(ctx, state["x"], spec.r, m, k, e, spec.verts, params, v=state["v"], angv=state["angv"],
                 active=state["active"])


def force_error(ctx, spec=None, settle=20000):
    """Aggregate relative error of the per-contact forces of one step taken in
    both precisions from the same state (the polydisperse bed after `settle`
    FP64 steps by default).  Returns (rel_err, contacts)."""
    spec = spec or scenes.polydisperse_bed(n=30000)
    s64 = _scene(ctx, spec, GDEM_FP64)
    s64.step(settle)
    st = s64.download()
    s64.close()
    a = _scene(ctx, spec, GDEM_FP64, st)
    b = _scene(ctx, spec, GDEM_MIXED, st)
    ra = a.step_record()
    rb = b.step_record()
    a.close()
    b.close()
    if not np.array_equal(ra["count"], rb["count"]):
        raise RuntimeError("mixed and FP64 steps found different contact sets")
    f64 = np.concatenate([ra["fn"][i, :c] + ra["ft"][i, :c] for i, c in enumerate(ra["count"])])
    f32 = np.concatenate([rb["fn"][i, :c] + rb["ft"][i, :c] for i, c in enumerate(rb["count"])])
    return float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64)), int(ra["count"].sum())


def restitution(ctx, precision, cor=0.5, r=0.005, speed=1.0):
    """Rebound/impact speed of a sphere striking a box floor head-on (no
    gravity)."""
    verts = scenes.box_mesh((-0.05, -0.05, -0.05), (0.05, 0.05, 0.05))
    params = dict(dt=2e-6, mu=0.5, g=(0., 0., 0.), bbox=scenes.bbox_rule(verts, [-0.05, 0.05] * 3),
                  cell=4 * r, skin=r, refresh=r / 2, precision=precision)
    x = np.array([[0.0, 0.0, -0.05 + r + 1e-5]])
    m, k, e = engine.material([r], 2500., 1e7, cor)
    sc = Scene(ctx, x, [r], m, k, e, verts, params, v=np.array([[0., 0., -speed]]))
    sc.step(3000)
    vz = float(sc.download()["v"][0, 2])
    sc.close()
    return vz / speed


def packing_fraction(ctx, precision, n=2000, steps=30000, seed=9):
    """Solid fraction of a box bed settled from a loose lattice: total sphere
    volume over (floor area x bed height), the height taken as the mean top
    (z + r) of the highest 5 % of the spheres above the floor."""
    spec = scenes.box_scene(seed, n)
    sc = _scene(ctx, spec, precision)
    sc.step(steps)
    d = sc.download()
    sc.close()
    lo = spec.verts.reshape(-1, 3).min(axis=0)
    hi = spec.verts.reshape(-1, 3).max(axis=0)
    act = d["active"] == 1
    top = np.sort(d["x"][act, 2] + spec.r[act])
    h = float(top[-max(1, len(top) // 20):].mean() - lo[2])
    vol = float(np.sum(4.0 / 3.0 * math.pi * spec.r[act] ** 3))
    return vol / ((hi[0] - lo[0]) * (hi[1] - lo[1]) * h), float(np.abs(d["v"][act]).max())


def report(ctx):
    """All three measures, mixed against FP64."""
    ferr, ncont = force_error(ctx)
    e64, e32 = restitution(ctx, GDEM_FP64), restitution(ctx, GDEM_MIXED)
    p64, v64 = packing_fraction(ctx, GDEM_FP64)
    p32, v32 = packing_fraction(ctx, GDEM_MIXED)
    return dict(force_rel_err=ferr, force_contacts=ncont, restitution_fp64=e64, restitution_mixed=e32,
                restitution_rel_err=abs(e32 - e64) / abs(e64), restitution_closed_form=0.5 ** (math.pi / 3.1415),
                packing_fp64=p64, packing_mixed=p32, packing_rel_err=abs(p32 - p64) / p64,
                packing_max_speed=max(v64, v32))
