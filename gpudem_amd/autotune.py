"""Autotuner of the batched engine's free parameters.

The lists are supersets of the contacts whenever refresh <= skin / 2 and
cell >= 2 r_max + skin, so the Verlet skin, the cell size and the number of
steps per captured hipGraph change only the speed, never a trajectory (the
GPU tests check this bit for bit).  The tuner times each candidate on a
sample of the real workload -- a few warm-up steps, then a timed window --
and keeps the fastest.  The window starts after 1,000 steps by default: the
first few hundred steps of freshly admitted scenes (no contacts yet) favour
settings that lose later.  The workgroup size is fixed at 256 (the step
kernel's LDS history columns and the XCD tile mapping are laid out for it).
"""
import itertools
import json
import math
import os
import time

import numpy as np

from . import engine
from .engine import Batch

SKIN_FACTORS = (0.05, 0.1, 0.2, 0.4)
CELL_FACTORS = (1.0, 1.25)
GRAPH_STEPS = (50, 200)
DEFAULT = dict(skin_factor=0.1, cell_factor=1.0, graph_steps=50)

# Choices are cached per (GPU architecture, scene-size bucket, solid-fraction
# bucket): the tuned settings depend on how many particles a scene holds
# (launch overhead against work per step) and how dense it is (list lengths,
# rebuild rate).  GDEM_TUNE_CACHE names the cache file; by default it sits
# next to this module.
CACHE = os.environ.get("GDEM_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json"))


def solid_fraction(spec):
    """Sphere volume over the volume of the mesh's bounding box."""
    v = spec.verts.reshape(-1, 3)
    box = float(np.prod(v.max(axis=0) - v.min(axis=0)))
    return float(np.sum(4.0 / 3.0 * math.pi * spec.r ** 3)) / box


def cache_key(specs, arch=None):
    """(arch, N bucket, phi bucket) of a set of scenes: N bucket is log2 of the
    median sphere count rounded, phi bucket the median solid fraction in
    steps of 0.05."""
    arch = arch or engine.device_arch()
    n = float(np.median([s.n for s in specs]))
    phi = float(np.median([solid_fraction(s) for s in specs]))
    return f"{arch}|n2^{int(round(math.log2(max(n, 1.0))))}|phi{int(round(phi / 0.05)) * 5:02d}"


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def cached_choice(specs, arch=None):
    """The cached choice for these scenes' bucket, or None."""
    e = _load().get(cache_key(specs, arch))
    return None if e is None else {k: e[k] for k in DEFAULT}


def store_choice(specs, best, rate, arch=None):
    """Records a tuned choice for these scenes' bucket."""
    table = _load()
    table[cache_key(specs, arch)] = dict(best, particle_steps_per_s=rate)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def with_tuning(spec, skin_factor, cell_factor):
    """The scene's params with skin = f r_max, refresh = skin/2 and
    cell = cell_factor (2 r_max + skin)."""
    rmax = float(spec.r.max())
    skin = skin_factor * rmax
    return dict(spec.params, skin=skin, refresh=0.5 * skin, cell=cell_factor * (2 * rmax + skin))


def time_batch(ctx, specs, precision, skin_factor, cell_factor, graph_steps, warm, steps, caps):
    # whole graphs only, and the first one (captured and instantiated on the
    # host) inside the warm-up
    g = max(1, graph_steps)
    warm = max(warm, g) // g * g + (g if max(warm, g) % g else 0)
    steps = max(steps, g) // g * g
    total = sum(s.n for s in specs)
    b = Batch(ctx, precision=precision, max_scenes=len(specs), particle_capacity=total + 64 * len(specs),
              graph_steps=graph_steps, **caps)
    try:
        for s in specs:
            m, k, e = engine.material(s.r, s.density, s.young, s.cor)
            sid = b.admit(s.x, s.r, m, k, e, s.verts, with_tuning(s, skin_factor, cell_factor), warm + steps,
                          v=s.v)
            if sid is None:
                raise RuntimeError("tuning batch does not fit")
        b.step(warm)
        ctx.sync()
        b.step(steps)
        ms = b.timing()
    finally:
        b.close()
    return ms / steps, total * steps / (ms / 1e3)


def tune_batch(ctx, specs, precision, sample=32, warm=1000, steps=1000, caps=None, candidates=None, log=None,
               store=True):
    """Times every (skin factor, cell factor, graph steps) candidate on every
    k-th scene of `specs` (about `sample` scenes).  Returns (best, table) where
    best = dict(skin_factor, cell_factor, graph_steps) and table lists
    (candidate, ms_per_step, particle_steps_per_s)."""
    caps = caps or dict(nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)
    stride = max(1, len(specs) // sample)
    probe = specs[::stride][:sample]
    cands = candidates or list(itertools.product(SKIN_FACTORS, CELL_FACTORS, GRAPH_STEPS))
    table = []
    t0 = time.perf_counter()
    for sf, cf, g in cands:
        ms, rate = time_batch(ctx, probe, precision, sf, cf, g, warm, steps, caps)
        table.append((dict(skin_factor=sf, cell_factor=cf, graph_steps=g), ms, rate))
        if log:
            log(f"# tune skin {sf} cell {cf} graph {g}: {ms:.3f} ms/step, {rate:.3e} particle-steps/s")
    best_row = max(table, key=lambda t: t[2])
    best = best_row[0]
    if log:
        log(f"# tuned in {time.perf_counter() - t0:.1f} s: {best}")
    if store:
        try:
            store_choice(specs, best, best_row[2])
        except OSError:
            pass  # a read-only install keeps working without the cache
    return best, table
