"""Throughput of the contact step on the 1,000,000-sphere helical chute.

Workload: 1,000,000 spheres (r ~ U[0.9,1.1] x 0.5 mm) on a 4-turn helical
chute meshed as 20,000 triangles, FP32 forces with FP64 integration, full
tangential history, dt = 2e-6 s.  Wall candidates come from a static per-cell
triangle table built once from the mesh (it yields the same candidate lists as
the BVH traversal, which stays selectable with --wall-broadphase 1).

Phases, in order:
  1. settle: --settle DEM steps (10,000 by default, whatever --warmup says)
     take the spheres from their placement lattice into a flowing, in-contact
     bed on the ramp; its state and contact histories are kept (snapshot);
  2. warm-up: W bench steps, untimed;
  3. window: K bench steps timed with HIP events on the engine's stream.
A bench step is a block of --block DEM steps (250 by default, about 1.6
rebuild intervals of the settled bed), so even a short window contains list
rebuilds and the contact pass of a settled bed; `value` counts DEM
particle-steps.  The force kernel of every 10th DEM step is bracketed by its
own event pair.  The run exits with status 3 (after printing its line) if
the window has no contacts, no rebuild, or an error bit.

`roofline` prices the force kernel's algorithmic bytes per launch -- 192 +
4 (n_nb + n_w) + 32 c bytes per active particle for the mixed form, 212 +
4 (n_nb + n_w) + 56 c for FP64, with the list and contact counts measured at
the end of the window -- against the 8 TB/s HBM peak, and reports a plain
device copy's bandwidth beside it; `traffic` is the HBM bytes per launch
that rocprofv3 FETCH_SIZE / WRITE_SIZE passes of this same command measured
(profiles/k_step_chute1M_traffic.json, scaled to the window's active count).
`cpu_baseline` times the C++/OpenMP oracle (a restatement of the reference's
cpu_dem_verlet_verlet) on the settled state, in a child process pinned with
OMP_PROC_BIND=close and one thread per physical core of the CPUs this
process may use.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_step_chute1M_traffic.json")
METRIC = "particle-steps/s at 1M spheres (% HBM roofline); batched p50/p99 scene latency"
# where the measured configuration departs from the stated workloads
DEVIATIONS = [
    "chute wall broadphase: a static per-cell triangle table built once from the mesh (same candidate "
    "lists as the LBVH traversal, which --wall-broadphase 1 selects)",
    "chute spheres start pre-placed along the ramp and are settled for --settle steps before the "
    "warm-up and the timed window",
    "particles are re-sorted at every list rebuild, not every 50 steps (the internal order never changes "
    "a result)",
    "batched step: five launches per step (four gated rebuild phases and one fused contact kernel) "
    "replayed from hipGraphs, not a persistent kernel",
]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(n_active):
    """HBM bytes per force-kernel launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (calibrated on a copy of known size), scaled to n_active."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    per = t["fetch_bytes_per_particle_step"] + t["write_bytes_per_particle_step"]
    return per * n_active, t


def algorithmic_bytes(mixed, nb, nw, c):
    base, per_contact = (192.0, 32.0) if mixed else (212.0, 56.0)
    return base + 4.0 * (nb + nw) + per_contact * c


def chute_scene(ctx, spec, precision, snap=None, chain=False):
    """The chute scene (direct launches; chain: the list rebuild as one kernel
    per phase, for per-phase profiles) from its placement or a snapshot."""
    from gpudem_amd import Scene, engine
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision, graph_steps=-1 if chain else 0)
    if snap is None:
        return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    sc = Scene(ctx, snap["x"], spec.r, m, k, e, spec.verts, params, v=snap["v"], angv=snap["angv"],
               active=snap["active"])
    sc.set_history(snap)
    return sc


def timed_window(sc, warmup_steps, steps, every=10):
    """Warm-up, then `steps` DEM steps timed with HIP events (force kernel of
    every `every`-th step bracketed); statistics before and after."""
    if warmup_steps:
        sc.step(warmup_steps)
    st0 = sc.status()
    sc.set_timing(every)
    t0 = time.perf_counter()
    sc.step(steps)
    total_ms, force_ms, launches = sc.timing()
    wall = time.perf_counter() - t0
    sc.set_timing(0)
    st1 = sc.status(check=False)
    return dict(stats0=st0, stats1=st1, total_ms=total_ms, force_ms=force_ms, launches=launches, wall_s=wall)


def summarize(run, steps, mixed):
    """Throughput and the force kernel's roofline over a timed window.  The
    algorithmic bytes per particle-step are the mean of those at the start and
    at the end of the window (the bed keeps densifying as it flows)."""
    s0, s1 = run["stats0"], run["stats1"]
    nact = s1["n_active"]
    force_avg_s = run["force_ms"] / 1e3 / max(1, run["launches"])
    B = 0.5 * (algorithmic_bytes(mixed, s0["mean_nb"], s0["mean_wall"], s0["mean_contacts"]) +
               algorithmic_bytes(mixed, s1["mean_nb"], s1["mean_wall"], s1["mean_contacts"]))
    achieved = B * nact / force_avg_s / 1e9 if force_avg_s > 0 else 0.0
    return dict(value=nact * steps / (run["total_ms"] / 1e3), ms_per_dem_step=run["total_ms"] / steps,
                force_avg_s=force_avg_s, B=B, achieved=achieved, nact=nact,
                rebuilds=s1["rebuilds"] - run["stats0"]["rebuilds"], errors=s1["errors"])


# ---- CPU baseline (child process) -------------------------------------------


def host_cpu():
    """CPU model and the physical cores among the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0))
    cores, sockets = set(), set()
    for c in allowed:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "core_id") as f:
                core = int(f.read())
            with open(base + "physical_package_id") as f:
                pkg = int(f.read())
        except (OSError, ValueError):
            core, pkg = c, 0
        cores.add((pkg, core))
        sockets.add(pkg)
    phys = max(1, len(cores))
    return dict(model=model, logical_cpus_allowed=len(allowed), physical_cores_allowed=phys,
                sockets=len(sockets), smt_threads_per_core=round(len(allowed) / phys, 2),
                machine_logical_cpus=os.cpu_count())


def cpu_leg(path, steps, warm):
    """Child process: the oracle on the saved state; prints one JSON line."""
    import numpy as np

    import oracle
    from gpudem_amd import scenes
    z = np.load(path)
    spec = scenes.helical_chute(n=int(z["n"]))
    orc = oracle.Sim(z["x"], spec.r, oracle.Mesh(spec.verts), spec.oracle_params(), v=z["v"], angv=z["angv"],
                     active=z["active"])
    orc.step(warm, gpu_order=False)
    t0 = time.perf_counter()
    orc.step(steps, gpu_order=False)
    dt = time.perf_counter() - t0
    nact = int(np.sum(orc.get("isActive") == 1))
    print(json.dumps(dict(value=nact * steps / dt, seconds=dt, threads=oracle.get_threads(), n_active=nact)))


def cpu_baseline(snap, n, steps, warm=2):
    import numpy as np
    info = host_cpu()
    threads = info["physical_cores_allowed"]
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit():
        threads = min(threads, int(cap))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "state.npz")
        np.savez(path, n=n, x=snap["x"], v=snap["v"], angv=snap["angv"], active=snap["active"])
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", path, "--cpu-steps",
                              str(steps), "--cpu-warm", str(warm)], env=env, capture_output=True, text=True,
                             check=True)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    return res, info, env


# ---- batched and served legs ------------------------------------------------

CAPS = dict(nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)


def _batch_roofline(value, mixed, scenes_stats):
    """Whole-step algorithmic bytes per particle-step (force pass only), weighted
    by each scene's particle-steps, and the HBM rate it implies at `value`."""
    w = sum(s["weight"] for s in scenes_stats)
    nb = sum(s["mean_nb"] * s["weight"] for s in scenes_stats) / w
    nw = sum(s["mean_wall"] * s["weight"] for s in scenes_stats) / w
    c = sum(s["mean_contacts"] * s["weight"] for s in scenes_stats) / w
    B = algorithmic_bytes(mixed, nb, nw, c)
    ach = value * B / 1e9
    return dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                bytes_per_particle_step=B, basis="whole batch step over all kernels"), \
        dict(mean_nb=nb, mean_wall=nw, mean_contacts=c)


def run_batched(ctx, specs, precision, tuned, steps):
    """All scenes admitted together and stepped `steps` timed batch steps (one
    launch per phase, hipGraph-captured).  Every scene ends at the same batch
    step, so its completion latency is the batch's run time."""
    from gpudem_amd import GDEM_MIXED, Batch, autotune, engine
    total = sum(s.n for s in specs)
    b = Batch(ctx, precision=precision, max_scenes=len(specs), particle_capacity=total + 64 * len(specs),
              graph_steps=tuned["graph_steps"], **CAPS)
    ids = []
    # one extra graph's worth of steps runs first, untimed: it captures and
    # instantiates the hipGraph; the timed window is `steps` of every scene
    lead = max(1, tuned["graph_steps"])
    t_a = time.perf_counter()
    for s in specs:
        m, k, e = engine.material(s.r, s.density, s.young, s.cor)
        ids.append(b.admit(s.x, s.r, m, k, e, s.verts,
                           autotune.with_tuning(s, tuned["skin_factor"], tuned["cell_factor"]), lead + steps,
                           v=s.v))
    ctx.sync()
    admit_s = time.perf_counter() - t_a
    b.step(lead)
    ctx.sync()
    t0 = time.perf_counter()
    b.step(steps)
    ms = b.timing()
    wall = time.perf_counter() - t0
    errors, rebuilds = 0, 0
    sst = []
    for sid, s in zip(ids, specs):
        st = b.status(sid, check=False)
        errors |= st["errors"]
        rebuilds += st["rebuilds"]
        sst.append(dict(st, weight=s.n * steps))
    info = b.info()
    b.close()
    value = total * steps / (ms / 1e3)
    roof, lists = _batch_roofline(value, precision == GDEM_MIXED, sst)
    return dict(value=value, ms_per_step=ms / steps, scenes=len(specs), particles=total, steps=steps,
                lead_steps=lead, scene_latency_p50_s=wall, scene_latency_p99_s=wall, admit_s=admit_s,
                errors=errors, rebuilds_per_scene=rebuilds / len(specs), list_stats=lists, roofline=roof,
                batch_bytes=info["bytes"])


def run_plain(ctx, specs, precision, lead, steps, graph_steps):
    """The single-scene path on the same scenes, one after another: each scene
    runs `lead` untimed steps, then `steps` steps replayed from hipGraphs of
    `graph_steps` steps, timed on the host clock around a synchronised run."""
    from gpudem_amd import Scene, engine
    work, secs = 0.0, 0.0
    for s in specs:
        m, k, e = engine.material(s.r, s.density, s.young, s.cor)
        sc = Scene(ctx, s.x, s.r, m, k, e, s.verts,
                   dict(s.params, precision=precision, graph_steps=graph_steps, **CAPS), v=s.v)
        sc.step(lead)
        ctx.sync()
        t0 = time.perf_counter()
        sc.step(steps)
        ctx.sync()
        secs += time.perf_counter() - t0
        work += s.n * steps
        sc.close()
    return dict(value=work / secs, scenes=len(specs), lead_steps=lead, steps=steps, graph_steps=graph_steps,
                seconds=secs)


def run_served(ctx, precision, nreq, budget, seed=5):
    from gpudem_amd import GDEM_MIXED, scenes, serve
    stream = scenes.served_stream(seed=seed, nreq=nreq)
    srv = serve.Server(ctx, precision, hbm_budget=budget, keep_results=False, **CAPS)
    done = srv.run(stream)
    out = serve.summarize(done, srv.makespan)
    srv.close()
    sst = [dict(r.stats, weight=r.spec.n * r.nsteps) for r in done if r.error is None]
    out["roofline"], out["list_stats"] = _batch_roofline(out["value"], precision == GDEM_MIXED, sst)
    out["n_range"] = [min(s.n for _, s, _ in stream), max(s.n for _, s, _ in stream)]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed bench steps (blocks of --block DEM steps)")
    ap.add_argument("--warmup", type=int, default=4, help="untimed bench steps after the settle phase")
    ap.add_argument("--block", type=int, default=250, help="DEM steps per bench step")
    ap.add_argument("--settle", type=int, default=10000, help="DEM steps that settle the bed first")
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--cpu-warm", type=int, default=2)
    ap.add_argument("--cpu-leg", help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--particles", type=int, default=1000000)
    ap.add_argument("--skin-factor", type=float, default=None,
                    help="chute Verlet skin in units of r_max (refresh = skin/2, cell = 2 r_max + skin)")
    ap.add_argument("--wall-broadphase", type=int, default=0,
                    help="0: per-cell table when it fits, 1: BVH traversal, 2: per-cell table")
    ap.add_argument("--batch-steps", type=int, default=5000, help="batch steps of the 256-scene mix")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=2000,
                    help="timed steps per scene of the plain single-scene comparison")
    ap.add_argument("--plain-sample", type=int, default=32, help="scenes of the mix timed on the plain path")
    ap.add_argument("--plain-graph-steps", type=int, default=50,
                    help="steps per hipGraph of the plain single-scene comparison (0: direct launches)")
    ap.add_argument("--served-requests", type=int, default=60)
    ap.add_argument("--served-budget-gb", type=float, default=8.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-FP64 error report")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (cached or default settings)")
    ap.add_argument("--batch-graph-steps", type=int, default=None,
                    help="override the tuned steps per hipGraph of the batched run (0: direct launches)")
    ap.add_argument("--save-state", help="write the chute's state and contact histories at the middle of "
                                         "the timed window to this .npz (for profiling runs)")
    ap.add_argument("--load-state", help="start the chute from this .npz (written by --save-state) instead "
                                         "of settling it")
    ap.add_argument("--chain-rebuild", action="store_true",
                    help="chute: list rebuild as one kernel per phase (per-phase profiles)")
    ap.add_argument("--no-batch-fp64", action="store_true", help="batched leg in mixed precision only")
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg(args.cpu_leg, args.cpu_steps, args.cpu_warm)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine steps one "
              "scene per device and this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    from gpudem_amd import GDEM_FP64, GDEM_MIXED, Context, scenes
    spec = scenes.helical_chute(n=args.particles) if args.skin_factor is None else \
        scenes.helical_chute(n=args.particles, skin_factor=args.skin_factor)
    spec.params["wall_broadphase"] = args.wall_broadphase
    log(f"# workload helical_chute n={spec.n} ntri={len(spec.verts)}")
    ctx = Context(0)
    copy_bytes = 1 << 30
    copy_ms = ctx.copy_probe(copy_bytes, 20)
    copy_bw = 2 * copy_bytes / (copy_ms / 1e3) / 1e9

    import numpy as np
    K, W, blk = args.steps, args.warmup, args.block
    if args.load_state:
        z = np.load(args.load_state)
        snap = {k: z[k] for k in z.files}
        sc = chute_scene(ctx, spec, GDEM_MIXED, snap, chain=args.chain_rebuild)
        log(f"# loaded the settled state from {args.load_state}")
    else:
        # 1. settle (mixed), keeping state and histories
        sc = chute_scene(ctx, spec, GDEM_MIXED, chain=args.chain_rebuild)
        t0 = time.perf_counter()
        sc.step(args.settle)
        ctx.sync()
        snap = sc.snapshot()
        st = sc.status()
        log(f"# settled {args.settle} steps in {time.perf_counter() - t0:.1f} s: contacts/particle "
            f"{st['mean_contacts']:.2f}, candidates {st['mean_nb']:.2f} + {st['mean_wall']:.2f}, "
            f"{st['rebuilds']} rebuilds")
    if args.save_state:
        # the middle of the timed window: a short profiled run from here sees
        # the window's mean list and contact counts
        sc.step(W * blk + K * blk // 2)
        np.savez(args.save_state, **sc.snapshot())
        log(f"# saved the state {W * blk + K * blk // 2} steps after the settle phase to {args.save_state}")
        sc.close()
        ctx.close()
        return
    # 2-3. warm-up and timed window, mixed (the same scene continues)
    mixed = timed_window(sc, W * blk, K * blk)
    sc.close()
    m = summarize(mixed, K * blk, True)
    s1 = mixed["stats1"]
    traffic, tsrc = measured_traffic(m["nact"])
    log(f"# mixed: {m['ms_per_dem_step']:.4f} ms/DEM step, force kernel {m['force_avg_s'] * 1e6:.1f} us "
        f"({mixed['launches']} sampled), B={m['B']:.1f} B/particle-step, rebuilds {m['rebuilds']}, "
        f"contacts {s1['mean_contacts']:.2f}")
    out = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": mixed["total_ms"] / K,
        "ms_per_dem_step": m["ms_per_dem_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 integration",
        "data": "synthetic (seeded helical chute and particle bed)",
        "config": {"workload": "helical_chute_1M_20k_tris_mixed", "particles": spec.n,
                   "triangles": int(len(spec.verts)), "dt": spec.params["dt"],
                   "dem_steps_per_step": blk, "settle_dem_steps": args.settle,
                   "timed_dem_steps": K * blk, "warmup_dem_steps": W * blk,
                   "skin": spec.params["skin"], "cell": spec.params["cell"],
                   "rebuild": "per-phase kernels" if args.chain_rebuild else "grid-barrier kernel",
                   "wall_broadphase": ["table", "bvh", "table"][args.wall_broadphase],
                   "deviations": DEVIATIONS},
        "errors": m["errors"],
        "list_stats": {"mean_nb": s1["mean_nb"], "mean_wall": s1["mean_wall"],
                       "mean_contacts": s1["mean_contacts"], "at": "end of the window",
                       "mean_contacts_at_start": mixed["stats0"]["mean_contacts"],
                       "rebuilds_in_window": m["rebuilds"], "n_active": m["nact"]},
        "roofline": {"bound": "hbm", "achieved": m["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": tsrc["source"] if tsrc else None,
                     "kernel": "k_step<float,float>", "bytes_per_particle_step": m["B"],
                     "avg_launch_us": m["force_avg_s"] * 1e6, "sampled_launches": mixed["launches"],
                     "copy_bw": copy_bw, "frac_of_copy": m["achieved"] / copy_bw},
    }
    ok = m["errors"] == 0 and s1["mean_contacts"] > 0 and m["rebuilds"] > 0
    if not args.no_fp64:
        f = chute_scene(ctx, spec, GDEM_FP64, snap)
        f64 = timed_window(f, W * blk, K * blk)
        f.close()
        d = summarize(f64, K * blk, False)
        ok = ok and d["errors"] == 0 and d["rebuilds"] > 0
        log(f"# fp64: {d['ms_per_dem_step']:.4f} ms/DEM step, force kernel {d['force_avg_s'] * 1e6:.1f} us")
        out["fp64_value"] = d["value"]
        out["fp64_roofline"] = {"achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": d["achieved"] / HBM_PEAK_GBS, "bytes_per_particle_step": d["B"],
                                "avg_launch_us": d["force_avg_s"] * 1e6, "sampled_launches": f64["launches"],
                                "kernel": "k_step<double,double>", "rebuilds_in_window": d["rebuilds"],
                                "mean_contacts": f64["stats1"]["mean_contacts"], "errors": d["errors"]}
    if not args.no_accuracy:
        from gpudem_amd import accuracy
        acc = accuracy.report(ctx)
        log(f"# mixed vs fp64: force {acc['force_rel_err']:.2e}, restitution {acc['restitution_rel_err']:.2e}, "
            f"packing {acc['packing_rel_err']:.2e}")
        out["mixed_precision_error"] = acc
    if not args.no_batch:
        from gpudem_amd import autotune
        mix = scenes.batch_mix(nscenes=args.batch_scenes, steps=args.batch_steps)
        log(f"# batch mix: {len(mix)} scenes, {sum(s.n for s in mix)} spheres")
        if args.no_tune:
            tuned = autotune.cached_choice(mix) or dict(autotune.DEFAULT)
        else:
            tuned, _ = autotune.tune_batch(ctx, mix, GDEM_MIXED, log=log)
        if args.batch_graph_steps is not None:
            tuned = dict(tuned, graph_steps=args.batch_graph_steps)
        bm = run_batched(ctx, mix, GDEM_MIXED, tuned, args.batch_steps)
        log(f"# batched mixed: {bm['value']:.3e} particle-steps/s, {bm['ms_per_step']:.3f} ms/step")
        bf = None
        if not args.no_batch_fp64:
            bf = run_batched(ctx, mix, GDEM_FP64, tuned, args.batch_steps)
            log(f"# batched fp64: {bf['value']:.3e} particle-steps/s")
        sample = mix[::max(1, len(mix) // args.plain_sample)][:args.plain_sample]
        pl = run_plain(ctx, sample, GDEM_FP64, bm["lead_steps"], args.plain_steps,
                       graph_steps=args.plain_graph_steps)
        log(f"# plain fp64 (single-scene path, {args.plain_graph_steps}-step hipGraphs, {len(sample)} scenes one at "
            f"a time): {pl['value']:.3e}")
        ok = ok and bm["errors"] == 0 and (bf is None or bf["errors"] == 0)
        out["batched"] = {"workload": "box_mix_256_logU1e3-5e4", "value": bm["value"],
                          "unit": "particle-steps/s", "dtype": "f32 forces / f64 integration",
                          "fp64_value": bf["value"] if bf else None, "plain_fp64_value": pl["value"],
                          "speedup_vs_plain_fp64": bm["value"] / pl["value"],
                          "scene_latency_p50_s": bm["scene_latency_p50_s"],
                          "scene_latency_p99_s": bm["scene_latency_p99_s"], "tuned": tuned,
                          "mixed": bm, "fp64": bf, "plain_fp64": pl}
    if not args.no_served:
        budget = int(args.served_budget_gb * (1 << 30))
        sv = {}
        for name, prec in (("mixed", GDEM_MIXED), ("fp64", GDEM_FP64)):
            sv[name] = run_served(ctx, prec, args.served_requests, budget)
            ok = ok and sv[name]["errors"] == 0 and sv[name]["failed"] == 0
            log(f"# served {name}: {sv[name]['value']:.3e} particle-steps/s, p50 {sv[name]['p50_s']:.2f} s, "
                f"p99 {sv[name]['p99_s']:.2f} s")
        out["served"] = {"workload": f"poisson20_boxes_logU500-2e5_{args.served_requests}req",
                         "hbm_budget_bytes": budget, "mixed": sv["mixed"], "fp64": sv["fp64"],
                         "p50_s": sv["mixed"]["p50_s"], "p99_s": sv["mixed"]["p99_s"]}
    ctx.close()
    if not args.no_cpu:
        cpu, info, env = cpu_baseline(snap, spec.n, args.cpu_steps, args.cpu_warm)
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "particle-steps/s", "cores": cpu["threads"],
                               "kind": "port",
                               "sample": f"oracle (C++/OpenMP restatement of cpu_dem_verlet_verlet, FP64) on "
                                         f"the settled 1M-sphere chute state after {args.settle} steps, "
                                         f"{args.cpu_steps} steps after {args.cpu_warm} warm-up steps, "
                                         f"{cpu['seconds']:.1f} s",
                               "host": dict(info, omp_num_threads=env["OMP_NUM_THREADS"],
                                            omp_proc_bind=env["OMP_PROC_BIND"], omp_places=env["OMP_PLACES"])}
        log(f"# cpu baseline: {cpu['value']:.3e} particle-steps/s on {cpu['threads']} threads ({info['model']})")
    print(json.dumps(out), flush=True)
    if not ok:
        log("# the timed window had no contacts, no rebuild or an error bit")
        sys.exit(3)


if __name__ == "__main__":
    main()
