"""Throughput of the contact step on the 1,000,000-sphere helical chute.

Workload: 1,000,000 spheres (r ~ U[0.9,1.1] x 0.5 mm) on a 4-turn helical
chute meshed as 20,000 triangles, FP32 forces with FP64 integration, full
tangential history, dt = 2e-6 s.  Wall candidates come from a static per-cell
triangle table built once from the mesh (it yields the same candidate lists as
the BVH traversal, which stays selectable with --wall-broadphase 1).  W warm-up steps let the bed
settle into contact, then K steps are timed with HIP events; the force kernel
(the dominant one) of every n-th timed step is bracketed by its own event pair
on the engine's stream.

Prints one JSON line.  `value` is particle-steps/s over the timed K steps with
all state resident in HBM.  `roofline` prices the force kernel's algorithmic
bytes per launch -- 192 + 4 (n_nb + n_w) + 32 c bytes per active particle for
the mixed form, 212 + 4 (n_nb + n_w) + 56 c for FP64, with the candidate and
contact counts measured at the end of the window -- against the 8 TB/s HBM
peak, and reports a plain device copy's bandwidth beside it; `traffic` is the
measured HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary.  `cpu_baseline`
times the C++/OpenMP oracle (a restatement of the reference's
cpu_dem_verlet_verlet) on the same settled state.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_step_chute1M_traffic.json")


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


def make_scene(ctx, spec, precision, graph_steps, state=None):
    from gpudem_amd import Scene, engine
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    params = dict(spec.params, precision=precision, graph_steps=graph_steps)
    params.setdefault("wall_broadphase", 0)
    if state is None:
        return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    return Scene(ctx, state["x"], spec.r, m, k, e, spec.verts, params, v=state["v"], angv=state["angv"],
                 active=state["active"])


def run_gpu(ctx, spec, precision, warmup, steps, graph_steps=0, every=10, state=None):
    sc = make_scene(ctx, spec, precision, graph_steps, state)
    if warmup:
        sc.step(warmup)
    ctx.sync()
    st0 = sc.status()
    out = dict(stats0=st0)
    if steps:
        sc.set_timing(every)
        t0 = time.perf_counter()
        sc.step(steps)
        ctx.sync()
        out["wall_s"] = time.perf_counter() - t0
        if every > 0:
            out["total_ms"], out["force_ms"], out["launches"] = sc.timing()
        else:  # untimed launches (hipGraph replay allowed): host clock only
            out["total_ms"], out["force_ms"], out["launches"] = out["wall_s"] * 1e3, 0.0, 0
    out["stats1"] = sc.status()
    out["state"] = sc.download()
    sc.close()
    return out


def run_cpu(spec, state, steps, warm):
    import numpy as np

    import oracle
    threads = oracle.get_threads()
    orc = oracle.Sim(state["x"], spec.r, oracle.Mesh(spec.verts), spec.oracle_params(), v=state["v"],
                     angv=state["angv"], active=state["active"])
    orc.step(warm, gpu_order=False)
    t0 = time.perf_counter()
    orc.step(steps, gpu_order=False)
    dt = time.perf_counter() - t0
    nact = int(np.sum(orc.get("isActive") == 1))
    return dict(value=nact * steps / dt, seconds=dt, threads=threads, n_active=nact)


def summarize(run, steps, mixed):
    s1 = run["stats1"]
    nact = s1["n_active"]
    force_avg_s = run["force_ms"] / 1e3 / max(1, run["launches"])
    B = algorithmic_bytes(mixed, s1["mean_nb"], s1["mean_wall"], s1["mean_contacts"])
    achieved = B * nact / force_avg_s / 1e9 if force_avg_s > 0 else 0.0
    return dict(value=nact * steps / (run["total_ms"] / 1e3), ms_per_step=run["total_ms"] / steps,
                force_avg_s=force_avg_s, B=B, achieved=achieved, nact=nact,
                rebuilds=s1["rebuilds"] - run["stats0"]["rebuilds"])


CAPS = dict(nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_batched(ctx, specs, precision, tuned, steps):
    """All scenes admitted together and stepped `steps` timed batch steps (one
    launch per phase, hipGraph-captured).  Every scene ends at the same batch
    step, so its completion latency is the batch's run time."""
    from gpudem_amd import Batch, autotune, engine
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
    errors, rebuilds, contacts = 0, 0, 0.0
    for sid, s in zip(ids, specs):
        st = b.status(sid, check=False)
        errors |= st["errors"]
        rebuilds += st["rebuilds"]
        contacts += st["mean_contacts"] * st["n_active"]
    info = b.info()
    b.close()
    return dict(value=total * steps / (ms / 1e3), ms_per_step=ms / steps, scenes=len(specs), particles=total,
                steps=steps, scene_latency_p50_s=wall, scene_latency_p99_s=wall, admit_s=admit_s,
                errors=errors, rebuilds_per_scene=rebuilds / len(specs), mean_contacts=contacts / total,
                batch_bytes=info["bytes"])


def run_plain(ctx, specs, precision, steps):
    """The single-scene path on the same scenes, one after another."""
    from gpudem_amd import Scene, engine
    work, secs = 0.0, 0.0
    for s in specs:
        m, k, e = engine.material(s.r, s.density, s.young, s.cor)
        sc = Scene(ctx, s.x, s.r, m, k, e, s.verts, dict(s.params, precision=precision, **CAPS), v=s.v)
        ctx.sync()
        t0 = time.perf_counter()
        sc.step(steps)
        ctx.sync()
        secs += time.perf_counter() - t0
        work += s.n * steps
        sc.close()
    return dict(value=work / secs, scenes=len(specs), steps=steps, seconds=secs)


def run_served(ctx, precision, nreq, budget, seed=5):
    from gpudem_amd import scenes, serve
    stream = scenes.served_stream(seed=seed, nreq=nreq)
    srv = serve.Server(ctx, precision, hbm_budget=budget, keep_results=False, **CAPS)
    done = srv.run(stream)
    out = serve.summarize(done, srv.makespan)
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=10000)
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--particles", type=int, default=1000000)
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="> 0: steps captured per hipGraph; 0: direct launches; "
                         "< 0: rebuild as per-phase kernels, direct launches")
    ap.add_argument("--skin-factor", type=float, default=None,
                    help="chute Verlet skin in units of r_max (refresh = skin/2, cell = 2 r_max + skin)")
    ap.add_argument("--wall-broadphase", type=int, default=0,
                    help="0: per-cell table when it fits, 1: BVH traversal, 2: per-cell table")
    ap.add_argument("--sample-every", type=int, default=10,
                    help="bracket the force kernel of every n-th timed step with HIP events "
                         "(0: no events, graph replay allowed, host clock)")
    ap.add_argument("--batch-steps", type=int, default=5000, help="batch steps of the 256-scene mix")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=300,
                    help="steps per scene of the plain single-scene comparison")
    ap.add_argument("--plain-sample", type=int, default=32, help="scenes of the mix timed on the plain path")
    ap.add_argument("--served-requests", type=int, default=60)
    ap.add_argument("--served-budget-gb", type=float, default=8.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-FP64 error report")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (skin 0.1, cell 1.0, graph 100)")
    ap.add_argument("--batch-graph-steps", type=int, default=None,
                    help="override the tuned steps per hipGraph of the batched run (0: direct launches)")
    ap.add_argument("--save-state", help="write the mixed run's final state to this .npz")
    ap.add_argument("--load-state", help="start both runs from this .npz instead of the lattice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine steps one "
              "scene per device and this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from gpudem_amd import GDEM_FP64, GDEM_MIXED, Context, scenes
    spec = scenes.helical_chute(n=args.particles) if args.skin_factor is None else \
        scenes.helical_chute(n=args.particles, skin_factor=args.skin_factor)
    spec.params["wall_broadphase"] = args.wall_broadphase
    print(f"# workload helical_chute n={spec.n} ntri={len(spec.verts)}", file=sys.stderr, flush=True)
    state = None
    if args.load_state:
        z = np.load(args.load_state)
        state = {k: z[k] for k in ("x", "v", "angv", "active")}
    ctx = Context(0)
    copy_bytes = 1 << 30
    copy_ms = ctx.copy_probe(copy_bytes, 20)
    copy_bw = 2 * copy_bytes / (copy_ms / 1e3) / 1e9

    mixed = run_gpu(ctx, spec, GDEM_MIXED, args.warmup, args.steps, args.graph_steps, args.sample_every, state)
    if args.save_state:
        np.savez(args.save_state, **mixed["state"])
    if not args.steps:
        print(f"# settled in {args.warmup} steps", file=sys.stderr, flush=True)
        return
    m = summarize(mixed, args.steps, True)
    s1 = mixed["stats1"]
    traffic, tsrc = measured_traffic(m["nact"])
    print(f"# mixed: {m['ms_per_step']:.4f} ms/step, force kernel {m['force_avg_s'] * 1e6:.1f} us, "
          f"B={m['B']:.1f} B/particle-step, rebuilds {m['rebuilds']}", file=sys.stderr, flush=True)
    out = {
        "metric": "particle-steps/s at 1M spheres (% HBM roofline); batched p50/p99 scene latency",
        "value": m["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 integration",
        "data": "synthetic (seeded helical chute and particle bed)",
        "config": {"workload": "helical_chute_1M_20k_tris_mixed", "particles": spec.n,
                   "triangles": int(len(spec.verts)), "dt": spec.params["dt"],
                   "skin": spec.params["skin"], "cell": spec.params["cell"],
                   "rebuild": "per-phase kernels" if args.graph_steps < 0 else "grid-barrier kernel",
                   "wall_broadphase": ["table", "bvh", "table"][args.wall_broadphase]},
        "list_stats": {"mean_nb": s1["mean_nb"], "mean_wall": s1["mean_wall"],
                       "mean_contacts": s1["mean_contacts"], "rebuilds_in_window": m["rebuilds"],
                       "n_active": m["nact"]},
        "roofline": {"bound": "hbm", "achieved": m["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": tsrc["source"] if tsrc else None,
                     "kernel": "k_step<float,float>", "bytes_per_particle_step": m["B"],
                     "avg_launch_us": m["force_avg_s"] * 1e6, "sampled_launches": mixed["launches"],
                     "copy_bw": copy_bw, "frac_of_copy": m["achieved"] / copy_bw},
    }
    if not args.no_fp64:
        f64 = run_gpu(ctx, spec, GDEM_FP64, args.warmup, args.steps, args.graph_steps, args.sample_every, state)
        d = summarize(f64, args.steps, False)
        out["fp64_value"] = d["value"]
        out["fp64_roofline"] = {"achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": d["achieved"] / HBM_PEAK_GBS, "bytes_per_particle_step": d["B"],
                                "avg_launch_us": d["force_avg_s"] * 1e6, "kernel": "k_step<double,double>"}
    if not args.no_accuracy:
        from gpudem_amd import accuracy
        acc = accuracy.report(ctx)
        log(f"# mixed vs fp64: force {acc['force_rel_err']:.2e}, restitution {acc['restitution_rel_err']:.2e}, "
            f"packing {acc['packing_rel_err']:.2e}")
        out["mixed_precision_error"] = acc
    if not args.no_batch:
        from gpudem_amd import GDEM_FP64 as F64, autotune
        mix = scenes.batch_mix(nscenes=args.batch_scenes, steps=args.batch_steps)
        log(f"# batch mix: {len(mix)} scenes, {sum(s.n for s in mix)} spheres")
        if args.no_tune:
            tuned = dict(skin_factor=0.1, cell_factor=1.0, graph_steps=100)
        else:
            tuned, _ = autotune.tune_batch(ctx, mix, GDEM_MIXED, log=log)
        if args.batch_graph_steps is not None:
            tuned = dict(tuned, graph_steps=args.batch_graph_steps)
        bm = run_batched(ctx, mix, GDEM_MIXED, tuned, args.batch_steps)
        log(f"# batched mixed: {bm['value']:.3e} particle-steps/s, {bm['ms_per_step']:.3f} ms/step")
        bf = run_batched(ctx, mix, F64, tuned, args.batch_steps)
        log(f"# batched fp64: {bf['value']:.3e} particle-steps/s")
        sample = mix[::max(1, len(mix) // args.plain_sample)][:args.plain_sample]
        pl = run_plain(ctx, sample, F64, args.plain_steps)
        log(f"# plain fp64 (single-scene path, {len(sample)} scenes one at a time): {pl['value']:.3e}")
        out["batched"] = {"workload": "box_mix_256_logU1e3-5e4", "value": bm["value"],
                          "unit": "particle-steps/s", "dtype": "f32 forces / f64 integration",
                          "fp64_value": bf["value"], "plain_fp64_value": pl["value"],
                          "speedup_vs_plain_fp64": bm["value"] / pl["value"],
                          "scene_latency_p50_s": bm["scene_latency_p50_s"],
                          "scene_latency_p99_s": bm["scene_latency_p99_s"], "tuned": tuned,
                          "mixed": bm, "fp64": bf, "plain_fp64": pl}
    if not args.no_served:
        budget = int(args.served_budget_gb * (1 << 30))
        sv = {}
        for name, prec in (("mixed", GDEM_MIXED), ("fp64", GDEM_FP64)):
            sv[name] = run_served(ctx, prec, args.served_requests, budget)
            log(f"# served {name}: {sv[name]['value']:.3e} particle-steps/s, p50 {sv[name]['p50_s']:.2f} s, "
                f"p99 {sv[name]['p99_s']:.2f} s")
        out["served"] = {"workload": f"poisson20_boxes_logU500-2e5_{args.served_requests}req",
                         "hbm_budget_bytes": budget, "mixed": sv["mixed"], "fp64": sv["fp64"],
                         "p50_s": sv["mixed"]["p50_s"], "p99_s": sv["mixed"]["p99_s"]}
    ctx.close()
    if not args.no_cpu:
        cpu = run_cpu(spec, mixed["state"], args.cpu_steps, 2)
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "particle-steps/s", "cores": cpu["threads"],
                               "kind": "port",
                               "sample": f"oracle (C++/OpenMP restatement of cpu_dem_verlet_verlet, FP64) on "
                                         f"the settled 1M-sphere chute state, {args.cpu_steps} steps "
                                         f"after 2 warm-up steps, {cpu['seconds']:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
