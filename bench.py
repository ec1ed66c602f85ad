"""Throughput of the contact step on the 1,000,000-sphere helical chute.

Workload: 1,000,000 spheres (r ~ U[0.9,1.1] x 0.5 mm) on a 4-turn helical
chute meshed as 20,000 triangles, FP32 forces with FP64 integration, LBVH wall
broadphase, full tangential history, dt = 2e-6 s.  W warm-up steps let the bed
settle into contact, then K steps are timed with HIP events; the force kernel
(the dominant one) is bracketed by its own event pair on the engine's stream.

Prints one JSON line.  `value` is particle-steps/s over the timed K steps
with all state resident in HBM.  `roofline` prices the force kernel's
algorithmic bytes per launch (192 + 4 (n_nb + n_w) + 32 c bytes per active
particle for the mixed form) against the 8 TB/s HBM peak.  `cpu_baseline`
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


def algorithmic_bytes(mixed, nb, nw, c):
    base, per_contact = (192.0, 32.0) if mixed else (212.0, 56.0)
    return base + 4.0 * (nb + nw) + per_contact * c


def run_gpu(spec, precision, warmup, steps):
    from gpudem_amd import Context, Scene, engine
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    ctx = Context(0)
    params = dict(spec.params, precision=precision)
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, params)
    sc.step(warmup)
    ctx.sync()
    st0 = sc.status()
    sc.set_timing(True)
    t0 = time.perf_counter()
    sc.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    total_ms, force_ms, launches = sc.timing()
    st1 = sc.status()
    state = sc.download()
    sc.close()
    ctx.close()
    return dict(total_ms=total_ms, force_ms=force_ms, launches=launches, wall_s=wall, stats0=st0,
                stats1=st1, state=state)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=5000)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--particles", type=int, default=1000000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the engine steps one "
              "scene per device and this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    from gpudem_amd import GDEM_FP64, GDEM_MIXED, scenes
    spec = scenes.helical_chute(n=args.particles)
    print(f"# workload helical_chute n={spec.n} ntri={len(spec.verts)}", file=sys.stderr, flush=True)

    mixed = run_gpu(spec, GDEM_MIXED, args.warmup, args.steps)
    s1 = mixed["stats1"]
    nact = s1["n_active"]
    value = nact * args.steps / (mixed["total_ms"] / 1e3)
    ms_per_step = mixed["total_ms"] / args.steps
    force_avg_s = mixed["force_ms"] / 1e3 / max(1, mixed["launches"])
    B = algorithmic_bytes(True, s1["mean_nb"], s1["mean_wall"], s1["mean_contacts"])
    achieved = B * nact / force_avg_s / 1e9
    print(f"# mixed: {ms_per_step:.4f} ms/step, force kernel {force_avg_s * 1e6:.1f} us, "
          f"B={B:.1f} B/particle-step, rebuilds {s1['rebuilds'] - mixed['stats0']['rebuilds']}",
          file=sys.stderr, flush=True)

    out = {
        "metric": "particle-steps/s at 1M spheres (% HBM roofline); batched p50/p99 scene latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 integration",
        "data": "synthetic (seeded helical chute and particle bed)",
        "config": {"workload": "helical_chute_1M_20k_tris_mixed", "particles": spec.n,
                   "triangles": int(len(spec.verts)), "dt": spec.params["dt"],
                   "skin": spec.params["skin"], "cell": spec.params["cell"]},
        "list_stats": {"mean_nb": s1["mean_nb"], "mean_wall": s1["mean_wall"],
                       "mean_contacts": s1["mean_contacts"],
                       "rebuilds_in_window": s1["rebuilds"] - mixed["stats0"]["rebuilds"],
                       "n_active": nact},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_step<float,float>", "bytes_per_particle_step": B,
                     "avg_launch_us": force_avg_s * 1e6},
    }
    if not args.no_fp64:
        f64 = run_gpu(spec, GDEM_FP64, args.warmup, args.steps)
        t = f64["stats1"]
        out["fp64_value"] = t["n_active"] * args.steps / (f64["total_ms"] / 1e3)
        fa = f64["force_ms"] / 1e3 / max(1, f64["launches"])
        B64 = algorithmic_bytes(False, t["mean_nb"], t["mean_wall"], t["mean_contacts"])
        out["fp64_roofline"] = {"achieved": B64 * t["n_active"] / fa / 1e9, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "bytes_per_particle_step": B64,
                                "avg_launch_us": fa * 1e6}
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
