"""Runs a reference case (a demSettings.json) the way main.cu does: build
the scene, step to endTime, and every outStep steps write the frame
(result_%012d.bin), the tracked particles' rows (pid<i>.txt) and, at the
start, initialPos.csv (main.cu:160-169, 258-338; solver_output.cu).

The last step of each output interval is a recorded step, so the pid rows
carry that step's linear and angular accelerations.  The reference's GPU
loop leaves those columns at their initial host values, because
copyFromDevice copies only the SAVE_ON members; its CPU loop writes the
computed accelerations, as here.

    python -m gpudem_amd.run_case cases/100SimpleBox/demSettings.json results --steps 2200
"""
import argparse
import sys

from . import engine, scenes
from .engine import GDEM_FP64, GDEM_MIXED, Context, Scene

NUM_WRITE = 2  # particles tracked in pid<i>.txt (main.cu:163)


def run_case(ctx, settings_path, outdir, precision=GDEM_FP64, steps=None, num_write=NUM_WRITE, graph_steps=0):
    """Returns (scene spec, settings, number of frames written)."""
    spec, st = scenes.from_settings(settings_path)
    m, k, e = spec.material
    sc = Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision,
                                                               graph_steps=graph_steps),
               v=spec.v, tris_per_file=spec.tris_per_file)
    engine.output_init(outdir)
    for pid in range(min(num_write, spec.n)):
        engine.write_pid_header(outdir, pid)
    engine.write_initial_pos(outdir, spec.x, spec.r)
    total = st["steps"] if steps is None else int(steps)
    out_step = st["out_step"]
    done, frames = 0, 0
    while done + out_step <= total:
        if out_step > 1:
            sc.step(out_step - 1)
        rec = sc.step_record(cap=1)
        done += out_step
        # the frame is compacted on the device (16 B per active particle)
        engine.write_frame_compact(outdir, done, *sc.frame())
        if num_write > 0:
            d = sc.download()
            for pid in range(min(num_write, spec.n)):
                engine.write_pid_row(outdir, done, st["dt"], pid, d["x"][pid], d["v"][pid], rec["accel"][pid],
                                     d["angv"][pid], rec["ang_accel"][pid], d["active"][pid])
        frames += 1
    if total > done:
        sc.step(total - done)
    st_run = sc.status(check=False)
    sc.close()
    if st_run["rc"]:
        raise engine.GdemError(st_run["rc"], f"kernel error bits {st_run['errors']:#x}")
    return spec, st, frames


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("settings")
    ap.add_argument("outdir")
    ap.add_argument("--steps", type=int, default=None, help="stop after this many steps (default: endTime/dt)")
    ap.add_argument("--mixed", action="store_true", help="FP32 forces / FP64 integration")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = Context(a.device)
    spec, st, frames = run_case(ctx, a.settings, a.outdir, GDEM_MIXED if a.mixed else GDEM_FP64, a.steps)
    print(f"{spec.n} particles, dt {st['dt']:.6e}, {frames} frames written to {a.outdir}", file=sys.stderr)
    ctx.close()


if __name__ == "__main__":
    main()
