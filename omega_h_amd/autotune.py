"""Lane-count autotuner for adapt() workloads.

    python -m omega_h_amd.autotune --workload levelset64
    python -m omega_h_amd.autotune --workload batch --boxes 64

Coordinate descent over a context's lane counts (Context.TUNING: the threads
that share one collapse side, one swap product side, one swap candidate and
one edge of the first independent-set round).  A setting is timed as the
median wall time of --reps adapt() calls, each on a fresh copy of the
workload's input mesh made before the clock starts, and must leave the same
mesh as the defaults (the four digests of vertex lists and classification,
coordinates, metric, lengths and qualities); a knob keeps its fastest value
before the next knob is tried, for --sweeps rounds or until a round changes
nothing.  The result (the chosen lane counts, the default and tuned times and
every trial) goes to omega_h_amd/tuning/<workload>.json, which bench.py
applies to its context.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")

CHOICES = {
    "coarsen_lanes": [2, 4, 8, 16, 32],
    "side_lanes": [4, 8, 16, 32, 64],
    "choose_lanes": [2, 4, 8, 16, 32],
    "first_lanes": [2, 4, 8, 16, 32],
}


def tuning_path(workload):
    return os.path.join(TUNING_DIR, workload + ".json")


def load_tuning(workload):
    """The lane counts the tuner chose for `workload`, or None."""
    path = tuning_path(workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["tuning"]


def levelset_input(ctx, n):
    from . import Mesh
    from . import workloads as W
    m = Mesh.box(ctx, 3, n)
    met, nc = W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    return m


def batch_input(ctx, nboxes, seed, max_box_elems):
    from . import Mesh
    from . import workloads as W
    boxes = W.batch_boxes(nboxes, seed, n_range=(8, 48), h_range=(0.005, 0.1),
                          max_nelems=max_box_elems)
    parts = []
    for n, v, _, _ in boxes:
        m = Mesh.box(ctx, 3, n)
        m.set_metric(np.tile(v, m.nents(0)), 6)
        parts.append(m)
    return Mesh.join(ctx, parts)


def time_setting(ctx, base, opts, reps, log=None):
    copies = [base.copy() for _ in range(reps + 1)]
    t0 = time.perf_counter()
    copies[0].adapt(opts)  # warm
    ctx.synchronize()
    if log:
        log(f"    warm-up adapt {(time.perf_counter() - t0) * 1e3:.2f} ms")
    times = []
    for m in copies[1:]:
        t0 = time.perf_counter()
        m.adapt(opts)
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
        if log:
            log(f"    adapt {times[-1] * 1e3:.2f} ms")
    return statistics.median(times), copies[-1].digests()


def tune(ctx, base, opts, reps, sweeps, log=print):
    current = ctx.tuning()
    defaults = dict(current)
    t_default, want = time_setting(ctx, base, opts, reps, log)
    log(f"defaults {defaults}: {t_default * 1e3:.2f} ms")
    best_t = t_default
    trials = [{"tuning": dict(current), "seconds": t_default}]
    for sweep in range(sweeps):
        changed = False
        for name, values in CHOICES.items():
            for v in values:
                if v == current[name]:
                    continue
                ctx.set_tuning(name, v)
                t, got = time_setting(ctx, base, opts, reps, log)
                if got != want:
                    raise RuntimeError(f"{name}={v} changed the adapted mesh")
                trials.append({"tuning": {**current, name: v}, "seconds": t})
                log(f"  sweep {sweep} {name}={v}: {t * 1e3:.2f} ms")
                if t < best_t:
                    best_t = t
                    current[name] = v
                    changed = True
            ctx.set_tuning(name, current[name])
        if not changed:
            break
    # the chosen setting timed again beside the defaults, so that the two
    # figures come from the same moment of the device
    ctx.apply_tuning(defaults)
    t_default, _ = time_setting(ctx, base, opts, reps)
    ctx.apply_tuning(current)
    t_tuned, _ = time_setting(ctx, base, opts, reps)
    return {"tuning": current, "defaults": defaults, "default_seconds": t_default,
            "tuned_seconds": t_tuned, "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=["levelset64", "levelset32", "batch"],
                    default="levelset64")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--boxes", type=int, default=64)
    ap.add_argument("--batch-seed", type=lambda x: int(x, 0), default=0x0A5A17E5EED5EED5)
    ap.add_argument("--max-box-elems", type=float, default=2.5e5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    from . import Context

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = Context(0)
    if args.workload == "batch":
        base = batch_input(ctx, args.boxes, args.batch_seed, args.max_box_elems)
        desc = (f"{args.boxes} joined boxes (batched form), seed {args.batch_seed:#x}, "
                f"metrics clamped to {args.max_box_elems:g} expected elements")
    else:
        n = 64 if args.workload == "levelset64" else 32
        base = levelset_input(ctx, n)
        desc = f"{n}^3 level-set box"
    opts = ctx.default_opts(3)
    opts.verbosity = 0
    log(f"tuning on {desc}: {base.nents(3)} tets")
    res = tune(ctx, base, opts, args.reps, args.sweeps, log)
    res["workload"] = desc
    res["reps"] = args.reps
    out = args.out or tuning_path(args.workload)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    log(f"chose {res['tuning']}: {res['tuned_seconds'] * 1e3:.2f} ms against "
        f"{res['default_seconds'] * 1e3:.2f} ms with the defaults; wrote {out}")
    print(json.dumps({k: v for k, v in res.items() if k != "trials"}))


if __name__ == "__main__":
    main()
