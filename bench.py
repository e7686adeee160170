#!/usr/bin/env python3
"""adapt() throughput on one MI355X.

Workload: build_box unit cube 64^3 cells, 6 tets per cell (1,572,864 tets),
anisotropic metric around the level set phi = z - 0.5 - 0.1 sin(2 pi x)
(normal eigen-length 0.002 + 0.048 (1 - exp(-|phi|/0.02)), tangential 0.05),
one full adapt() with default AdaptOpts (verbosity SILENT), float64 throughout.

A step is one adapt() of a fresh copy of that input mesh (the copies are made
before the timed region; the mesh lives in HBM the whole time).  `value` is
the adapt() throughput in the per-pass sense of the metric: the elements that
entered every pass that rebuilt the mesh, divided by the summed wall time of
those passes (each pass timed on the host clock with the stream drained at
both ends), in millions per second.  `ms_per_step` is the wall time of a
whole adapt() call, including the calls that changed nothing.

The `roofline` object covers the element-quality kernel
(measure_qualities, 3D, 6-component metric) on the input mesh, timed with HIP
events on the library's own stream.  Algorithmic bytes per launch:
24 B per tet (4 vertex indices + one FP64 quality) + 72 B per vertex
(3 FP64 coordinates + 6 FP64 metric components, each read once).  The edge
length kernel is timed the same way (16 B per edge + 72 B per vertex) and
reported on stderr.

The `cpu_baseline` object times the CPU restatement of the same algorithm
(single thread) on the first passes of the same workload, bounded to about
`--cpu-seconds` seconds of pass time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PASS_NAMES = ("refine_by_size", "coarsen_by_size", "swap_edges", "coarsen_slivers")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_rooflines(ctx, mesh, launches):
    """HIP-event timing, on the library's stream, of measure_qualities over
    every tet and measure_edges_metric over every edge of the mesh."""
    import numpy as np
    from omega_h_amd.hip import DeviceArray, Event
    dim = mesh.dim
    nverts = mesh.nents(0)
    ncomps = mesh.metric_ncomps
    coords = DeviceArray(mesh.coords())
    metric = DeviceArray(mesh.metric())
    stream = ctx.stream
    out = {}
    for name, d, fn, want, bytes_per_ent in (
            ("elem_qualities_kernel<3,6>", dim, ctx.measure_qualities, mesh.qualities(), 24),
            ("edge_lengths_kernel<3,6>", 1, ctx.measure_edges_metric, mesh.lengths(), 16)):
        n = mesh.nents(d)
        ev2v = DeviceArray(np.ascontiguousarray(mesh.verts(d), dtype=np.int32))
        res = DeviceArray(dtype=np.float64, n=n)

        def launch():
            fn(dim, ncomps, n, ev2v.ptr, coords.ptr, metric.ptr, res.ptr)

        for _ in range(3):
            launch()
        ctx.synchronize()
        t0, t1 = Event(), Event()
        t0.record(stream)
        for _ in range(launches):
            launch()
        t1.record(stream)
        t1.synchronize()
        ms = t0.elapsed_ms(t1) / launches
        if not np.array_equal(want.view(np.int64), res.numpy().view(np.int64)):
            raise RuntimeError(f"{name} on device pointers differs from the mesh's own values")
        nbytes = bytes_per_ent * n + 72 * nverts
        out[name] = {"kernel": name, "launch_ms": ms, "bytes": nbytes,
                     "achieved": nbytes / (ms * 1e-3) / 1e9}
    return out


def roofline_only(ctx, mesh, launches):
    import numpy as np
    from omega_h_amd.hip import DeviceArray
    for r in kernel_rooflines(ctx, mesh, launches).values():
        log(f"{r['kernel']}: {r['launch_ms'] * 1e3:.1f} us per launch, {r['bytes']} B, "
            f"{r['achieved']:.0f} GB/s")
    n = COPY_BYTES // 8
    a = DeviceArray(np.arange(n, dtype=np.float64))
    b = DeviceArray(dtype=np.float64, n=n)
    for _ in range(5):
        b.copy_from(a)
    log(f"copied {COPY_BYTES} B 5 times")


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` over the whole 64^3 input, from the
    counter summary in profiles/ (FETCH_SIZE + WRITE_SIZE, each scaled by
    what the counter reads for a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "measure_kernels_64cube_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
            return float(parts[2])
    return None


COPY_BYTES = 64 << 20


def constant_metric_request(rng, dim, elems, h_lo, h_hi):
    """A box of about `elems` elements under a constant anisotropic metric:
    a uniformly random rotation and eigen-lengths log-uniform in [h_lo, h_hi]."""
    import numpy as np
    n = max(1, round((elems / 2) ** 0.5)) if dim == 2 else max(1, round((elems / 6) ** (1 / 3)))
    g = rng.normal(size=(dim, dim))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))
    h = np.exp(rng.uniform(np.log(h_lo), np.log(h_hi), dim))
    m = q @ np.diag(1.0 / h ** 2) @ q.T
    if dim == 3:
        v = [m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[1, 2], m[0, 2]]
    else:
        v = [m[0, 0], m[1, 1], m[0, 1]]
    return dim, n, v


def percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p)) if xs else None


def serve_main(args):
    """Poisson stream of adapt requests (30% 2D, 70% 3D boxes, element counts
    log-uniform in [1e3, --max-elems], constant metrics as in the batched
    form) against the request server; p50/p99 latency and throughput."""
    import numpy as np
    import omega_h_amd as A
    ctx = A.Context(0)
    opts = {2: ctx.default_opts(2), 3: ctx.default_opts(3)}
    for o in opts.values():
        o.verbosity = 0  # SILENT, as the workload specifies
    rng = np.random.default_rng(args.seed)
    out = []
    for rate in args.rates:
        server = A.Server(nslots=args.slots, hbm_budget_bytes=args.budget)
        # one request of each dimension first, so that code paths are warm
        for d in (2, 3):
            server.wait(server.submit(*constant_metric_request(rng, d, 1e3, args.h_lo, args.h_hi),
                                      opts[d]))
        t = 0.0
        arrivals = []
        while True:
            t += rng.exponential(1.0 / rate)
            if t > args.duration:
                break
            dim = 2 if rng.random() < 0.3 else 3
            elems = float(np.exp(rng.uniform(np.log(1e3), np.log(args.max_elems))))
            arrivals.append((t, constant_metric_request(rng, dim, elems, args.h_lo, args.h_hi)))
        t0 = time.perf_counter()
        tickets = []
        for at, (dim, n, v) in arrivals:
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(server.submit(dim, n, v, opts[dim]))
        res = [server.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        server.close()
        lat = [r.latency_seconds for r in res]
        elems = sum(r.pass_elems for r in res)
        rec = {"rate": rate, "requests": len(res), "p50_s": percentile(lat, 50),
               "p99_s": percentile(lat, 99), "mean_queue_s": float(np.mean([r.queue_seconds for r in res])),
               "Mtets_per_s": elems / wall / 1e6, "wall_s": wall,
               "elems_out": int(sum(r.nelems_out for r in res))}
        log(f"serve rate {rate}/s: {rec}")
        out.append(rec)
    line = {
        "metric": "p50/p99 serve latency (adapt requests, continuous batching)",
        "value": out[-1]["p99_s"] if out else None, "unit": "s (p99 at the highest rate)",
        "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Poisson adapt requests, 30% 2D / 70% 3D boxes, elements "
                               f"log-uniform in [1e3, {args.max_elems:g}], constant metrics with "
                               f"eigen-lengths log-uniform in [{args.h_lo}, {args.h_hi}], "
                               f"{args.duration:g} s per rate",
                   "slots": args.slots, "hbm_budget_bytes": args.budget},
        "rates": out,
    }
    print(json.dumps(line), flush=True)


def batch_main(args):
    """Many independent boxes of uneven size (cells per side log-uniform in
    [8, 48] in 3D, constant rotated metrics) adapted concurrently by the
    request server, against the same boxes one at a time."""
    import numpy as np
    import omega_h_amd as A
    ctx = A.Context(0)
    opts = ctx.default_opts(3)
    opts.verbosity = 0  # SILENT, as the workload specifies
    rng = np.random.default_rng(args.seed)
    boxes = []
    for _ in range(args.boxes):
        n = int(round(np.exp(rng.uniform(np.log(8), np.log(48)))))
        d, _, v = constant_metric_request(rng, 3, 6 * n ** 3, args.h_lo, args.h_hi)
        boxes.append((n, v))
    out = {}
    for slots in (1, args.slots):
        server = A.Server(nslots=slots, hbm_budget_bytes=args.budget)
        server.wait(server.submit(3, 4, boxes[0][1], opts))
        t0 = time.perf_counter()
        tickets = [server.submit(3, n, v, opts) for n, v in boxes]
        res = [server.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
        server.close()
        elems = sum(r.pass_elems for r in res)
        out[slots] = {"wall_s": wall, "Mtets_per_s": elems / wall / 1e6,
                      "tets_in": int(sum(r.nelems_in for r in res)),
                      "tets_out": int(sum(r.nelems_out for r in res))}
        log(f"batch slots {slots}: {out[slots]}")
    line = {
        "metric": "Mtets adapted/s per pass, batched boxes",
        "value": out[args.slots]["Mtets_per_s"], "unit": "Mtets/s", "n_gpus": 1,
        "higher_is_better": True, "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.boxes} tet boxes, cells per side log-uniform in [8, 48], "
                               f"constant metrics with eigen-lengths log-uniform in "
                               f"[{args.h_lo}, {args.h_hi}]", "slots": args.slots},
        "one_slot": out[1], "concurrent": out[args.slots],
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(n, seconds):
    """Oracle passes, in adapt() order, on the same workload until `seconds`
    of pass time have been spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from omega_h_amd import workloads as W
    m = O.Mesh.box(3, n)
    met, nc = W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    opts = O.default_opts(3)
    spent = 0.0
    nin = 0
    done = []
    order = [0, 1] * 64
    for kind in order:
        n0 = m.nents(3)
        t0 = time.perf_counter()
        changed = getattr(m, PASS_NAMES[kind])(opts)
        dt = time.perf_counter() - t0
        if changed:
            spent += dt
            nin += n0
            done.append(PASS_NAMES[kind])
        if spent >= seconds:
            break
    return {"value": nin / spent / 1e6 if spent > 0 else None, "unit": "Mtets/s",
            "cores": 1, "kind": "port",
            "sample": f"first {len(done)} mesh-changing passes ({', '.join(done)}) of the "
                      f"{n}^3 level-set adapt, {spent:.1f} s of pass time, one thread"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per side of the box")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-launches", type=int, default=20)
    ap.add_argument("--mode", choices=["adapt", "serve", "batch"], default="adapt",
                    help="adapt: the 64^3 level-set adapt (default); serve: Poisson request "
                         "stream; batch: many boxes at once")
    ap.add_argument("--rates", type=float, nargs="+", default=[5.0, 20.0, 50.0])
    ap.add_argument("--duration", type=float, default=20.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget", type=float, default=200e9)
    ap.add_argument("--max-elems", type=float, default=2e5)
    ap.add_argument("--h-lo", type=float, default=0.02)
    ap.add_argument("--h-hi", type=float, default=0.1)
    ap.add_argument("--boxes", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the measure-kernel sweeps and a 64 MiB device copy (the "
                         "reference kernel for calibrating HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the adapt() path has no multi-GPU form)", file=sys.stderr)
        sys.exit(2)

    if args.mode == "serve":
        serve_main(args)
        return
    if args.mode == "batch":
        batch_main(args)
        return

    import omega_h_amd as A
    from omega_h_amd import workloads as W

    ctx = A.Context(0)
    base = A.Mesh.box(ctx, 3, args.n)
    met, nc = W.levelset_metric_3d(base.coords())
    base.set_metric(met, nc)
    opts = ctx.default_opts(3)
    opts.verbosity = 0  # SILENT, as the workload specifies
    nelems0 = base.nents(3)
    log(f"input: {args.n}^3 box, {nelems0} tets, {base.nents(0)} verts")
    if args.roofline_only:
        roofline_only(ctx, base, args.roofline_launches)
        return

    copies = [base.copy() for _ in range(args.warmup + args.steps)]
    for i in range(args.warmup):
        t0 = time.perf_counter()
        res, passes = copies[i].adapt(opts)
        log(f"warmup {i}: {len(passes)} passes, {copies[i].nents(3)} tets, "
            f"{time.perf_counter() - t0:.2f} s")

    ctx.synchronize()
    t0 = time.perf_counter()
    runs = []
    for i in range(args.steps):
        m = copies[args.warmup + i]
        runs.append(m.adapt(opts)[1])
    ctx.synchronize()
    wall = time.perf_counter() - t0

    nin = sum(p[1] for passes in runs for p in passes)
    secs = sum(p[3] for passes in runs for p in passes)
    value = nin / secs / 1e6
    final = copies[-1].nents(3)
    for kind in range(4):
        ps = [p for p in runs[-1] if p[0] == kind]
        if ps:
            log(f"{PASS_NAMES[kind]}: {len(ps)} passes, {sum(p[1] for p in ps)} tets in, "
                f"{sum(p[3] for p in ps):.3f} s")
    log(f"step: {len(runs[-1])} passes, {nelems0} -> {final} tets, {wall / args.steps:.3f} s")

    roofs = kernel_rooflines(ctx, base, args.roofline_launches)
    for r in roofs.values():
        log(f"{r['kernel']}: {r['launch_ms'] * 1e3:.1f} us per launch, {r['bytes']} B, "
            f"{r['achieved']:.0f} GB/s")
    roof = roofs["elem_qualities_kernel<3,6>"]
    line = {
        "metric": "Mtets adapted/s per pass (plain FP64 vs mixed) at 1.57M tets; "
                  "p50/p99 serve latency",
        "value": value,
        "unit": "Mtets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"build_box unit cube {args.n}^3 x 6 tets ({nelems0} tets), "
                               "level-set anisotropic metric, full adapt() FP64, default "
                               "AdaptOpts with verbosity SILENT",
                   "passes_per_step": len(runs[-1]), "final_tets": final},
        "roofline": {"bound": "hbm", "achieved": roof["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": roof["achieved"] / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch(roof["kernel"]) if args.n == 64 else None,
                     "kernel": roof["kernel"],
                     # the counters show HBM traffic within 6% of the algorithmic
                     # bytes; the kernel is bound by its FP64 instruction issue
                     "binding": "fp64 valu",
                     "launch_us": roof["launch_ms"] * 1e3, "bytes": roof["bytes"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(args.n, args.cpu_seconds)
        log(f"cpu baseline: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
