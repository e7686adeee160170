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

`mixed` times the same steps with the mixed-precision form (FP32 cavity
screens, every cavity not rejected by the 1e-3 band re-evaluated in FP64),
checks that it produces the FP64 mesh bit for bit, and reports the FP32
estimates' largest relative errors against FP64 over the input mesh.

The `roofline` object covers the adapt's most expensive kernel, the
coarsen quality sweep (coarsen_quality_lanes, through osh_coarsen_qualities)
over every edge of the input shorter than 1/sqrt(2), both collapse
directions, timed with HIP events on the library's own stream.  It is bound
by FP64 arithmetic: achieved = 222 FP64 operations per cavity element
evaluated (element_quality<3,3>: four 3x3 determinants, the volume, six
metric products and the mean-ratio formula) times the elements evaluated,
over the launch time, against MI355X's 78.6 TFLOP/s.  Divisions and cube
roots execute as instruction sequences, so the FP64 pipe's occupancy is
larger than that fraction: `fp64_issue_frac` divides the kernel's FP64 wave
instructions (counted by rocprofv3, profiles/) by what the pipe issues in the
measured launch time.  `traffic` is the sweep's HBM bytes per launch from the
FETCH_SIZE/WRITE_SIZE counters.  `roofline_measure` keeps the element-quality
measure kernel against HBM (24 B per tet + 72 B per vertex).

`levelset_drivers` compares, once each outside the timed steps, the bare
adapt() with the reference's idiom for a strongly anisotropic target
(src/aniso_test.cpp:15-27: the implied metric becomes the metric, the
level-set metric the target, then while (approach_metric) adapt): passes by
kind, time, final element count, minimum quality and the fraction of edges
with metric length in [1/sqrt(2), sqrt(2)].

The `cpu_baseline` object times the oracle (the CPU restatement, one
thread) on a bounded sample of the same workload's passes: the first coarsen
pass from the input, then swap and sliver passes from the mesh the device
reached after its last coarsen pass (the device's mesh is the oracle's, bit
for bit, at every pass), about `--cpu-seconds` seconds in all.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# FP64 operations (add, sub, mul, div, sqrt, cbrt: one each) of one
# element_quality<3,3> (maxdet_metric + metric_element_quality): 4 x 14 for
# the vertex metrics' determinants, 9 + 15 for the basis and volume, 2 for the
# size, 9 for the other edge vectors, 6 x 20 + 6 + 1 for the mean squared
# metric length, 5 for the mean ratio
FLOPS_PER_QUALITY = 222
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


def coarsen_roofline(ctx, mesh, launches, precision=0):
    """HIP-event timing, on the library's stream, of the coarsen quality sweep
    (osh_coarsen_qualities) over every edge shorter than 1/sqrt(2) of the
    mesh, both collapse directions."""
    import numpy as np
    from omega_h_amd.hip import DeviceArray, Event
    lengths = mesh.lengths()
    cands = np.nonzero(lengths < 2 ** -0.5)[0].astype(np.int32)
    codes = np.full(len(cands), 3, dtype=np.int8)
    ev2v = mesh.verts(1).reshape(-1, 2)
    v2t = np.diff(mesh.up(0, 3)[0])
    e2t = np.diff(mesh.up(1, 3)[0])
    # the elements around the collapsing vertex that survive the collapse:
    # all of its elements except those around the edge
    evals = int((v2t[ev2v[cands, 0]] + v2t[ev2v[cands, 1]] - 2 * e2t[cands]).sum())
    dc, dk = DeviceArray(cands), DeviceArray(codes)
    out = DeviceArray(dtype=np.float64, n=2 * len(cands))

    def launch():
        mesh.coarsen_qualities_ptr(len(cands), dc.ptr, dk.ptr, out.ptr, precision, 0.2)

    for _ in range(2):
        launch()
    ctx.synchronize()
    stream = ctx.stream
    t0, t1 = Event(), Event()
    t0.record(stream)
    for _ in range(launches):
        launch()
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_ms(t1) / launches
    flops = FLOPS_PER_QUALITY * evals
    return {"kernel": "coarsen_quality_lanes<3,3>", "launch_ms": ms, "sides": 2 * len(cands),
            "evaluations": evals, "flops": flops, "tflops": flops / (ms * 1e-3) / 1e12}


def quality_stats(mesh):
    """Minimum element quality and the fraction of edges of metric length in
    [1/sqrt(2), sqrt(2)]."""
    import numpy as np
    q = mesh.qualities()
    lengths = mesh.lengths()
    ok = (lengths >= 2 ** -0.5) & (lengths <= 2 ** 0.5)
    return {"min_quality": float(q.min()), "edges_in_length_goal": float(ok.mean()),
            "max_length": float(lengths.max())}


def levelset_drivers(ctx, base, opts, n):
    """The bare adapt() (as timed) against the reference's idiom for a strongly
    anisotropic target (implied metric + while (approach_metric) adapt)."""
    import numpy as np
    from omega_h_amd import workloads as W
    out = {}
    m = base.copy()
    t0 = time.perf_counter()
    _, passes = m.adapt(opts)
    out["bare_adapt"] = {"adapt_calls": 1, "seconds": time.perf_counter() - t0,
                         "passes": [sum(1 for p in passes if p[0] == k) for k in range(4)],
                         "final_tets": m.nents(3), **quality_stats(m)}
    m = A_mod().Mesh.box(ctx, 3, n)
    target, nc = W.levelset_metric_3d(m.coords())
    t0 = time.perf_counter()
    m.implied_metrics(set_as_metric=True)
    m.add_vertex_tag("target_metric", nc, 1, target)
    calls, kinds = 0, [0, 0, 0, 0]
    nin = 0
    pass_secs = 0.0
    while m.approach_metric(opts):
        _, passes = m.adapt(opts)
        calls += 1
        for p in passes:
            kinds[p[0]] += 1
            nin += p[1]
            pass_secs += p[3]
    secs = time.perf_counter() - t0
    out["implied_then_approach"] = {
        "adapt_calls": calls, "seconds": secs, "passes": kinds, "final_tets": m.nents(3),
        "Mtets_per_s_per_pass": nin / pass_secs / 1e6 if pass_secs > 0 else None,
        **quality_stats(m)}
    for k, v in out.items():
        log(f"driver {k}: {v}")
    return out


def A_mod():
    import omega_h_amd as A
    return A


def roofline_only(ctx, mesh, launches):
    import numpy as np
    from omega_h_amd.hip import DeviceArray
    for r in kernel_rooflines(ctx, mesh, launches).values():
        log(f"{r['kernel']}: {r['launch_ms'] * 1e3:.1f} us per launch, {r['bytes']} B, "
            f"{r['achieved']:.0f} GB/s")
    cq = coarsen_roofline(ctx, mesh, launches)
    log(f"{cq['kernel']}: {cq['launch_ms'] * 1e3:.1f} us per launch, {cq['evaluations']} "
        f"element qualities, {cq['tflops']:.2f} TFLOP/s")
    n = COPY_BYTES // 8
    a = DeviceArray(np.arange(n, dtype=np.float64))
    b = DeviceArray(dtype=np.float64, n=n)
    for _ in range(5):
        b.copy_from(a)
    log(f"copied {COPY_BYTES} B 5 times")


TRAFFIC_SUMMARIES = ("measure_kernels_64cube_hbm_traffic.txt",
                     "coarsen_quality_lanes_64cube_counters.txt")


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` over the whole 64^3 input, from the
    counter summaries in profiles/ (FETCH_SIZE + WRITE_SIZE, each scaled by
    what the counter reads for a device copy of known size), or None."""
    for name in TRAFFIC_SUMMARIES:
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                return float(parts[2])
    return None


def fp64_wave_instructions(kernel):
    """FP64 wave instructions per launch of `kernel` over the 64^3 input, from
    the counter summary in profiles/ (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64),
    or None."""
    path = os.path.join(ROOT, "profiles", "coarsen_quality_lanes_64cube_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "fp64_wave_instructions" and parts[1] == kernel:
            return float(parts[2])
    return None


# FP64 wave instructions MI355X issues per second: one per CU per clock
# (256 CUs at 2.4 GHz; 64 lanes each, the 78.6 TFLOP/s peak with FMAs)
FP64_ISSUE_PER_S = 256 * 2.4e9


COPY_BYTES = 64 << 20


def constant_metric_request(rng, dim, elems, h_lo, h_hi, clamp=None):
    from omega_h_amd import workloads as W
    return W.serve_request(rng, dim, elems, h_lo, h_hi, clamp)


def apply_tuning(ctx, args, workload):
    """The lane counts bench runs with (see --tuning); returns them."""
    from omega_h_amd import autotune
    if args.tuning == "none":
        chosen = None
    elif args.tuning == "auto":
        chosen = autotune.load_tuning(workload) if workload else None
    else:
        with open(args.tuning) as f:
            chosen = json.load(f)["tuning"]
    if chosen:
        ctx.apply_tuning(chosen)
    return {"lanes": ctx.tuning(), "source": "autotuner" if chosen else "defaults"}


def named_ranges(ctx, base, opts):
    """One adapt() of a copy of `base` with the context's timer tree on and
    every range waiting for the stream at its end (so a range is charged the
    device work it enqueued; the run is slower than a timed one).  Prints the
    TOP-DOWN/BOTTOM-UP listings and returns each range name's self time and
    calls (the roots of the bottom-up tree) and the adapt's total."""
    import omega_h_amd as A
    m = base.copy()
    ctx.synchronize()
    ctx.profile_clear()
    ctx.set_profile(A.Context.PROFILE_TREE | A.Context.PROFILE_SYNC)
    try:
        m.adapt(opts)
    finally:
        ctx.set_profile(0)
    log(ctx.profile_report(chop=0.5, percent=False))
    frames = ctx.profile_frames()
    kids = {}
    for i, f in enumerate(frames):
        kids.setdefault(f[1], []).append(i)
    selfs = {}
    for i, (name, _, secs, calls) in enumerate(frames):
        own = max(secs - sum(frames[k][2] for k in kids.get(i, [])), 0.0)
        s, c = selfs.get(name, (0.0, 0))
        selfs[name] = (s + own, c + calls)
    total = sum(f[2] for f in frames if f[1] == -1)
    ctx.profile_clear()
    return {"total_s": total,
            "self_s": {k: round(v[0], 6) for k, v in sorted(selfs.items(), key=lambda kv: -kv[1][0])},
            "calls": {k: v[1] for k, v in selfs.items()}}


def percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p)) if xs else None


def serve_main(args):
    """Poisson stream of adapt requests (30% 2D, 70% 3D boxes, element counts
    log-uniform in [1e3, --max-elems], constant metrics as in the batched
    form, each clamped by the reference's element-count rule to
    --clamp-elems expected elements) against the request server; p50/p99
    latency and throughput."""
    import numpy as np
    import omega_h_amd as A
    from omega_h_amd import workloads as W
    ctx = A.Context(0)
    opts = {2: ctx.default_opts(2), 3: ctx.default_opts(3)}
    for o in opts.values():
        o.verbosity = 0  # SILENT, as the workload specifies
        o.precision = 1 if args.precision == "mixed" else 0
    rng = np.random.default_rng(args.seed)
    out = []
    for rate in args.rates:
        server = A.Server(nslots=args.slots, hbm_budget_bytes=args.budget)
        # one request of each dimension first, so that code paths are warm
        for d in (2, 3):
            server.wait(server.submit(*constant_metric_request(rng, d, 1e3, args.h_lo, args.h_hi,
                                                               args.clamp_elems), opts[d]))
        arrivals = W.serve_stream(rng, rate, args.duration, args.max_elems, args.h_lo,
                                  args.h_hi, args.clamp_elems)
        t0 = time.perf_counter()
        tickets = []
        next_note = 10.0
        for at, (dim, n, v) in arrivals:
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(server.submit(dim, n, v, opts[dim]))
            if at >= next_note:
                log(f"  rate {rate}/s: {len(tickets)} requests submitted by {at:.0f} s")
                next_note += 10.0
        res = []
        for tk in tickets:
            res.append(server.wait(tk))
            if len(res) % 200 == 0:
                log(f"  rate {rate}/s: {len(res)} of {len(tickets)} done at "
                    f"{time.perf_counter() - t0:.0f} s")
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
        "n_gpus": 1, "higher_is_better": False, "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 screens, f64 decisions",
        "data": "synthetic",
        "config": {"precision": args.precision,
                   "workload": f"Poisson adapt requests, 30% 2D / 70% 3D boxes, elements "
                               f"log-uniform in [1e3, {args.max_elems:g}], constant metrics with "
                               f"eigen-lengths log-uniform in [{args.h_lo}, {args.h_hi}] "
                               f"clamped to {args.clamp_elems:g} expected elements by the "
                               f"reference's element-count rule, {args.duration:g} s per rate",
                   "slots": args.slots, "hbm_budget_bytes": args.budget},
        "rates": out,
    }
    print(json.dumps(line), flush=True)


def batch_main(args):
    """The batched form: --boxes independent tet boxes (cells per side n
    log-uniform in [8, 48]; each with a constant metric rotated by a random
    unit quaternion, eigen-lengths log-uniform in [0.005, 0.1]; numpy PCG64
    seeded with --batch-seed), each metric clamped by the reference's
    element-count rule to --max-box-elems expected elements, adapted four
    ways: one after another on one stream; osh_adapt_batch on one stream
    (joined meshes, one launch per kernel for all boxes, finished boxes set
    aside as they appear); and both again spread over --streams contexts
    (one HIP stream and one host thread each, boxes dealt out by size), so
    that the passes of different boxes overlap on the device.  The inputs are
    built before the clock starts.  `value` is the fastest way's throughput:
    the work of adapting every box alone (elements entering each of its
    mesh-changing passes) over the way's wall time, so that all ways are
    priced on the same work (a joined pass also counts the elements of parts
    it leaves alone; that count is kept as `pass_tets`).  Every box must end
    bit for bit as it does adapted alone (four digests), in FP64 and in the
    mixed form."""
    import threading

    import numpy as np
    import omega_h_amd as A
    from omega_h_amd import workloads as W
    n_range = tuple(args.n_range)
    boxes = W.batch_boxes(args.boxes, args.batch_seed, n_range=n_range, h_range=(0.005, 0.1),
                          max_nelems=args.max_box_elems)
    nclamped = sum(1 for b in boxes if b[3] < 1.0)
    expected_raw = sum(b[2] for b in boxes)
    expected = sum(min(b[2], args.max_box_elems) for b in boxes)
    tets_in = sum(6 * b[0] ** 3 for b in boxes)
    log(f"{len(boxes)} boxes, {tets_in} tets in, expected {expected_raw:.3g} tets unclamped, "
        f"{expected:.3g} after clamping {nclamped} boxes")
    nstreams = max(1, args.streams)
    ctxs = [A.Context(0) for _ in range(nstreams)]
    tuning = apply_tuning(ctxs[0], args, "batch")
    for c in ctxs[1:]:
        c.apply_tuning(tuning["lanes"])
    order = sorted(range(len(boxes)), key=lambda i: -boxes[i][0])

    def build(ctx, ids):
        out = []
        for i in ids:
            n, v, _, _ = boxes[i]
            m = A.Mesh.box(ctx, 3, n)
            m.set_metric(np.tile(v, m.nents(0)), 6)
            out.append(m)
        return out

    def run(way, s, precision):
        groups = [order[k::s] for k in range(s)]
        meshes = [build(ctxs[k], groups[k]) for k in range(s)]
        opts = []
        for k in range(s):
            o = ctxs[k].default_opts(3)
            o.verbosity = 0
            o.precision = precision
            opts.append(o)
        for c in ctxs[:s]:
            c.synchronize()
        nin = [0] * s

        def work(k):
            if way == "joined":
                nin[k] = A.adapt_batch(meshes[k], opts[k])[2][1]
            else:
                nin[k] = sum(p[1] for m in meshes[k] for p in m.adapt(opts[k])[1])
            ctxs[k].synchronize()

        t0 = time.perf_counter()
        if s == 1:
            work(0)
        else:
            th = [threading.Thread(target=work, args=(k,)) for k in range(s)]
            for t in th:
                t.start()
            for t in th:
                t.join()
        wall = time.perf_counter() - t0
        by_box = {}
        for k in range(s):
            for i, m in zip(groups[k], meshes[k]):
                by_box[i] = m
        return {"s": wall, "pass_tets": sum(nin),
                "tets_out": sum(m.nents(3) for m in by_box.values())}, by_box

    ways = [("one_at_a_time", "alone", 1), ("joined", "joined", 1),
            (f"one_at_a_time_{nstreams}_streams", "alone", nstreams),
            (f"joined_{nstreams}_streams", "joined", nstreams)]
    results = {}
    reference = None
    same = True
    for precision in (0, 1):
        for name, way, s in ways:
            if s > 1 and nstreams == 1:
                continue
            run(way, s, precision)  # warm
            rec, by_box = run(way, s, precision)
            digests = [by_box[i].digests() for i in range(len(boxes))]
            if reference is None:
                reference = digests
            same = same and digests == reference
            key = name if precision == 0 else name + "_mixed"
            results[key] = rec
            log(f"batch {key}: {rec}")
    work = results["one_at_a_time"]["pass_tets"]
    for rec in results.values():
        rec["Mtets_per_s"] = work / rec["s"] / 1e6
    best = max((k for k in results if not k.endswith("_mixed")),
               key=lambda k: results[k]["Mtets_per_s"])
    line = {
        "metric": "Mtets adapted/s, batched boxes (elements entering each box's own "
                  "mesh-changing passes over the wall time)",
        "value": results[best]["Mtets_per_s"], "unit": "Mtets/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": results[best]["s"] * 1e3,
        "higher_is_better": True, "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(boxes)} tet boxes, cells per side log-uniform in "
                               f"[{n_range[0]}, {n_range[1]}] (seed {args.batch_seed:#x}), "
                               "constant metrics rotated "
                               "by random unit quaternions, eigen-lengths log-uniform in "
                               f"[0.005, 0.1], clamped to {args.max_box_elems:g} expected "
                               "elements per box", "boxes_clamped": nclamped,
                   "expected_tets_unclamped": expected_raw, "expected_tets": expected,
                   "streams": nstreams, "tuning": tuning},
        "tets_in": tets_in,
        "pass_tets_alone": work,
        "best": best,
        "ways": results,
        "speedup_vs_one_at_a_time": results[best]["Mtets_per_s"] /
        results["one_at_a_time"]["Mtets_per_s"],
        "every_way_and_precision_same_meshes": same,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(ctx, base, opts_gpu, n, seconds):
    """The oracle on a bounded sample of the workload's passes: the first
    coarsen pass from the input, then (from the mesh the device reached after
    satisfy_lengths, which is the oracle's own mesh there, bit for bit) swap
    and sliver passes in satisfy_quality's order, until `seconds` of pass
    time have been spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from omega_h_amd import workloads as W
    opts = O.default_opts(3)
    opts.verbosity = 0
    done = []
    spent = 0.0
    nin = 0
    m = O.Mesh.box(3, n)
    met, nc = W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    n0 = m.nents(3)
    t0 = time.perf_counter()
    changed = m.coarsen_by_size(opts)
    dt = time.perf_counter() - t0
    if changed:
        spent += dt
        nin += n0
        done.append("coarsen_by_size")
    # satisfy_lengths on the device, then the same mesh in the oracle
    g = base.copy()
    while True:
        r = g.refine_by_size(opts_gpu)
        c = g.coarsen_by_size(opts_gpu)
        if not (r or c):
            break
    m = O.Mesh.copy_of(g)
    qdesired = opts.min_quality_desired
    while spent < seconds and m.qualities().min() < qdesired:
        for kind in (2, 3):
            n0 = m.nents(3)
            t0 = time.perf_counter()
            changed = getattr(m, PASS_NAMES[kind])(opts)
            dt = time.perf_counter() - t0
            if changed:
                spent += dt
                nin += n0
                done.append(PASS_NAMES[kind])
                break
        else:
            break
    counts = {k: done.count(k) for k in dict.fromkeys(done)}
    return {"value": nin / spent / 1e6 if spent > 0 else None, "unit": "Mtets/s",
            "cores": 1, "kind": "port",
            "sample": f"{len(done)} mesh-changing passes of the {n}^3 level-set adapt "
                      f"({', '.join(f'{v} {k}' for k, v in counts.items())}): the first "
                      f"coarsen pass from the input, then satisfy_quality's passes from "
                      f"the mesh after the last coarsen pass; {spent:.1f} s of pass time, "
                      f"one thread"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per side of the box")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-drivers", action="store_true")
    ap.add_argument("--roofline-launches", type=int, default=20)
    ap.add_argument("--mode", choices=["adapt", "serve", "batch"], default="adapt",
                    help="adapt: the 64^3 level-set adapt (default); serve: Poisson request "
                         "stream; batch: many boxes at once")
    ap.add_argument("--rates", type=float, nargs="+", default=[5.0, 20.0, 50.0])
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="adapt precision of the serving requests")
    ap.add_argument("--duration", type=float, default=20.0)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--budget", type=float, default=200e9)
    ap.add_argument("--max-elems", type=float, default=2e6)
    ap.add_argument("--clamp-elems", type=float, default=2.5e5)
    ap.add_argument("--h-lo", type=float, default=0.005)
    ap.add_argument("--h-hi", type=float, default=0.1)
    ap.add_argument("--boxes", type=int, default=256)
    ap.add_argument("--batch-seed", type=lambda x: int(x, 0), default=0x0A5A17E5EED5EED5)
    ap.add_argument("--max-box-elems", type=float, default=2.5e5)
    ap.add_argument("--n-range", type=int, nargs=2, default=[8, 48],
                    help="cells per side of the batched boxes (log-uniform)")
    ap.add_argument("--streams", type=int, default=8,
                    help="contexts (HIP streams and host threads) of the batched form")
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--tuning", default="auto",
                    help="lane counts: 'auto' (omega_h_amd/tuning/<workload>.json when "
                         "the autotuner wrote one), 'none' (the defaults) or a JSON path")
    ap.add_argument("--ranges", action="store_true",
                    help="after the timed steps, one more adapt() with named ranges (timer "
                         "tree, each range waiting for the stream): its listings go to "
                         "stderr and each range's self time into the JSON line")
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

    import numpy as np

    import omega_h_amd as A
    from omega_h_amd import workloads as W

    ctx = A.Context(0)
    tuning = apply_tuning(ctx, args, "levelset64" if args.n == 64 else None)
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

    def timed(precision):
        o = ctx.default_opts(3)
        o.verbosity = 0  # SILENT, as the workload specifies
        o.precision = precision
        copies = [base.copy() for _ in range(args.warmup + args.steps)]
        for i in range(args.warmup):
            t0 = time.perf_counter()
            copies[i].adapt(o)
            log(f"{'mixed' if precision else 'fp64'} warmup {i}: {copies[i].nents(3)} tets, "
                f"{time.perf_counter() - t0:.2f} s")
        ctx.synchronize()
        t0 = time.perf_counter()
        runs = []
        for i in range(args.steps):
            runs.append(copies[args.warmup + i].adapt(o)[1])
        ctx.synchronize()
        wall = time.perf_counter() - t0
        nin = sum(p[1] for passes in runs for p in passes)
        secs = sum(p[3] for passes in runs for p in passes)
        return nin / secs / 1e6, wall, runs, copies[-1]

    value, wall, runs, last = timed(0)
    final = last.nents(3)
    for kind in range(4):
        ps = [p for p in runs[-1] if p[0] == kind]
        if ps:
            log(f"{PASS_NAMES[kind]}: {len(ps)} passes, {sum(p[1] for p in ps)} tets in, "
                f"{sum(p[3] for p in ps):.3f} s")
    log(f"step: {len(runs[-1])} passes, {nelems0} -> {final} tets, {wall / args.steps:.3f} s")
    mixed = None
    if not args.no_mixed:
        mvalue, mwall, mruns, mlast = timed(1)
        l32, q32, unsure = base.screen_estimates()
        l64, q64 = base.lengths(), base.qualities()
        ok = unsure == 0
        mixed = {
            "value": mvalue, "unit": "Mtets/s", "ms_per_step": mwall / args.steps * 1e3,
            "speedup_vs_fp64": mvalue / value,
            "step_speedup_vs_fp64": wall / mwall,
            "same_mesh_as_fp64": mlast.digests() == last.digests(),
            "same_passes_as_fp64": [p[:3] for p in mruns[-1]] == [p[:3] for p in runs[-1]],
            "screen_band_rel": 1e-3,
            "fp32_estimate_error_on_input": {
                "length_rel_max": float((np.abs(l32 - l64) / np.abs(l64)).max()),
                "quality_rel_max": float((np.abs(q32 - q64)[ok] /
                                          np.maximum(np.abs(q64[ok]), 1e-3)).max()),
                "quality_unsure_frac": float(1.0 - ok.mean()),
                "log_exp": "not on this workload's path (no refinement passes)"},
        }
        log(f"mixed: {mixed}")

    cq = coarsen_roofline(ctx, base, args.roofline_launches)
    cq32 = coarsen_roofline(ctx, base, args.roofline_launches, precision=1)
    log(f"{cq['kernel']}: {cq['launch_ms'] * 1e3:.1f} us per launch, {cq['evaluations']} "
        f"element qualities, {cq['tflops']:.2f} TFLOP/s; mixed "
        f"{cq32['launch_ms'] * 1e3:.1f} us")
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
                   "passes_per_step": len(runs[-1]), "final_tets": final,
                   "tuning": tuning},
        "roofline": {"bound": "fp64", "achieved": cq["tflops"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": cq["tflops"] / FP64_PEAK_TFLOPS,
                     "traffic": traffic_per_launch(cq["kernel"]) if args.n == 64 else None,
                     "kernel": cq["kernel"], "launch_us": cq["launch_ms"] * 1e3,
                     "evaluations": cq["evaluations"], "flops_per_evaluation": FLOPS_PER_QUALITY,
                     "fp64_issue_frac": (fp64_wave_instructions(cq["kernel"]) /
                                         (FP64_ISSUE_PER_S * cq["launch_ms"] * 1e-3))
                     if args.n == 64 and fp64_wave_instructions(cq["kernel"]) else None,
                     "mixed_launch_us": cq32["launch_ms"] * 1e3},
        "roofline_measure": {"bound": "hbm", "achieved": roof["achieved"], "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": roof["achieved"] / HBM_PEAK_GBS,
                             "traffic": traffic_per_launch(roof["kernel"])
                             if args.n == 64 else None,
                             "kernel": roof["kernel"], "launch_us": roof["launch_ms"] * 1e3,
                             "bytes": roof["bytes"]},
        "mixed": mixed,
        "levelset_drivers": None,
        "cpu_baseline": None,
    }
    if args.ranges:
        line["ranges"] = named_ranges(ctx, base, opts)
    if not args.no_drivers:
        o = ctx.default_opts(3)
        o.verbosity = 0
        line["levelset_drivers"] = levelset_drivers(ctx, base, o, args.n)
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(ctx, base, opts, args.n, args.cpu_seconds)
        log(f"cpu baseline: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
