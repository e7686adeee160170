"""The request server (continuous batching of adapt() requests, one worker
context per slot): every result is the same alone and in the mix, and the
same as the oracle's."""
import numpy as np
import pytest

import oracle_lib as O

A = pytest.importorskip("omega_h_amd")


def fnv1a(*arrays):
    h = 1469598103934665603
    for a in arrays:
        for b in np.ascontiguousarray(a).view(np.uint8):
            h ^= int(b)
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def requests():
    rng = np.random.default_rng(11)
    out = []
    for k in range(12):
        dim = 2 if k % 3 == 0 else 3
        n = int(rng.integers(6, 20)) if dim == 2 else int(rng.integers(2, 5))
        q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
        h = np.exp(rng.uniform(np.log(0.08), np.log(0.3), dim))
        m = q @ np.diag(1.0 / h ** 2) @ q.T
        v = O.symm2vector(m) if k % 2 else np.array([1.0 / h[0] ** 2])
        out.append((dim, n, v))
    return out


def test_submit_rejects_bad_requests_before_running_them():
    try:
        s = A.Server(nslots=1, hbm_budget_bytes=1e9)
    except A.OshError:
        pytest.skip("no GPU")
    with pytest.raises(A.OshError):
        s.submit(4, 3, [1.0], A.AdaptOpts())
    with pytest.raises(A.OshError):
        s.submit(3, 3, [1.0, 2.0], A.AdaptOpts())
    s.close()


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix():
    ctx = A.Context(0)
    reqs = requests()
    server = A.Server(nslots=4, hbm_budget_bytes=50e9)
    print(f"serving {len(reqs)} requests on 4 slots", flush=True)
    tickets = [server.submit(d, n, v, ctx.default_opts(d), want_digest=True) for d, n, v in reqs]
    mixed = [server.wait(t) for t in tickets]
    server.close()
    for (d, n, v), r in zip(reqs, mixed):
        g = A.Mesh.box(ctx, d, n)
        g.set_metric(np.tile(v, g.nents(0)), len(v))
        g.adapt(ctx.default_opts(d))
        assert r.digest == g.digest(), (d, n)
        assert r.nelems_out == g.nents(d)
        assert r.latency_seconds >= r.run_seconds >= r.pass_seconds > 0
    # and the oracle's result for the first few
    for (d, n, v), r in list(zip(reqs, mixed))[:4]:
        o = O.Mesh.box(d, n)
        o.set_metric(np.tile(v, o.nents(0)), len(v))
        o.adapt(O.default_opts(d))
        assert r.digest == fnv1a(o.verts(d), o.coords()), (d, n)


@pytest.mark.gpu
def test_admission_under_a_tiny_budget_runs_requests_one_at_a_time():
    ctx = A.Context(0)
    reqs = requests()[:5]
    server = A.Server(nslots=3, hbm_budget_bytes=1.0)
    tickets = [server.submit(d, n, v, ctx.default_opts(d), want_digest=True) for d, n, v in reqs]
    res = [server.wait(t) for t in tickets]
    server.close()
    # every request exceeds the budget, so each ran alone, in arrival order:
    # a request's queue time covers the run times of those before it
    for i in range(1, len(res)):
        assert res[i].queue_seconds >= res[i - 1].run_seconds * 0.9
    assert all(r.status == 0 for r in res)


@pytest.mark.gpu
def test_timed_out_request_fails_and_frees_its_slot():
    """A request past its time limit fails with a status (it does not hold
    its slot); requests after the limit is lifted run normally."""
    ctx = A.Context(0)
    reqs = requests()[:3]
    server = A.Server(nslots=1, hbm_budget_bytes=50e9)
    server.set_timeout(1e-9)
    late = [server.submit(d, n, v, ctx.default_opts(d)) for d, n, v in reqs]
    for t in late:
        with pytest.raises(A.OshError, match="timed out"):
            server.wait(t)
    server.set_timeout(0)
    ok = [server.wait(server.submit(d, n, v, ctx.default_opts(d))) for d, n, v in reqs]
    server.close()
    assert all(r.status == 0 and r.nelems_out > 0 for r in ok)


@pytest.mark.gpu
def test_serving_stream_alone_equals_mix():
    """A short Poisson stream of the serving workload (30 % 2D / 70 % 3D
    boxes, element counts log-uniform in [1e3, 2e6], rotated metrics with
    eigen-lengths log-uniform in [0.005, 0.1], each clamped to 2e4 expected
    elements by the reference's element-count rule) submitted at its arrival
    times to a 4-slot server: every result equals the same request adapted
    alone, and the smallest ones equal the oracle's."""
    import time

    from omega_h_amd import workloads as W
    ctx = A.Context(0)
    rng = np.random.default_rng(2024)
    stream = W.serve_stream(rng, 20.0, 0.6, 2e6, 0.005, 0.1, clamp=2e4)
    assert len(stream) >= 6
    opts = {d: ctx.default_opts(d) for d in (2, 3)}
    for o in opts.values():
        o.verbosity = 0
    server = A.Server(nslots=4, hbm_budget_bytes=100e9)
    print(f"serving a stream of {len(stream)} requests", flush=True)
    t0 = time.perf_counter()
    tickets = []
    for at, (d, n, v) in stream:
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(server.submit(d, n, v, opts[d], want_digest=True))
    mixed = [server.wait(t) for t in tickets]
    server.close()
    sizes = []
    for (_, (d, n, v)), r in zip(stream, mixed):
        g = A.Mesh.box(ctx, d, n)
        g.set_metric(np.tile(v, g.nents(0)), len(v))
        g.adapt(opts[d])
        assert r.digest == g.digest(), (d, n)
        assert r.nelems_out == g.nents(d)
        sizes.append(g.nents(d))
    smallest = sorted(range(len(stream)), key=lambda i: stream[i][1][1])[:2]
    for i in smallest:
        d, n, v = stream[i][1]
        o = O.Mesh.box(d, n)
        o.set_metric(np.tile(v, o.nents(0)), len(v))
        oo = O.default_opts(d)
        oo.verbosity = 0
        o.adapt(oo)
        assert mixed[i].digest == fnv1a(o.verts(d), o.coords()), (d, n)


# request 1386 of the serving workload's 20 requests/s stream (seed 20260101,
# after the 5 requests/s stream): a 60^3 box (1,296,000 tets) under a constant
# anisotropic metric, which adapt() leaves after 4,639 swap and 283 sliver
# passes, unable to reach min_quality_desired
LONG_REQUEST = (3, 60, [5790.258056459252, 2135.76209205912, 494.46986710579586,
                        3207.769618264246, 774.5229738639886, 1204.9854838345045])


@pytest.mark.gpu
def test_adapt_with_more_passes_than_the_stats_list():
    """An adapt() of more passes than osh_adapt_stats lists (OSH_MAX_PASSES)
    runs to its end, alone and as a served request, and the totals count
    every pass."""
    dim, n, v = LONG_REQUEST
    ctx = A.Context(0)
    opts = ctx.default_opts(dim)
    opts.verbosity = 0
    m = A.Mesh.box(ctx, dim, n)
    m.set_metric(np.tile(v, m.nents(0)), len(v))
    did, passes = m.adapt(opts)
    total, elems_in, _ = m.last_adapt_totals
    assert did == 1 and len(passes) == 4096
    assert total == 4922 and elems_in >= sum(p[1] for p in passes)
    assert np.bincount([p[0] for p in passes], minlength=4)[:2].tolist() == [0, 0]
    server = A.Server(nslots=1)
    r = server.wait(server.submit(dim, n, v, opts, want_digest=True))
    server.close()
    assert r.npasses == 4922 and r.nelems_out == m.nents(dim)
