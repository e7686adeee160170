"""Metric intersection and gradation limiting on the GPU against the oracle,
bit for bit (src/Omega_h_metric_intersect.hpp:18-269,
src/Omega_h_metric.cpp:156-220)."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def metric_pairs(dim, n, rng):
    """Random SPD pairs; three in four have one or two null eigenvalues so
    that every degenerate branch is taken."""
    a, b = [], []
    for k in range(n):
        pair = []
        for _ in range(2):
            q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
            lam = np.exp(rng.uniform(-3, 3, dim))
            if k % 4 >= 1:
                lam[rng.integers(dim)] = 0.0
            if k % 4 >= 2 and dim == 3:
                lam[rng.integers(dim)] = 0.0
            pair.append(O.symm2vector(q @ np.diag(lam) @ q.T))
        a.append(pair[0])
        b.append(pair[1])
    return np.concatenate(a), np.concatenate(b)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_intersect_metrics_identical(ctx, dim):
    a, b = metric_pairs(dim, 2000, np.random.default_rng(dim))
    got = ctx.intersect_metrics(dim, a, b)
    nc = dim * (dim + 1) // 2
    want = np.concatenate([O.kat_intersect(dim, a[i:i + nc], b[i:i + nc])
                           for i in range(0, len(a), nc)])
    bad = np.nonzero(got.view(np.int64) != want.view(np.int64))[0]
    assert len(bad) == 0, (len(bad), bad[:5])


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 8)])
def test_limit_metric_gradation_identical(ctx, dim, n):
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    met, nc = W.banded_aniso_metric(o.coords(), dim, np.random.default_rng(7))
    o.set_metric(met, nc)
    g.set_metric(met, nc)
    so = o.limit_metric_gradation(0.3, 1e-2)
    sg = g.limit_metric_gradation(0.3, 1e-2)
    assert so == sg and so > 1
    assert np.array_equal(o.metric().view(np.int64), g.metric().view(np.int64))
    # the main metric then adapts identically
    ro, po = o.adapt(O.default_opts(dim))
    rg, pg = g.adapt(ctx.default_opts(dim))
    assert [p[:3] for p in po] == [p[:3] for p in pg]
    assert np.array_equal(o.verts(dim), g.verts(dim))
    assert np.array_equal(o.coords().view(np.int64), g.coords().view(np.int64))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_interpolate_metrics_identical(ctx, dim):
    rng = np.random.default_rng(10 + dim)
    a, b = metric_pairs(dim, 500, rng)
    # interpolation needs full-rank metrics: lift the null eigenvalues
    nc = dim * (dim + 1) // 2
    eye = O.symm2vector(np.eye(dim))
    a = a + np.tile(eye, len(a) // nc)
    b = b + np.tile(eye, len(b) // nc)
    for t in (0.5, 0.125):
        got = ctx.interpolate_metrics(dim, a, b, t)
        want = O.interpolate_metrics(dim, a, b, t)
        assert np.array_equal(got.view(np.int64), want.view(np.int64))


@pytest.mark.parametrize("dim,n", [(2, 8), (3, 3)])
def test_approach_metric_loop_identical(ctx, dim, n):
    """The reference's idiom for a target metric (src/aniso_test.cpp:15-27):
    while (approach_metric) adapt, with a linear vertex tag along for the
    ride; every step bit for bit."""
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    c = o.coords().reshape(-1, dim)
    if dim == 2:
        init, nc = np.full(o.nents(0), 64.0), 1
        target = 1.0 / (0.03 + 0.1 * c[:, 0]) ** 2
    else:
        init, nc = np.tile([9.0, 9.0, 9.0, 0.0, 0.0, 0.0], o.nents(0)), 6
        target, _ = W.banded_aniso_metric(o.coords(), 3, np.random.default_rng(2))
    for m in (o, g):
        m.set_metric(init, nc)
        m.add_vertex_tag("target_metric", nc, 1, target)
        m.add_vertex_tag("xyz", dim, 0, o.coords())
    steps = 0
    while True:
        ro = o.approach_metric(O.default_opts(dim))
        rg = g.approach_metric(ctx.default_opts(dim))
        assert ro == rg
        if ro == 0:
            break
        o.adapt(O.default_opts(dim))
        g.adapt(ctx.default_opts(dim))
        steps += 1
        assert np.array_equal(o.verts(dim), g.verts(dim))
        assert np.array_equal(o.metric().view(np.int64), g.metric().view(np.int64))
        assert np.array_equal(o.vertex_tag("xyz").view(np.int64),
                              g.vertex_tag("xyz").view(np.int64))
    assert steps >= 2
    # the linear tag started as the coordinates and moved with them
    assert np.array_equal(g.vertex_tag("xyz"), g.coords())


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 6)])
def test_implied_metrics_identical(ctx, dim, n):
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    assert np.array_equal(o.implied_metrics().view(np.int64),
                          g.implied_metrics().view(np.int64))


def test_adapt_to_a_target_from_the_implied_metric(ctx):
    """src/aniso_test.cpp:15-27 on a 3D box: the implied metric becomes the
    metric, the level-set metric the target, then approach + adapt until
    the target is reached; identical to the oracle at every step."""
    o = O.Mesh.box(3, 3)
    g = A.Mesh.box(ctx, 3, 3)
    implied = g.implied_metrics(set_as_metric=True)
    target, nc = W.levelset_metric_3d(o.coords(), h_min=0.02, h_max=0.2, width=0.1)
    o.set_metric(implied, 6)
    for m in (o, g):
        m.add_vertex_tag("target_metric", nc, 1, target)
    steps = 0
    while True:
        ro = o.approach_metric(O.default_opts(3))
        rg = g.approach_metric(ctx.default_opts(3))
        assert ro == rg
        if ro == 0:
            break
        o.adapt(O.default_opts(3))
        g.adapt(ctx.default_opts(3))
        steps += 1
        assert np.array_equal(o.verts(3), g.verts(3))
        assert np.array_equal(o.metric().view(np.int64), g.metric().view(np.int64))
    assert steps >= 1


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 4)])
def test_inheritance_tags_identical(ctx, dim, n):
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    met, nc = W.iso_ring_metric_2d(o.coords()) if dim == 2 else W.levelset_metric_3d(o.coords())
    rng = np.random.default_rng(3)
    tags = [rng.random(o.nents(d) * 2) for d in range(dim + 1)]
    for m in (o, g):
        m.set_metric(met, nc)
        m.add_inherit_tag("field", 2, tags)
    o.adapt(O.default_opts(dim))
    g.adapt(ctx.default_opts(dim))
    for d in range(dim + 1):
        assert np.array_equal(o.inherit_tag("field", d), g.inherit_tag("field", d)), d


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


@pytest.mark.parametrize("dim,n", [(2, 8), (3, 3)])
def test_warp_and_adapt_loop_identical(ctx, dim, n):
    """do adapt while warp_to_limit (src/bend_test.cpp:55-65,
    src/Omega_h_approach.cpp:36-81) for the shear x += 2 z^2: return codes,
    coordinates, the warp left over and the mesh after every step, bit for bit."""
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    c = o.coords().reshape(-1, dim)
    w = np.zeros_like(c)
    w[:, 0] = 2.0 * c[:, dim - 1] ** 2
    for m in (o, g):
        m.set_metric(np.full(m.nents(0), float(n * n)), 1)
        m.add_vertex_tag("warp", dim, 0, w.ravel())
    steps = 0
    while True:
        o.adapt(O.default_opts(dim))
        g.adapt(ctx.default_opts(dim))
        ro = o.warp_to_limit(O.default_opts(dim))
        rg = g.warp_to_limit(ctx.default_opts(dim))
        assert ro == rg
        assert np.array_equal(o.verts(dim), g.verts(dim))
        assert same_bits(o.coords(), g.coords())
        wo, wg = o.vertex_tag("warp"), g.vertex_tag("warp")
        assert (wo is None) == (wg is None)
        if wo is not None:
            assert same_bits(wo, wg)
        if ro == 0:
            break
        steps += 1
        assert steps < 10
    assert steps >= 3
    assert same_bits(o.qualities(), g.qualities())


def test_warp_to_limit_single_vertex_and_errors(ctx):
    o = O.Mesh.box(2, 4)
    g = A.Mesh.box(ctx, 2, 4)
    c = o.coords().reshape(-1, 2)
    v = int(np.argmin(np.abs(c - 0.5).sum(1)))
    w = np.zeros_like(c)
    w[v] = [0.375, 0.0]
    for m in (o, g):
        m.set_metric(np.full(m.nents(0), 16.0), 1)
        m.add_vertex_tag("warp", 2, 0, w.ravel())
    assert o.warp_to_limit(O.default_opts(2)) == g.warp_to_limit(ctx.default_opts(2)) == 1
    assert same_bits(o.coords(), g.coords())
    assert same_bits(o.vertex_tag("warp"), g.vertex_tag("warp"))
    # exit_on_stall drops the warp where the last try left it
    g = A.Mesh.box(ctx, 2, 4)
    g.set_metric(np.full(g.nents(0), 16.0), 1)
    g.add_vertex_tag("warp", 2, 0, w.ravel())
    assert g.warp_to_limit(ctx.default_opts(2), exit_on_stall=True, max_niters=1) == 1
    assert g.vertex_tag("warp") is None
    assert g.coords().reshape(-1, 2)[v, 0] == 0.5 + 0.375 / 2
    # without it a stall is an error, as is an unacceptable starting mesh
    g.add_vertex_tag("warp", 2, 0, w.ravel())
    g.set_coords(c.ravel())
    with pytest.raises(A.OshError, match="warp step"):
        g.warp_to_limit(ctx.default_opts(2), max_niters=1)
    opts = ctx.default_opts(2)
    opts.min_quality_allowed = 0.9
    g.set_coords(c.ravel())
    with pytest.raises(A.OshError, match="minimum element quality"):
        g.warp_to_limit(opts)
    assert g.warp_to_limit(ctx.default_opts(2)) == 1
    assert g.vertex_tag("warp") is not None


def jittered_pair(ctx, dim, n, amp, seed):
    """The same jittered box (interior vertices moved by up to amp / n per
    axis) on both sides."""
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    c = o.coords().reshape(-1, dim).copy()
    inner = np.all((c > 1e-12) & (c < 1 - 1e-12), axis=1)
    c[inner] += np.random.default_rng(seed).uniform(-amp, amp, (inner.sum(), dim)) / n
    o.set_coords(c.ravel())
    g.set_coords(c.ravel())
    return o, g


def assert_meshes_same(o, g, dim):
    assert [o.nents(d) for d in range(dim + 1)] == [g.nents(d) for d in range(dim + 1)]
    assert np.array_equal(o.verts(dim), g.verts(dim))
    assert same_bits(o.coords(), g.coords())
    assert o.metric_ncomps == g.metric_ncomps
    assert same_bits(o.metric(), g.metric())


@pytest.mark.parametrize("dim,n,amp,seed", [(2, 8, 0.0, 0), (2, 8, 0.4, 2), (3, 4, 0.35, 3)])
def test_implied_isos_and_ill_metric_identical(ctx, dim, n, amp, seed):
    """get_implied_isos and compute_ill_metric (src/Omega_h_metric.cpp:316-321,
    src/Omega_h_fix.cpp:16-79) for both isotropies, bit for bit."""
    o, g = jittered_pair(ctx, dim, n, amp, seed)
    assert same_bits(o.implied_isos(), g.implied_isos())
    for iso in (0, 1):
        o, g = jittered_pair(ctx, dim, n, amp, seed)
        o.ill_metric(O.default_opts(dim), iso)
        g.ill_metric(ctx.default_opts(dim), iso)
        assert_meshes_same(o, g, dim)
        assert same_bits(o.qualities(), g.qualities())


@pytest.mark.parametrize("dim,n,amp,seed,iso", [(2, 8, 0.4, 2, 0), (2, 8, 0.4, 2, 1),
                                                (3, 4, 0.35, 3, 0), (3, 4, 0.35, 3, 1)])
def test_fix_identical(ctx, dim, n, amp, seed, iso):
    """fix (src/Omega_h_fix.cpp:90-134): the ill metric, then adapt() with the
    allowed quality raised, call by call the same."""
    o, g = jittered_pair(ctx, dim, n, amp, seed)
    co = o.fix(O.default_opts(dim), iso)
    cg = g.fix(ctx.default_opts(dim), iso)
    assert co == cg >= 1
    assert_meshes_same(o, g, dim)


@pytest.mark.parametrize("dim,n,amp,seed,nc", [(2, 8, 0.4, 2, 1), (2, 8, 0.4, 2, 3),
                                               (3, 4, 0.35, 3, 1), (3, 4, 0.35, 3, 6)])
def test_grade_fix_adapt_identical(ctx, dim, n, amp, seed, nc):
    """grade_fix_adapt (src/Omega_h_fix.cpp:136-176) from a jittered box with no
    metric, to a banded target: the repair and the approach loop, bit for bit."""
    o, g = jittered_pair(ctx, dim, n, amp, seed)
    c = o.coords().reshape(-1, dim)
    h = 0.5 / n + 0.5 * np.abs(c[:, 0] - 0.5) / n
    if nc == 1:
        target = 1.0 / h ** 2
    else:
        t = np.zeros((len(c), nc))
        t[:, :dim] = (1.0 / h ** 2)[:, None]
        t[:, 0] *= 2.0
        target = t.ravel()
    so = o.grade_fix_adapt(O.default_opts(dim), target, nc)
    sg = g.grade_fix_adapt(ctx.default_opts(dim), target, nc)
    assert so == sg >= 1
    assert_meshes_same(o, g, dim)
    assert g.vertex_tag("target_metric") is None


@pytest.mark.parametrize("dim,n", [(2, 16), (3, 4)])
def test_pointwise_elem_tags_identical(ctx, dim, n):
    """An element tag carried pointwise (src/Omega_h_transfer.cpp:324-335,
    515-583) through a whole adapt(): refinement, coarsening and swap
    products, bit for bit."""
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    met, nc = W.iso_ring_metric_2d(o.coords()) if dim == 2 else W.levelset_metric_3d(o.coords())
    vals = np.random.default_rng(3).normal(size=o.nents(dim) * 2)
    for m in (o, g):
        m.set_metric(met, nc)
        m.add_elem_tag("field", 2, vals)
    ro, po = o.adapt(O.default_opts(dim))
    rg, pg = g.adapt(ctx.default_opts(dim))
    assert ro == rg and [p[:3] for p in po] == [p[:3] for p in pg]
    assert {p[0] for p in po} >= ({0, 1} if dim == 2 else {0, 1, 2, 3})
    assert_meshes_same(o, g, dim)
    assert same_bits(o.elem_tag("field"), g.elem_tag("field"))
    # the values are the original ones, moved around
    assert np.isin(g.elem_tag("field"), vals).all()


def test_verbosity_reports(ctx, capfd):
    """The reports of adapt() at the reference's verbosity levels
    (src/Omega_h_adapt.cpp:101-288, src/Omega_h_histogram.cpp:56-85)."""
    assert ctx.default_opts(3).verbosity == 2  # EACH_REBUILD, as AdaptOpts(dim)
    o = O.Mesh.box(2, 16)
    met, nc = W.iso_ring_metric_2d(o.coords())
    g = A.Mesh.box(ctx, 2, 16)
    g.set_metric(met, nc)
    opts = ctx.default_opts(2)
    did, passes = g.adapt(opts)
    out = capfd.readouterr().out
    assert "before adapting:" in out and "addressing edge lengths" in out
    assert "512 triangles, quality [" in out and "edges, length [" in out
    assert ("refining " in out) and ("coarsening " in out)
    assert "adapting took " in out
    assert out.count("triangles, quality [") == 1 + len(passes)
    g = A.Mesh.box(ctx, 2, 16)
    g.set_metric(met, nc)
    opts.verbosity = 0
    g.adapt(opts)
    assert capfd.readouterr().out == ""
