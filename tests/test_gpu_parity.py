"""Parity of the MI355X path (through the C ABI) with the oracle.

Integer work (connectivity, codes, classification, pass sequence, entity
counts) must be identical, and floating-point results (coordinates, metrics,
lengths, qualities) are compared bit for bit: both sides evaluate the same
IEEE operation sequence (no contraction, same transcendental methods).
"""
import math

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.dtype.kind == "f":
        # bit-exact, except that any two NaNs match (their sign bit is the
        # platform's default NaN, not a result)
        diff = a.view(np.int64) != b.view(np.int64)
        bad = np.nonzero(diff & ~(np.isnan(a) & np.isnan(b)))[0]
    else:
        bad = np.nonzero(a != b)[0]
    assert len(bad) == 0, (what, len(bad), bad[:5], a[bad[:5]], b[bad[:5]])


def compare_meshes(o, g, tags=True):
    dim = o.dim
    assert g.dim == dim
    for d in range(dim + 1):
        assert o.nents(d) == g.nents(d), (d, o.nents(d), g.nents(d))
    for d in range(1, dim + 1):
        assert_same(o.verts(d), g.verts(d), f"verts {d}")
        od, oc = o.down(d)
        gd, gc = g.down(d)
        assert_same(od, gd, f"down {d}")
        if d >= 2:
            assert_same(oc, gc, f"codes {d}")
    assert_same(o.coords(), g.coords(), "coords")
    for x, y in zip(o.dual(), g.dual()):
        assert_same(x, y, "dual")
    assert_same(o.sizes(), g.sizes(), "sizes")
    if tags:
        for d in range(dim + 1):
            oi, od_ = o.classification(d)
            gi, gd_ = g.classification(d)
            assert_same(oi, gi, f"class_id {d}")
            assert_same(od_, gd_, f"class_dim {d}")
        assert_same(o.metric(), g.metric(), "metric")
        assert_same(o.lengths(), g.lengths(), "lengths")
        assert_same(o.qualities(), g.qualities(), "qualities")


def boxes(ctx, dim, n):
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    c = o.coords()
    met, nc = W.iso_ring_metric_2d(c) if dim == 2 else W.levelset_metric_3d(c)
    o.set_metric(met, nc)
    g.set_metric(met, nc)
    return o, g


@pytest.mark.parametrize("which,name,lo,hi", [(0, "om_log", -300, 300), (1, "om_exp", -700, 700),
                                              (2, "om_cbrt", -300, 300), (3, "om_acos", -1, 1),
                                              (4, "om_cos", -4.5, 4.5)])
def test_device_transcendentals_match_oracle_bitwise(ctx, which, name, lo, hi):
    rng = np.random.default_rng(which)
    if which in (0, 2):
        xs = 10.0 ** rng.uniform(lo, hi, 200000)
        if which == 2:
            xs *= rng.choice([-1.0, 1.0], len(xs))
    else:
        xs = rng.uniform(lo, hi, 200000)
    xs = np.concatenate([xs, [1.0, 0.5, 2.0, 1e-310, 3.0, -1.0, 0.0]])
    if which == 3:
        xs = np.clip(xs, -1.0, 1.0)
    got = ctx.eval_math(which, xs)
    f = getattr(O.lib(), name)
    want = np.array([f(float(x)) for x in xs])
    assert_same(want, got, name)


@pytest.mark.parametrize("dim,n", [(2, 1), (2, 5), (3, 1), (3, 2), (3, 5)])
def test_build_box_internal_identical(ctx, dim, n):
    o = O.Mesh.box(dim, n, internal=True)
    g = A.Mesh.box(ctx, dim, n, internal=True)
    compare_meshes(o, g, tags=False)
    for low, high in [(0, 1), (1, dim), (0, dim), (dim - 1, dim)]:
        for x, y in zip(o.up(low, high), g.up(low, high)):
            assert_same(x, y, f"up {low} {high}")
    for d in (0, 1):
        for x, y in zip(o.star(d), g.star(d)):
            assert_same(x, y, f"star {d}")


def test_reference_kats_on_device(ctx):
    # src/unit_mesh.cpp:424-447 and :382-391
    eev, loop = A.Mesh.box(ctx, 3, 1, internal=True).find_loop(6)
    assert list(eev) == [7, 0] and list(loop) == [2, 3, 1, 5, 4, 6]
    g = A.Mesh.box(ctx, 2, 1, internal=True)
    g.set_metric(np.tile([1.0, 1.0, 0.0], g.nents(0)), 3)
    q = g.refine_qualities(np.arange(g.nents(1)))
    assert np.allclose(q, [0.494872, 0.494872, 0.866025, 0.494872, 0.494872], atol=1e-4)
    g = A.Mesh.from_elems(ctx, 3, [0, 1, 2, 3], 4)
    a, b = g.star(1)
    assert list(b) == [1, 3, 4, 2, 5, 3, 0, 2, 5, 4, 0, 4, 5, 1, 3, 0, 1, 5,
                       4, 2, 2, 0, 3, 5, 1, 1, 2, 4, 3, 0]


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 6), (3, 16)])
def test_box_and_measures_identical(ctx, dim, n):
    o, g = boxes(ctx, dim, n)
    compare_meshes(o, g)


def test_indset_identical(ctx):
    o, g = boxes(ctx, 3, 6)
    rng = np.random.default_rng(4)
    for d in (0, 1):
        n = o.nents(d)
        q = rng.integers(0, 5, n).astype(np.float64)
        c = (rng.random(n) < 0.7).astype(np.int8)
        assert_same(o.indset(d, q, c), g.indset(d, q, c), f"indset {d}")


# verbosity 2 (the default, EACH_REBUILD) reports every pass; 0 is silent
@pytest.mark.parametrize("verbosity", [2, 0], ids=["each_rebuild", "silent"])
@pytest.mark.parametrize("dim,n", [(2, 32), (3, 4), (3, 8)])
def test_adapt_identical(ctx, dim, n, verbosity):
    o, g = boxes(ctx, dim, n)
    ro, po = o.adapt(O.default_opts(dim))
    go = ctx.default_opts(dim)
    go.verbosity = verbosity
    rg, pg = g.adapt(go)
    assert ro == rg
    assert [p[:3] for p in po] == [p[:3] for p in pg]
    compare_meshes(o, g)


def test_pass_by_pass_identical_3d(ctx):
    o, g = boxes(ctx, 3, 12)
    oo = O.default_opts(3)
    go = ctx.default_opts(3)
    for name in ("refine_by_size", "refine_by_size", "coarsen_by_size", "refine_by_size",
                 "coarsen_by_size"):
        assert getattr(o, name)(oo) == getattr(g, name)(go), name
        compare_meshes(o, g)
    o.swap_edges(oo)
    g.swap_edges(go)
    compare_meshes(o, g)
    o.coarsen_slivers(oo)
    g.coarsen_slivers(go)
    compare_meshes(o, g)


def test_isotropic_3d_adapt_identical(ctx):
    o = O.Mesh.box(3, 6)
    g = A.Mesh.box(ctx, 3, 6)
    c = o.coords().reshape(-1, 3)
    h = 0.05 + 0.2 * np.abs(np.linalg.norm(c - 0.5, axis=1) - 0.3)
    met = 1.0 / (h * h)
    o.set_metric(met, 1)
    g.set_metric(met, 1)
    ro, po = o.adapt(O.default_opts(3))
    go = ctx.default_opts(3)
    go.verbosity = 0
    rg, pg = g.adapt(go)
    assert [p[:3] for p in po] == [p[:3] for p in pg]
    compare_meshes(o, g)


def test_anisotropic_2d_adapt_identical(ctx):
    o = O.Mesh.box(2, 20)
    g = A.Mesh.box(ctx, 2, 20)
    met, nc = W.rotated_constant_metric(o.nents(0), 2, np.random.default_rng(9))
    o.set_metric(met, nc)
    g.set_metric(met, nc)
    ro, po = o.adapt(O.default_opts(2))
    go = ctx.default_opts(2)
    go.verbosity = 0
    rg, pg = g.adapt(go)
    assert [p[:3] for p in po] == [p[:3] for p in pg]
    compare_meshes(o, g)


@pytest.mark.parametrize("dim,n", [(2, 12), (3, 8)])
def test_measure_entry_points_on_device_pointers(ctx, dim, n):
    from omega_h_amd.hip import DeviceArray
    o, g = boxes(ctx, dim, n)
    nc = g.metric_ncomps
    coords, metric = DeviceArray(g.coords()), DeviceArray(g.metric())
    for d, fn, want in ((dim, ctx.measure_qualities, o.qualities()),
                        (1, ctx.measure_edges_metric, o.lengths())):
        nents = g.nents(d)
        ev2v = DeviceArray(g.verts(d))
        out = DeviceArray(dtype=np.float64, n=nents)
        fn(dim, nc, nents, ev2v.ptr, coords.ptr, metric.ptr, out.ptr)
        ctx.synchronize()  # the entry points are asynchronous on the context's stream
        assert_same(want, out.numpy(), f"measure dim {d}")
        # a subset through the entity list
        sub = np.arange(1, nents, 3, dtype=np.int32)
        dsub = DeviceArray(sub)
        out2 = DeviceArray(dtype=np.float64, n=len(sub))
        fn(dim, nc, len(sub), ev2v.ptr, coords.ptr, metric.ptr, out2.ptr, dsub.ptr)
        ctx.synchronize()
        assert_same(want[sub], out2.numpy(), f"measure subset dim {d}")


def test_measure_rejects_misaligned_metric(ctx):
    from omega_h_amd.hip import DeviceArray
    o, g = boxes(ctx, 3, 2)
    coords, metric = DeviceArray(g.coords()), DeviceArray(g.metric())
    ev2v = DeviceArray(g.verts(3))
    out = DeviceArray(dtype=np.float64, n=g.nents(3))
    # the check runs on the host before any launch
    with pytest.raises(A.OshError, match="aligned"):
        ctx.measure_qualities(3, 6, g.nents(3), ev2v.ptr, coords.ptr, metric.ptr + 8, out.ptr)
