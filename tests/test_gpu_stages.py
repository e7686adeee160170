"""The stage entry points of the C ABI on the device against the oracle, bit
for bit: midpoint metrics (get_mident_metrics), the swap choices of listed
interior edges (swap3d_qualities with the choice of a triangulation),
invert_adj (with the reference's known answer, src/unit_array_algs.cpp:
139-148), reflect_down of rotated and flipped copies of a mesh's own
elements, and collect_marked."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def pair(ctx, dim, n, metric, jitter=0.0):
    """The same box on both sides; with `jitter`, interior vertices moved by
    up to that many cell sizes along each axis (poor elements for swaps to
    improve)."""
    o = O.Mesh.box(dim, n)
    g = A.Mesh.box(ctx, dim, n)
    if jitter:
        x = o.coords().reshape(-1, dim).copy()
        inner = np.all((x > 1e-12) & (x < 1 - 1e-12), axis=1)
        rng = np.random.default_rng(n)
        x[inner] += rng.uniform(-jitter, jitter, (inner.sum(), dim)) / n
        o.set_coords(x.ravel())
        g.set_coords(x.ravel())
    met, nc = metric(o.coords())
    o.set_metric(met, nc)
    g.set_metric(met, nc)
    return o, g


METRICS = {
    "iso2d": (2, lambda c: W.iso_ring_metric_2d(c)),
    "aniso2d": (2, lambda c: W.rotated_constant_metric(len(c) // 2, 2,
                                                      np.random.default_rng(3))),
    "levelset3d": (3, lambda c: W.levelset_metric_3d(c)),
    "banded3d": (3, lambda c: W.banded_aniso_metric(c, 3, np.random.default_rng(5))),
    # edge length 1/5: a jittered 5^3 box has many improving swaps under it
    "iso3d": (3, lambda c: (np.full(len(c) // 3, 25.0), 1)),
}


@pytest.mark.parametrize("name", list(METRICS))
def test_mident_metrics(ctx, name):
    dim, metric = METRICS[name]
    o, g = pair(ctx, dim, 6, metric)
    edges = np.arange(o.nents(1), dtype=np.int32)[::-1].copy()
    want = o.mident_metrics(edges)
    got = g.mident_metrics(edges)
    assert np.array_equal(want.view(np.int64), got.view(np.int64))


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", ["iso3d", "banded3d"])
def test_swap3d_qualities(ctx, name, precision):
    dim, metric = METRICS[name]
    o, g = pair(ctx, dim, 5, metric, jitter=0.2)
    _, edim = g.classification(1)
    edges = np.flatnonzero(edim == 3).astype(np.int32)
    opts = ctx.default_opts(3)
    opts.precision = precision
    q, cfg = g.swap3d_qualities(opts, edges)
    wq, wcfg = o.swap3d_qualities(opts.max_length_allowed, edges)
    # the decisions (which edges a swap would keep, at which quality)
    assert np.array_equal(q.view(np.int64), wq.view(np.int64))
    assert (wq > 0).any(), "the case must have improving swaps"
    if precision == 0:
        assert np.array_equal(cfg, wcfg)
    else:
        # a candidate the FP32 screen rejects reports no triangulation
        kept = wq > 0
        assert np.array_equal(cfg[kept], wcfg[kept])


def test_swap3d_qualities_refuses_boundary_edges(ctx):
    _, g = pair(ctx, 3, 3, METRICS["levelset3d"][1])
    _, edim = g.classification(1)
    bdry = np.flatnonzero(edim < 3)[:1].astype(np.int32)
    with pytest.raises(A.OshError, match="interior"):
        g.swap3d_qualities(ctx.default_opts(3), bdry)


def test_invert_adj_known_answer(ctx):
    a2ab, ab2b, codes = ctx.invert_adj([0, 1, 2, 2, 3, 0], 3, 4)
    assert a2ab.tolist() == [0, 2, 3, 5, 6]
    assert ab2b.tolist() == [0, 1, 0, 0, 1, 1]
    mc = O.make_code
    assert codes.tolist() == [mc(0, 0, 0), mc(0, 0, 2), mc(0, 0, 1), mc(0, 0, 2),
                              mc(0, 0, 0), mc(0, 0, 1)]


@pytest.mark.parametrize("deg,nhigh,nlow", [(1, 0, 5), (2, 1000, 37), (3, 5000, 900),
                                            (6, 20000, 4000)])
def test_invert_adj_random(ctx, deg, nhigh, nlow):
    rng = np.random.default_rng(deg * 100 + nlow)
    hl2l = rng.integers(0, nlow, nhigh * deg).astype(np.int32)
    codes = rng.integers(0, 8, nhigh * deg).astype(np.int8)  # flip and rotation bits
    for dc in (None, codes):
        want = O.invert_adj(hl2l, deg, nlow, dc)
        got = ctx.invert_adj(hl2l, deg, nlow, dc)
        for w, h in zip(want, got):
            assert np.array_equal(w, h)


@pytest.mark.parametrize("dim,high", [(2, 2), (3, 2), (3, 3)])
def test_reflect_down_of_rotated_copies(ctx, dim, high):
    _, g = pair(ctx, dim, 4, METRICS["levelset3d" if dim == 3 else "iso2d"][1])
    hv = g.verts(high).reshape(-1, high + 1)
    rng = np.random.default_rng(high)
    # each simplex rotated, half of them also flipped (two vertices swapped)
    rot = np.array([np.roll(r, k) for r, k in zip(hv, rng.integers(0, high + 1, len(hv)))])
    flip = rng.random(len(hv)) < 0.5
    rot[flip, 0], rot[flip, 1] = rot[flip, 1].copy(), rot[flip, 0].copy()
    pv = rot.ravel().astype(np.int32)
    down, codes = g.reflect_down(high, pv)
    want_down, want_codes = O.reflect_down(pv, g.verts(high - 1), g.nents(0), high, high - 1)
    assert np.array_equal(down, want_down)
    assert np.array_equal(codes, want_codes)
    # unrotated copies find exactly the mesh's own sides
    d0, _ = g.reflect_down(high, g.verts(high))
    assert np.array_equal(d0, g.down(high)[0])


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 100000])
def test_collect_marked(ctx, n):
    rng = np.random.default_rng(n)
    marks = (rng.random(n) < 0.3).astype(np.int8) * rng.integers(1, 5, n).astype(np.int8)
    assert np.array_equal(ctx.collect_marked(marks), np.flatnonzero(marks))
