"""The batched form on the GPU: meshes joined into one (osh_mesh_join),
adapted as one mesh, split back (osh_mesh_split).  Every part must equal,
bit for bit (the four mesh digests: vertex lists and classification,
coordinates, metric, lengths and qualities), the same mesh adapted alone on
the GPU and adapted alone by the oracle."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def make(mod, ctx, dim, n, v):
    m = mod.Mesh.box(dim, n) if ctx is None else mod.Mesh.box(ctx, dim, n)
    m.set_metric(np.tile(v, m.nents(0)), len(v))
    return m


@pytest.mark.parametrize("dim,seed,nparts,n_range,h_range,cap,noracle", [
    (3, 11, 4, (2, 4), (0.08, 0.6), 2500, 4),
    (2, 5, 5, (3, 10), (0.02, 0.4), 3000, 5),
    # a sample of the 256-box workload's generator: boxes of 8 to 16 cells per
    # side, eigen-lengths 0.005 to 0.1, clamped to 30,000 expected elements
    (3, 0x0A5A17E5EED5EED5, 12, (8, 16), (0.005, 0.1), 30000, 2),
])
def test_joined_adapt_equals_separate_adapts(ctx, dim, seed, nparts, n_range, h_range, cap,
                                             noracle):
    boxes = W.batch_boxes(nparts, seed, n_range=n_range, h_range=h_range, max_nelems=cap,
                          dim=dim)
    opts = ctx.default_opts(dim)
    opts.verbosity = 0
    alone = []
    for n, v, _, _ in boxes:
        m = make(A, ctx, dim, n, v)
        m.adapt(opts)
        alone.append(m)
    joined = A.Mesh.join(ctx, [make(A, ctx, dim, n, v) for n, v, _, _ in boxes])
    joined.adapt(opts)
    counts = joined.part_counts(nparts)
    parts = joined.split(nparts)
    for k, (p, m) in enumerate(zip(parts, alone)):
        assert [p.nents(d) for d in range(dim + 1)] == [m.nents(d) for d in range(dim + 1)]
        assert counts[k].tolist() == [m.nents(d) for d in range(dim + 1)]
        assert p.digests() == m.digests(), f"part {k}"
    oo = O.default_opts(dim)
    oo.verbosity = 0
    for k in range(noracle):
        n, v, _, _ = boxes[k]
        o = make(O, None, dim, n, v)
        o.adapt(oo)
        assert o.digests() == parts[k].digests(), f"part {k} against the oracle"


@pytest.mark.parametrize("dim,seed,nparts,n_range,h_range,cap", [
    (3, 0x0A5A17E5EED5EED5, 16, (4, 14), (0.005, 0.1), 20000),
    (2, 7, 12, (4, 40), (0.005, 0.1), 8000),
])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_compacting_batch_equals_separate_adapts(ctx, dim, seed, nparts, n_range, h_range, cap,
                                                 precision):
    """osh_adapt_batch (joined meshes, frozen parts set aside as they appear)
    leaves every mesh exactly where its own adapt() leaves it; a mesh that
    already meets the goals is left alone and reported as such."""
    boxes = W.batch_boxes(nparts, seed, n_range=n_range, h_range=h_range, max_nelems=cap,
                          dim=dim)
    # a box whose edges are all within the length goals (metric edge
    # lengths 0.8, 1.13 and 1.39 for the box's three edge directions) and
    # whose elements are good: adapt() has nothing to do
    n_ok = 5
    h = 1.25 / n_ok
    ok = np.array([1.0 / h ** 2] + [1.0 / h ** 2] * (dim - 1) + [0.0] * (3 if dim == 3 else 1))
    boxes.insert(3, (n_ok, ok, 0.0, 1.0))
    opts = ctx.default_opts(dim)
    opts.verbosity = 0
    opts.precision = precision
    alone, want_did = [], []
    for n, v, _, _ in boxes:
        m = make(A, ctx, dim, n, v)
        want_did.append(m.adapt(opts)[0])
        alone.append(m)
    assert want_did[3] == 0 and sum(want_did) >= len(boxes) - 1
    batch = [make(A, ctx, dim, n, v) for n, v, _, _ in boxes]
    did, passes, _ = A.adapt_batch(batch, opts)
    assert did == want_did
    assert len(passes) > 0
    for k, (b, m) in enumerate(zip(batch, alone)):
        assert b.digests() == m.digests(), f"mesh {k}"
