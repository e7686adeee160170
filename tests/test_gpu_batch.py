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
