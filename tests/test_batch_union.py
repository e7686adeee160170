"""The batched form's premise, checked on the oracle: adapting one disjoint
mesh made of several meshes (each numbered after the previous ones) gives,
for every part, exactly the mesh that adapting that part alone gives, and
each part stays one contiguous block of every dimension.  Parts have
different sizes and metrics, so they need different numbers of passes: the
joined run's extra passes must leave finished parts untouched."""
import numpy as np
import pytest

import oracle_lib as O
from omega_h_amd import workloads as W


def parts_of(dim, seed, nparts, n_range, h_range, cap):
    boxes = W.batch_boxes(nparts, seed, n_range=n_range, h_range=h_range, max_nelems=cap,
                          dim=dim)
    meshes = []
    for n, v, _, _ in boxes:
        m = O.Mesh.box(dim, n)
        m.set_metric(np.tile(v, m.nents(0)), len(v))
        meshes.append(m)
    return meshes


def snapshot(m):
    d = m.dim
    return {"verts": [m.verts(k) for k in range(1, d + 1)],
            "coords": m.coords(), "metric": m.metric(),
            "class": [m.classification(k) for k in range(d + 1)],
            "lengths": m.lengths(), "qualities": m.qualities()}


def split(m, nparts):
    """Slices of the joined mesh by the part of each entity's first vertex."""
    d = m.dim
    vpart = m.vertex_tag("part").astype(np.int64)
    assert np.array_equal(vpart, np.sort(vpart))
    parts_of_dim = [vpart] + [vpart[m.verts(k).reshape(-1, k + 1)[:, 0]] for k in range(1, d + 1)]
    first = []
    for k in range(d + 1):
        p = parts_of_dim[k]
        assert np.all(np.diff(p) >= 0), "a part is not contiguous"
        first.append(np.searchsorted(p, np.arange(nparts + 1)))
    coords, metric = m.coords(), m.metric()
    nc = m.metric_ncomps
    lengths, quals = m.lengths(), m.qualities()
    out = []
    for j in range(nparts):
        s = {}
        s["verts"] = [m.verts(k)[first[k][j] * (k + 1):first[k][j + 1] * (k + 1)] - first[0][j]
                      for k in range(1, d + 1)]
        s["coords"] = coords[first[0][j] * d:first[0][j + 1] * d]
        s["metric"] = metric[first[0][j] * nc:first[0][j + 1] * nc]
        s["class"] = [tuple(a[first[k][j]:first[k][j + 1]] for a in m.classification(k))
                      for k in range(d + 1)]
        s["lengths"] = lengths[first[1][j]:first[1][j + 1]]
        s["qualities"] = quals[first[d][j]:first[d][j + 1]]
        out.append(s)
    return out


def same(a, b):
    assert len(a["verts"]) == len(b["verts"])
    for x, y in zip(a["verts"], b["verts"]):
        assert np.array_equal(x, y)
    for k in ("coords", "metric", "lengths", "qualities"):
        assert np.array_equal(a[k].view(np.int64), b[k].view(np.int64)), k
    for (i1, d1), (i2, d2) in zip(a["class"], b["class"]):
        assert np.array_equal(i1, i2) and np.array_equal(d1, d2)


@pytest.mark.parametrize("dim,seed,nparts,n_range,h_range,cap", [
    (3, 11, 4, (2, 4), (0.08, 0.6), 2500),
    (2, 5, 5, (3, 10), (0.02, 0.4), 3000),
])
def test_joined_adapt_equals_separate_adapts(dim, seed, nparts, n_range, h_range, cap):
    alone = parts_of(dim, seed, nparts, n_range, h_range, cap)
    opts = O.default_opts(dim)
    opts.verbosity = 0
    joined = O.Mesh.join(parts_of(dim, seed, nparts, n_range, h_range, cap))
    npass = []
    for m in alone:
        _, passes = m.adapt(opts)
        npass.append(len(passes))
    _, jpasses = joined.adapt(opts)
    # parts need different pass counts; the joined run takes at least the most
    assert len(set(npass)) > 1
    assert len(jpasses) >= max(npass)
    for s, m in zip(split(joined, nparts), alone):
        same(s, snapshot(m))
