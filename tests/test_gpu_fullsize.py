"""The 64^3 level-set workload (1,572,864 tets, the benchmark's input) at full size.

The oracle runs the first two mesh-changing passes of adapt() on the same input
(about 20 s of host time) and the device mesh must match it bit for bit.  The
whole adapt() (226 passes) is compared with the oracle's own whole run, pass by
pass, through the digests the oracle recorded in
tests/golden/adapt_levelset_3d_64.json (tests/golden/make_adapt_digests.py):
every pass's kind and element counts, and the vertex lists, classification,
coordinate, metric, length and quality bits after it.  Properties that need no
oracle are checked too: a second run on a fresh copy gives the same mesh and
pass sequence, every element keeps a positive quality, the Euler characteristic
of the ball is 1, every face has one or two tets and a face with one tet is
classified on the boundary.
"""
import numpy as np
import pytest

import adapt_fixtures as F
import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402

from test_gpu_parity import compare_meshes  # noqa: E402

N = 64


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def levelset_box(mod, ctx=None):
    m = mod.Mesh.box(3, N) if ctx is None else mod.Mesh.box(ctx, 3, N)
    met, nc = W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    return m


def test_first_passes_match_the_oracle(ctx):
    o = levelset_box(O)
    g = levelset_box(A, ctx)
    oo, go = O.default_opts(3), ctx.default_opts(3)
    # adapt() starts with refine_by_size (nothing to refine here) and
    # coarsen_by_size, in this order, until neither changes the mesh
    done = 0
    for name in ("refine_by_size", "coarsen_by_size", "refine_by_size", "coarsen_by_size"):
        changed = getattr(o, name)(oo)
        assert getattr(g, name)(go) == changed, name
        done += changed
    assert done == 2
    compare_meshes(o, g)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_full_adapt_matches_oracle_pass_by_pass(ctx, precision):
    fx = F.load("levelset_3d_64")
    m = levelset_box(A, ctx)
    opts = ctx.default_opts(3)
    opts.verbosity = 0
    opts.precision = precision
    did, passes, digests = m.adapt_digests(opts)
    assert did == fx["adapted"]
    msg = F.first_difference(fx["passes"], F.rows_of(passes, digests))
    assert msg is None, msg
    assert [m.nents(d) for d in range(4)] == fx["final_nents"]


def test_full_adapt_properties(ctx):
    fx = F.load("levelset_3d_64")
    base = levelset_box(A, ctx)
    opts = ctx.default_opts(3)
    runs = []
    for _ in range(2):
        m = base.copy()
        changed, passes = m.adapt(opts)
        assert changed
        runs.append((m.digest(), [p[:3] for p in passes], m))
    assert runs[0][0] == runs[1][0]
    assert runs[0][1] == runs[1][1]
    # osh_adapt takes the same passes as the oracle's run
    assert [list(p) for p in runs[0][1]] == [r[:3] for r in fx["passes"]]
    m = runs[0][2]
    nt = m.nents(3)
    assert nt == fx["final_nents"][3]
    assert m.qualities().min() > 0.0
    assert m.nents(0) - m.nents(1) + m.nents(2) - nt == 1
    a2ab, _, _ = m.up(2, 3)
    deg = np.diff(a2ab)
    assert deg.min() >= 1 and deg.max() <= 2
    _, fdim = m.classification(2)
    assert np.all(fdim[deg == 1] < 3)
    assert np.all(fdim[deg == 2] == 3)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_silent_adapt_ends_where_the_oracle_does(ctx, precision):
    """adapt() with no per-pass hook and no reports (as bench times it): the
    same passes, and a final mesh whose four digests equal those the oracle
    recorded after its last pass."""
    fx = F.load("levelset_3d_64")
    m = levelset_box(A, ctx)
    opts = ctx.default_opts(3)
    opts.verbosity = 0
    opts.precision = precision
    did, passes = m.adapt(opts)
    assert did == fx["adapted"]
    assert [list(p[:3]) for p in passes] == [r[:3] for r in fx["passes"]]
    assert ["%016x" % d for d in m.digests()] == fx["passes"][-1][3:7]
