"""adapt() with AdaptOpts::should_prevent_coarsen_flip on the GPU against the
oracle, bit for bit (the four mesh digests), on boxes whose boundary
vertices are moved off their faces so that the flip check vetoes
collapses (src/Omega_h_coarsen_flip.cpp:68-150), and on flat boxes, where it
must veto nothing."""
import pytest

import oracle_lib as O
from test_oracle_coarsen_flip import jittered_box

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


@pytest.mark.parametrize("dim,n,jitter,h,seed", [
    (3, 8, 0.45, 0.4, 1),
    (3, 8, 0.2, 0.4, 1),
    (3, 6, 0.0, 0.4, 0),
    (2, 16, 0.45, 0.2, 2),
])
def test_flip_prevention_matches_oracle(ctx, dim, n, jitter, h, seed):
    o = jittered_box(O, dim, n, jitter, h, seed)
    g = jittered_box(A, dim, n, jitter, h, seed, ctx=ctx)
    oo = O.default_opts(dim)
    oo.verbosity = 0
    oo.should_prevent_coarsen_flip = 1
    go = ctx.default_opts(dim)
    go.verbosity = 0
    go.should_prevent_coarsen_flip = 1
    assert bool(o.adapt(oo)) == bool(g.adapt(go)[0])
    assert [g.nents(d) for d in range(dim + 1)] == [o.nents(d) for d in range(dim + 1)]
    assert g.digests() == o.digests()
