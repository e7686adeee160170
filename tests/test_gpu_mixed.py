"""The FP32 estimates behind the mixed-precision form, against FP64.

The mixed form rejects a cavity only when an FP32 estimate falls short of its
threshold by more than the screening band (1e-3 relative) and evaluates every
other cavity again in FP64, so the adapted mesh is the FP64 one (the pass-by-
pass digest tests run both forms).  That is sound when the estimates are
within the band; here their relative error is measured on the 64^3
level-set input (normal eigen-lengths down to 0.002 against 0.05
tangentially) and on boxes under rotated anisotropic metrics, and must stay
below 2e-4, a fifth of the band.  Elements whose max-determinant vertex
metric is not clear in FP32 are excluded: the screen marks them and always
re-decides them in FP64.  Qualities below 1e-3 are compared absolutely.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402

TOL = 2e-4


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def errors(m):
    l32, q32, unsure = m.screen_estimates()
    l64, q64 = m.lengths(), m.qualities()
    lrel = np.abs(l32 - l64) / np.abs(l64)
    ok = unsure == 0
    qerr = np.abs(q32 - q64)[ok] / np.maximum(np.abs(q64[ok]), 1e-3)
    return lrel.max(), qerr.max(), float(np.mean(unsure))


def test_levelset_64_estimates(ctx):
    m = A.Mesh.box(ctx, 3, 64)
    met, nc = W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    lrel, qrel, frac = errors(m)
    print("level set 64^3: length %.2e, quality %.2e, unsure %.4f" % (lrel, qrel, frac))
    assert lrel < TOL and qrel < TOL


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rotated_metric_estimates(ctx, seed):
    rng = np.random.default_rng(seed)
    m = A.Mesh.box(ctx, 3, 16)
    coords = m.coords().reshape(-1, 3)
    # jitter the interior so that elements are not all alike
    inner = np.all((coords > 1e-9) & (coords < 1 - 1e-9), axis=1)
    coords[inner] += rng.uniform(-0.02, 0.02, size=(inner.sum(), 3))
    m.set_coords(coords.ravel())
    met, nc = W.rotated_constant_metric(m.nents(0), 3, rng)
    m.set_metric(met, nc)
    lrel, qrel, frac = errors(m)
    print("seed %d: length %.2e, quality %.2e, unsure %.4f" % (seed, lrel, qrel, frac))
    assert lrel < TOL and qrel < TOL


@pytest.mark.parametrize("n", [12, 20])
def test_coarsen_qualities_match_the_oracle(ctx, n):
    """osh_coarsen_qualities (the coarsen pass's dominant sweep) on every
    other edge of the box, both directions, bit for bit against the oracle;
    the mixed form gives the same values except -1 where the FP64 value is
    below the threshold."""
    import oracle_lib as O
    o = O.Mesh.box(3, n)
    met, nc = W.levelset_metric_3d(o.coords())
    o.set_metric(met, nc)
    g = A.Mesh.box(ctx, 3, n)
    g.set_metric(met, nc)
    cands = np.arange(0, g.nents(1), 2, dtype=np.int32)
    codes = np.full(len(cands), 3, dtype=np.int8)
    codes[::5] = 1
    codes[1::5] = 2
    want = o.coarsen_qualities(cands, codes)
    got = g.coarsen_qualities(cands, codes)
    assert np.array_equal(want.view(np.int64), got.view(np.int64))
    qmin = 0.2
    mixed = g.coarsen_qualities(cands, codes, precision=1, min_quality=qmin)
    kept = mixed != -1.0
    assert np.array_equal(mixed[kept].view(np.int64), want[kept].view(np.int64))
    assert np.all(want[~kept] < qmin)
    print("screen rejected %d of %d directions" % ((~kept & (want != -1.0)).sum(),
                                                     (want != -1.0).sum()))
