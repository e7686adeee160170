"""The oracle's prevent_coarsen_flip (src/Omega_h_coarsen_flip.cpp:68-150),
pinned by the reference's known answers for the most normal normal
(src/unit_math.cpp:171-195), by closed forms for it (unit vectors on a
circle of the sphere: the circle's axis), and by what the rule must leave
alone: on a box every exposed side is flat, so no collapse can turn a side
and adapt() must give the same mesh with the option on as with it off.  On
boxes whose boundary vertices are moved off their faces the option must
veto collapses (the adapted mesh changes)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O


def most_normal(vectors):
    N = np.ascontiguousarray(np.asarray(vectors, dtype=np.float64).ravel())
    out = np.zeros(3)
    dp = C.POINTER(C.c_double)
    rc = O.lib().om_most_normal(C.c_int(3), C.c_int(len(N) // 3), N.ctypes.data_as(dp),
                                out.ctypes.data_as(dp))
    return rc, out


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("vectors,expect", [
    ([[1, 0, 0], [0, 1, 0], [0, 0, 1]], [1, 1, 1]),
    ([[1, 0, 0], [0, 0, 1], unit([1, 1, 1]), [0, 1, 0]], [1, 1, 1]),
    ([[1, 0, 0], [0, 1, 0], [0, 1, 0]], [1, 1, 0]),
    ([[1, 0, 0], [1, 0, 0], [1, 0, 0]], [1, 0, 0]),
])
def test_most_normal_reference_known_answers(vectors, expect):
    rc, c = most_normal(vectors)
    assert rc == 0
    np.testing.assert_allclose(c, unit(expect), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_most_normal_of_a_circle_is_its_axis(seed):
    # k unit vectors at one angle from an axis (a circle on the sphere),
    # spread so that no half of the circle holds them all (consecutive gaps
    # below 180 degrees), plus vectors strictly inside: the smallest cap
    # holding them all is the circle's, centred on the axis
    rng = np.random.default_rng(seed)
    axis = unit(rng.normal(size=3))
    t1 = unit(np.cross(axis, rng.normal(size=3)))
    t2 = np.cross(axis, t1)
    ang = rng.uniform(0.1, 1.2)
    k = 3 + seed % 3
    phis = 2 * np.pi * (np.arange(k) + rng.uniform(-0.15, 0.15, k)) / k
    vs = [np.cos(ang) * axis + np.sin(ang) * (np.cos(p) * t1 + np.sin(p) * t2) for p in phis]
    for _ in range(3):
        p, r = rng.uniform(0, 2 * np.pi), rng.uniform(0, 0.8 * ang)
        vs.append(np.cos(r) * axis + np.sin(r) * (np.cos(p) * t1 + np.sin(p) * t2))
    order = rng.permutation(len(vs))
    rc, c = most_normal([vs[i] for i in order])
    assert rc == 0
    np.testing.assert_allclose(c, axis, atol=1e-9)


def test_most_normal_of_a_pair_is_the_bisector():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, b = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        if a @ b < -0.5:
            continue
        rc, c = most_normal([a, b])
        assert rc == 0
        np.testing.assert_allclose(c, unit(a + b), atol=1e-12)


def jittered_box(mod, dim, n, jitter, h, seed, ctx=None):
    """A box whose vertices inside each boundary face are moved normal to
    that face by up to `jitter` cell sizes, with the isotropic metric of
    edge length h."""
    m = mod.Mesh.box(dim, n) if ctx is None else mod.Mesh.box(ctx, dim, n)
    x = m.coords().reshape(-1, dim).copy()
    rng = np.random.default_rng(seed)
    for ax in range(dim):
        for side in (0.0, 1.0):
            inner = np.abs(x[:, ax] - side) < 1e-12
            for ax2 in range(dim):
                if ax2 != ax:
                    inner &= (x[:, ax2] > 1e-12) & (x[:, ax2] < 1 - 1e-12)
            x[inner, ax] += rng.uniform(-jitter, jitter, inner.sum()) / n
    m.set_coords(x.ravel())
    m.set_metric(np.full(m.nents(0), 1.0 / h ** 2), 1)
    return m


def adapted(dim, n, jitter, h, seed, flag):
    m = jittered_box(O, dim, n, jitter, h, seed)
    o = O.default_opts(dim)
    o.verbosity = 0
    o.should_prevent_coarsen_flip = flag
    m.adapt(o)
    return m


@pytest.mark.parametrize("dim,n,h", [(2, 16, 0.15), (2, 16, 0.4), (3, 6, 0.15), (3, 6, 0.4)])
def test_flat_boundaries_are_never_vetoed(dim, n, h):
    a = adapted(dim, n, 0.0, h, 0, 0)
    b = adapted(dim, n, 0.0, h, 0, 1)
    assert a.digests() == b.digests()


def test_curved_boundaries_are_vetoed():
    a = adapted(3, 8, 0.45, 0.4, 1, 0)
    b = adapted(3, 8, 0.45, 0.4, 1, 1)
    assert a.digests() != b.digests()
    assert b.nents(3) > a.nents(3)
