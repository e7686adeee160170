"""The oracle's metric intersection (all of Barral's degenerate cases) and
gradation limiting, pinned by closed forms and invariants.

src/Omega_h_metric_intersect.hpp:18-269, src/Omega_h_metric.cpp:156-220.
"""
import math

import numpy as np
import pytest

import oracle_lib as O


def inter(dim, a, b):
    return O.vector2symm(O.kat_intersect(dim, O.symm2vector(np.asarray(a, float)),
                                         O.symm2vector(np.asarray(b, float))))


def rot3(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q


@pytest.mark.parametrize("case,a,b,want", [
    # 3D case 2.a: two rank-1 metrics along the same axis -> the larger one
    ("2a", np.diag([4, 0, 0]), np.diag([9, 0, 0]), np.diag([9, 0, 0])),
    # 3D case 2.b: rank-1 along x and y -> diag(a, b, 0)
    ("2b", np.diag([4, 0, 0]), np.diag([0, 9, 0]), np.diag([4, 9, 0])),
    # 3D case 3.a: rank 2 and rank 1 in its plane -> elementwise max in the plane
    ("3a", np.diag([4, 5, 0]), np.diag([7, 0, 0]), np.diag([7, 5, 0])),
    # 3D case 4.b: two rank-2 metrics with different null axes -> elementwise max
    ("4b", np.diag([4, 5, 0]), np.diag([7, 0, 2]), np.diag([7, 5, 2])),
    # 3D case 4.a: same null axis -> intersection within the plane
    ("4a", np.diag([4, 5, 0]), np.diag([1, 8, 0]), np.diag([4, 8, 0])),
    # all null -> null
    ("null", np.zeros((3, 3)), np.zeros((3, 3)), np.zeros((3, 3))),
    # 2D, one full metric: the normal procedure, elementwise max for diagonals
    ("2d", np.diag([4, 1]), np.diag([0, 9]), np.diag([4, 9])),
])
def test_degenerate_cases_closed_forms(case, a, b, want):
    dim = len(want)
    got = inter(dim, a, b)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13), (case, got)
    # the intersection is symmetric in its arguments
    assert np.allclose(inter(dim, b, a), want, rtol=1e-13, atol=1e-13), case


def test_case5a_keeps_the_reference_rule():
    # 2D, both rank 1 with orthogonal null directions: the reference returns
    # max(l1, l2) along the first metric's range (src/Omega_h_metric_intersect.hpp:51-53)
    assert np.allclose(inter(2, np.diag([4, 0]), np.diag([0, 9])), np.diag([9, 0]))


def test_case5b_invariants():
    # 2D, both rank 1 with non-orthogonal null directions: P = [v1, v2] and
    # M = P^-T diag(v1.m2.v1, v2.m1.v2) P^-1 (src/Omega_h_metric_intersect.hpp:54-64)
    for t in (0.3, 0.7, 1.1):
        u = np.array([math.cos(t), math.sin(t)])
        m1 = np.diag([4.0, 0.0])
        m2 = 9.0 * np.outer(u, u)
        got = inter(2, m1, m2)
        v1 = np.array([0.0, 1.0])
        v2 = np.array([-math.sin(t), math.cos(t)])
        assert np.isclose(v1 @ got @ v1, v1 @ m2 @ v1, rtol=1e-12)
        assert np.isclose(v2 @ got @ v2, v2 @ m1 @ v2, rtol=1e-12)
        assert abs(v1 @ got @ v2) < 1e-12 * np.abs(got).max()


def test_case2b_contains_both_metrics():
    # rank-1 metrics along different axes: in the basis (u1 x u, u2 x u, u)
    # both are diagonal, so the result contains both exactly
    rng = np.random.default_rng(3)
    for _ in range(20):
        r = rot3(rng)
        u1, u2 = r[:, 0], (r[:, 0] + r[:, 1]) / math.sqrt(2)
        a = 3.0 * np.outer(u1, u1)
        b = 5.0 * np.outer(u2, u2)
        m = inter(3, a, b)
        for x in (a, b):
            w = np.linalg.eigvalsh(m - x)
            assert w.min() > -1e-9 * np.abs(m).max()


def test_case3_defining_properties():
    # m1 of rank 2 (null axis w1), m2 of rank 1 (range u2).  The reference
    # branches on the signed product u2 . w1 of the eigenvectors it computed
    # (src/Omega_h_metric_intersect.hpp:134): below 1e-10 it intersects
    # within m1's range plane (3.a), otherwise (3.b) it returns
    # M = P^-T diag(v2.m1.v2, w2.m1.w2, w1.m2.w1) P^-1, P = [v2, w2, w1], with
    # (v2, w2) an orthonormal basis of m2's null plane
    rng = np.random.default_rng(4)
    seen = set()
    for _ in range(40):
        r = rot3(rng)
        a = r @ np.diag([2.0, 6.0, 0.0]) @ r.T
        u = (r[:, 2] + r[:, 0]) / math.sqrt(2)
        b = 4.0 * np.outer(u, u)
        m = inter(3, a, b)
        q1, l1 = O.kat_decompose_eigen(a)
        q2, l2 = O.kat_decompose_eigen(b)
        w1 = q1[:, int(np.argmin(l1))]
        u2 = q2[:, int(np.argmax(l2))]
        scale = np.abs(m).max()
        if u2 @ w1 < 1e-10:
            seen.add("3a")
            # within the range plane of m1: M kills w1 and contains m1 there
            assert np.abs(m @ w1).max() < 1e-10 * scale
            w = np.linalg.eigvalsh(m - a)
            assert w.min() > -1e-9 * scale
        else:
            seen.add("3b")
            assert np.isclose(w1 @ m @ w1, w1 @ b @ w1, rtol=1e-12)
            qq, _ = np.linalg.qr(np.column_stack([u2, rng.normal(size=3), rng.normal(size=3)]))
            plane = qq[:, 1:]
            for k in range(2):
                assert abs(plane[:, k] @ m @ w1) < 1e-10 * scale
            assert np.isclose(np.trace(plane.T @ m @ plane), np.trace(plane.T @ a @ plane),
                              rtol=1e-12)
    assert seen == {"3a", "3b"}


def test_full_rank_intersection_contains_both():
    rng = np.random.default_rng(5)
    for _ in range(50):
        r1, r2 = rot3(rng), rot3(rng)
        a = r1 @ np.diag(np.exp(rng.uniform(-3, 3, 3))) @ r1.T
        b = r2 @ np.diag(np.exp(rng.uniform(-3, 3, 3))) @ r2.T
        m = inter(3, a, b)
        for x in (a, b):
            w = np.linalg.eigvalsh(m - x)
            assert w.min() > -1e-9 * np.abs(m).max()


def star_graph(m):
    a2ab, ab2b = m.star(0)
    return a2ab, ab2b


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 4)])
def test_gradation_of_isotropic_metric_is_a_fixed_point(dim, n):
    rate = 0.5
    m = O.Mesh.box(dim, n)
    c = m.coords().reshape(-1, dim)
    # a sharp isotropic spike at one corner
    h = np.where(np.linalg.norm(c, axis=1) < 0.3, 0.05, 0.3)
    m.set_metric(1.0 / (h * h), 1)
    before = m.metric().copy()
    steps = m.limit_metric_gradation(rate, 1e-3)
    assert steps > 1
    after = m.metric()
    # intersection only ever increases a metric
    assert np.all(after >= before * (1 - 1e-12))
    # converged: every edge satisfies the gradation bound, m_v >= m_u / (1 + rate d)^2,
    # with d the neighbour's metric distance, up to the tolerance
    a2ab, ab2b = star_graph(m)
    for v in range(m.nents(0)):
        for j in range(a2ab[v], a2ab[v + 1]):
            u = ab2b[j]
            d = math.sqrt(after[u] * float(np.sum((c[u] - c[v]) ** 2)))
            assert after[v] >= after[u] / (1.0 + rate * d) ** 2 * (1 - 2e-3)


def test_gradation_leaves_a_graded_metric_alone():
    m = O.Mesh.box(3, 3)
    m.set_metric(np.full(m.nents(0), 4.0), 1)
    assert m.limit_metric_gradation(1.0, 1e-6) == 1
    assert np.allclose(m.metric(), 4.0, rtol=1e-13)


def test_interpolate_metrics_closed_forms():
    rng = np.random.default_rng(8)
    a = np.exp(rng.uniform(-2, 2, 50))
    b = np.exp(rng.uniform(-2, 2, 50))
    for t in (0.0, 0.25, 0.5, 1.0):
        got = O.interpolate_metrics(1, a, b, t)
        assert np.allclose(got, a ** (1 - t) * b ** t, rtol=1e-13)
    # full 3D: t = 0 and t = 1 give the end points
    ms = []
    for _ in range(2):
        r = rot3(rng)
        ms.append(O.symm2vector(r @ np.diag(np.exp(rng.uniform(-3, 3, 3))) @ r.T))
    for t, want in ((0.0, ms[0]), (1.0, ms[1])):
        got = O.interpolate_metrics(3, ms[0], ms[1], t)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 4)])
def test_vertex_tags_follow_the_passes(dim, n):
    from omega_h_amd import workloads as W
    m = O.Mesh.box(dim, n)
    met, nc = W.iso_ring_metric_2d(m.coords()) if dim == 2 else W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    # coordinates as a linear tag and the metric as a metric tag are moved
    # exactly as the mesh moves its own coordinates and metric
    m.add_vertex_tag("xyz", dim, O.Mesh.LINEAR, m.coords())
    m.add_vertex_tag("m2", nc, O.Mesh.METRIC, met)
    m.adapt(O.default_opts(dim))
    assert np.array_equal(m.vertex_tag("xyz"), m.coords())
    assert np.array_equal(m.vertex_tag("m2"), m.metric())


def test_approach_metric_reaches_the_target():
    m = O.Mesh.box(2, 8)
    c = m.coords().reshape(-1, 2)
    h0 = 1.0 / 8
    m.set_metric(np.full(m.nents(0), 1.0 / h0 ** 2), 1)
    target = 1.0 / (0.03 + 0.1 * c[:, 0]) ** 2
    m.add_vertex_tag("target_metric", 1, O.Mesh.METRIC, target)
    # a copy of the target that is only transported, never consumed
    m.add_vertex_tag("copy", 1, O.Mesh.METRIC, target)
    opts = O.default_opts(2)
    steps = 0
    while True:
        r = m.approach_metric(opts)
        assert r >= 0
        if r == 0:
            break
        m.adapt(opts)
        steps += 1
        assert steps < 50
    assert steps >= 2
    assert m.vertex_tag("target_metric") is None
    # the final metric is the target itself, transported to the new vertices
    assert np.array_equal(m.metric(), m.vertex_tag("copy"))


@pytest.mark.parametrize("dim", [2, 3])
def test_element_implied_metric_gives_unit_edges(dim):
    rng = np.random.default_rng(12 + dim)
    for _ in range(50):
        pts = [rng.random(dim) for _ in range(dim + 1)]
        m = O.vector2symm(O.element_implied_metric(pts))
        lens = [math.sqrt((pts[j] - pts[i]) @ m @ (pts[j] - pts[i]))
                for i in range(dim + 1) for j in range(i + 1, dim + 1)]
        assert np.allclose(lens, 1.0, rtol=1e-9)


def test_implied_metric_of_the_regular_tet_is_the_identity():
    pts = [np.array([0.0, 0, 0]), np.array([1.0, 0, 0]),
           np.array([0.5, math.sqrt(3) / 2, 0]),
           np.array([0.5, math.sqrt(3) / 6, math.sqrt(2.0 / 3.0)])]
    assert np.allclose(O.element_implied_metric(pts), [1, 1, 1, 0, 0, 0], atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_implied_metrics_scale_with_the_mesh(dim):
    # the same box, twice as large: every implied metric is a quarter
    a = O.Mesh.box(dim, 3)
    b = O.Mesh.box(dim, 3, x=2.0, y=2.0, z=2.0)
    ma, mb = a.implied_metrics(), b.implied_metrics()
    assert np.allclose(mb, ma / 4.0, rtol=1e-12, atol=1e-12)
    # and they are positive definite
    nc = dim * (dim + 1) // 2
    for i in range(0, len(ma), nc):
        assert np.linalg.eigvalsh(O.vector2symm(ma[i:i + nc])).min() > 0


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 3)])
def test_inheritance_tags_follow_the_classification(dim, n):
    """A tag on every dimension that starts as a copy of class_id is
    transferred exactly as class_id is (src/Omega_h_transfer.cpp:20-34)."""
    from omega_h_amd import workloads as W
    m = O.Mesh.box(dim, n)
    met, nc = W.iso_ring_metric_2d(m.coords()) if dim == 2 else W.levelset_metric_3d(m.coords())
    m.set_metric(met, nc)
    m.add_inherit_tag("cid", 1, [m.classification(d)[0].astype(np.float64)
                                 for d in range(dim + 1)])
    m.adapt(O.default_opts(dim))
    for d in range(dim + 1):
        assert np.array_equal(m.inherit_tag("cid", d), m.classification(d)[0].astype(np.float64))


def center_vertex_mesh():
    m = O.Mesh.box(2, 4)
    m.set_metric(np.full(m.nents(0), 16.0), 1)
    c = m.coords().reshape(-1, 2)
    v = int(np.argmin(np.abs(c - 0.5).sum(1)))
    return m, c, v


def okay(m, opts):
    return (m.qualities().min() >= opts.min_quality_allowed and
            not m.lengths().max() > opts.max_length_allowed)


def test_warp_to_limit_halves_until_acceptable():
    # src/Omega_h_approach.cpp:36-81: the centre vertex of a 4x4 box is
    # pushed 3/8 towards its neighbour; the move halves k times until the
    # mesh is acceptable, and what was not moved stays in the tag (dyadic
    # values, so every sum is exact)
    m, c, v = center_vertex_mesh()
    w = np.zeros_like(c)
    w[v] = [0.375, 0.0]
    m.add_vertex_tag("warp", 2, 0, w.ravel())
    opts = O.default_opts(2)
    assert m.warp_to_limit(opts) == 1
    moved = m.coords().reshape(-1, 2) - c
    k = int(round(math.log2(0.375 / moved[v, 0])))
    assert k >= 1 and moved[v, 0] == 0.375 / 2 ** k and moved[v, 1] == 0.0
    assert np.count_nonzero(moved) == 1
    rem = m.vertex_tag("warp").reshape(-1, 2)
    assert rem[v, 0] == 0.375 - 0.375 / 2 ** k and np.count_nonzero(rem) == 1
    assert okay(m, opts)
    # one halving fewer would not have been acceptable
    p = O.Mesh.box(2, 4)
    p.set_metric(np.full(p.nents(0), 16.0), 1)
    q = c.copy()
    q[v, 0] += 0.375 / 2 ** (k - 1)
    p.set_coords(q.ravel())
    assert not okay(p, opts)


def test_warp_to_limit_translation_in_one_step():
    m, c, _ = center_vertex_mesh()
    w = np.tile([0.3, -0.7], (len(c), 1))
    m.add_vertex_tag("warp", 2, 0, w.ravel())
    assert m.warp_to_limit(O.default_opts(2)) == 1
    assert np.array_equal(m.coords(), (c + w).ravel())
    assert m.vertex_tag("warp") is None
    assert m.warp_to_limit(O.default_opts(2)) == 0


def test_warp_to_limit_stall_and_unacceptable_start():
    m, c, v = center_vertex_mesh()
    w = np.zeros_like(c)
    w[v] = [0.375, 0.0]
    m.add_vertex_tag("warp", 2, 0, w.ravel())
    opts = O.default_opts(2)
    # more halvings than allowed: the reference fails ...
    assert m.warp_to_limit(opts, max_niters=1) == -2
    # ... or, with exit_on_stall, drops the warp where the last try left it
    m, c, v = center_vertex_mesh()
    m.add_vertex_tag("warp", 2, 0, w.ravel())
    assert m.warp_to_limit(opts, exit_on_stall=True, max_niters=1) == 1
    assert m.vertex_tag("warp") is None
    assert m.coords().reshape(-1, 2)[v, 0] == 0.5 + 0.375 / 2
    opts.min_quality_allowed = 0.9
    m, c, v = center_vertex_mesh()
    m.add_vertex_tag("warp", 2, 0, w.ravel())
    assert m.warp_to_limit(opts) == -1


@pytest.mark.parametrize("dim,n", [(2, 8), (3, 3)])
def test_warp_and_adapt_loop_finishes(dim, n):
    # the reference's idiom, src/bend_test.cpp:55-65: do adapt while
    # warp_to_limit, here for the shear x += 2 z^2 (y^2 in 2D)
    m = O.Mesh.box(dim, n)
    m.set_metric(np.full(m.nents(0), float(n * n)), 1)
    c = m.coords().reshape(-1, dim)
    w = np.zeros_like(c)
    w[:, 0] = 2.0 * c[:, dim - 1] ** 2
    m.add_vertex_tag("warp", dim, 0, w.ravel())
    opts = O.default_opts(dim)
    steps = 0
    while True:
        m.adapt(opts)
        r = m.warp_to_limit(opts)
        assert r in (0, 1)
        if r == 0:
            break
        steps += 1
        assert steps < 10
    assert steps >= 3
    assert m.qualities().min() >= opts.min_quality_allowed
    # the shear moves along x only: the top face went to x in [2, 3]
    x = m.coords().reshape(-1, dim)
    top = x[:, dim - 1] == 1.0
    assert top.any() and np.all(x[top, 0] >= 2.0) and np.all(x[top, 0] <= 3.0)
    assert np.all((x[:, dim - 1] >= 0.0) & (x[:, dim - 1] <= 1.0))


@pytest.mark.parametrize("dim", [2, 3])
def test_max_eigenvalue_against_numpy(dim):
    # get_max_eigenvalues (src/Omega_h_eigen.cpp:7-40): the root of largest
    # magnitude of the characteristic polynomial
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(300):
        q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
        lam = np.exp(rng.uniform(0, 4, dim))
        if np.min(np.abs(np.subtract.outer(lam, lam))[~np.eye(dim, dtype=bool)]) < 0.01:
            continue
        a = q @ np.diag(lam) @ q.T
        got = O.max_eigenvalue(dim, O.symm2vector(a))
        assert abs(got - lam.max()) <= 1e-7 * lam.max()
        checked += 1
    assert checked > 200
    # the tolerance 5e-5 is absolute and the matrix is not normalised: small
    # eigenvalues closer than it merge into their mean (2D: |disc| < 5e-5)
    if dim == 2:
        a = np.diag([0.0233376, 0.0215307])
        assert O.max_eigenvalue(2, O.symm2vector(a)) == pytest.approx(0.02243415, 1e-9)
    # equal eigenvalues: the double and triple root branches
    assert O.max_eigenvalue(dim, O.symm2vector(np.eye(dim) * 3.0)) == pytest.approx(3.0, 1e-12)
    if dim == 3:
        a = np.diag([2.0, 5.0, 5.0])
        assert O.max_eigenvalue(3, O.symm2vector(a)) == pytest.approx(5.0, 1e-12)


@pytest.mark.parametrize("dim,n", [(2, 6), (3, 3)])
def test_implied_isos_keep_the_volume(dim, n):
    # get_implied_isos (src/Omega_h_metric.cpp:316-321): root<dim>(det) of
    # the implied metric
    m = O.Mesh.box(dim, n)
    full = m.implied_metrics().reshape(m.nents(0), -1)
    dets = np.array([np.linalg.det(O.vector2symm(v)) for v in full])
    want = np.cbrt(dets) if dim == 3 else np.sqrt(dets)
    assert np.allclose(m.implied_isos(), want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [3, 8])
def test_ill_metric_of_the_uniform_2d_box(n):
    # every triangle of the 2D box is a right isoceles triangle with legs 1/n
    # and the same diagonal, so each implies M = n^2 [[1, -1/2], [-1/2, 1]]
    # (unit edges (1/n, 0), (0, 1/n), (1/n, 1/n)); projection, smoothing and
    # grading keep a uniform metric, and every element is perfect in it
    opts = O.default_opts(2)
    m = O.Mesh.box(2, n)
    m.ill_metric(opts, 0)
    want = np.tile([n * n, n * n, -0.5 * n * n], m.nents(0))
    assert np.allclose(m.metric(), want, rtol=1e-12, atol=1e-12 * n * n)
    assert np.allclose(m.qualities(), 1.0, rtol=1e-12)
    # isotropic by length: the largest eigenvalue, 3/2 n^2
    m = O.Mesh.box(2, n)
    m.ill_metric(opts, 1)
    assert m.metric_ncomps == 1
    assert np.allclose(m.metric(), 1.5 * n * n, rtol=1e-12)


def jittered_box(dim, n, amp, seed):
    """A box whose interior vertices moved by up to amp / n per axis."""
    m = O.Mesh.box(dim, n)
    c = m.coords().reshape(-1, dim).copy()
    inner = np.all((c > 1e-12) & (c < 1 - 1e-12), axis=1)
    c[inner] += np.random.default_rng(seed).uniform(-amp, amp, (inner.sum(), dim)) / n
    m.set_coords(c.ravel())
    assert m.sizes().min() > 0
    return m


@pytest.mark.parametrize("dim,n,amp,seed,iso", [(2, 8, 0.4, 2, 0), (2, 8, 0.4, 2, 1),
                                                (3, 4, 0.35, 3, 1)])
def test_fix_raises_the_minimum_quality(dim, n, amp, seed, iso):
    # fix (src/Omega_h_fix.cpp:120-134): in the ill metric the mesh is below
    # the allowed quality; the adapt() calls of fix_for_given_metric raise
    # its minimum and stop once it improves by less than 1e-3
    opts = O.default_opts(dim)
    m = jittered_box(dim, n, amp, seed)
    m.ill_metric(opts, iso)
    q0 = m.qualities().min()
    m = jittered_box(dim, n, amp, seed)
    calls = m.fix(opts, iso)
    assert calls >= 2
    assert m.qualities().min() > q0
    assert m.sizes().min() > 0


@pytest.mark.parametrize("dim,n,amp,seed,nc", [(2, 8, 0.4, 2, 1), (2, 8, 0.4, 2, 3),
                                               (3, 4, 0.35, 3, 1), (3, 4, 0.35, 3, 6)])
def test_grade_fix_adapt_reaches_a_uniform_target(dim, n, amp, seed, nc):
    # grade_fix_adapt (src/Omega_h_fix.cpp:136-176) from a jittered box with
    # no metric: the implied metric is below the allowed quality (repaired),
    # then approach_metric + adapt reach the target, which grading left alone
    opts = O.default_opts(dim)
    m = jittered_box(dim, n, amp, seed)
    h2 = float((1.5 * n) ** 2)
    one = [h2] if nc == 1 else [h2] * dim + [0.0] * (nc - dim)
    steps = m.grade_fix_adapt(opts, np.tile(one, m.nents(0)), nc)
    assert steps >= 1
    assert m.vertex_tag("target_metric") is None
    assert m.metric_ncomps == nc
    assert np.allclose(m.metric(), np.tile(one, m.nents(0)), rtol=1e-12, atol=1e-12 * h2)
    assert m.qualities().min() >= opts.min_quality_allowed
    assert m.sizes().min() > 0
    assert abs(m.sizes().sum() - 1.0) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_barycentric_against_numpy(dim):
    # barycentric_from_global (src/Omega_h_shape.hpp:46-63)
    rng = np.random.default_rng(11)
    for _ in range(200):
        p = rng.normal(size=(dim + 1, dim))
        x = rng.normal(size=dim)
        got = O.barycentric(dim, x, p)
        lam = np.linalg.solve((p[1:] - p[0]).T, x - p[0])
        want = np.concatenate([[1.0 - lam.sum()], lam])
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9)
    p = rng.normal(size=(dim + 1, dim))
    for i in range(dim + 1):
        assert np.allclose(O.barycentric(dim, p[i], p), np.eye(dim + 1)[i], atol=1e-12)
    assert np.allclose(O.barycentric(dim, p.mean(0), p), 1.0 / (dim + 1), atol=1e-12)


def containing_elements(old, x):
    """Brute force: the old elements whose closed simplex holds x."""
    dim = old.dim
    ev = old.verts(dim).reshape(-1, dim + 1)
    c = old.coords().reshape(-1, dim)
    b = c[ev[:, 1:]] - c[ev[:, :1]]
    lam = np.linalg.solve(np.transpose(b, (0, 2, 1)), (x - c[ev[:, 0]])[:, :, None])[:, :, 0]
    bc = np.concatenate([1.0 - lam.sum(1, keepdims=True), lam], axis=1)
    return np.nonzero(bc.min(1) >= -1e-9)[0]


@pytest.mark.parametrize("dim,n,which", [(2, 6, "refine"), (2, 6, "coarsen"), (2, 6, "swap"),
                                         (3, 3, "refine"), (3, 3, "coarsen"),
                                         (3, 3, "swap")])
def test_pointwise_tags_come_from_the_element_under_the_centroid(dim, n, which):
    # transfer_pointwise_tmpl (src/Omega_h_transfer.cpp:515-583) picks, among
    # the old elements of the cavity, the one in which the product's centroid
    # lies deepest; refinement products take their parent's value.  Checked
    # against a search over every old element for one pass of each kind.
    # every coordinate off the box's faces jittered (boundary vertices slide
    # within their faces) so that centroids rarely fall on old faces
    m = O.Mesh.box(dim, n)
    c = m.coords().reshape(-1, dim).copy()
    rng = np.random.default_rng(5)
    free = (c > 1e-12) & (c < 1 - 1e-12)
    amp = 0.4 if which == "swap" else 0.3
    c[free] += rng.uniform(-amp, amp, free.sum()) / n
    m.set_coords(c.ravel())
    h = {"refine": 0.5 / n, "coarsen": 2.5 / n, "swap": 1.0 / n}[which]
    met = np.full(m.nents(0), h ** -2)
    m.set_metric(met, 1)
    vals = rng.normal(size=(m.nents(dim), 2))
    m.add_elem_tag("field", 2, vals.ravel())
    ev_old = m.verts(dim).reshape(-1, dim + 1)
    old = O.Mesh.from_elems(dim, ev_old.ravel(), m.nents(0))
    old.set_coords(m.coords())
    opts = O.default_opts(dim)
    did = {"refine": m.refine_by_size, "coarsen": m.coarsen_by_size,
           "swap": m.swap_edges}[which](opts)
    assert did
    new_vals = m.elem_tag("field").reshape(-1, 2)
    ev = m.verts(dim).reshape(-1, dim + 1)
    x = m.coords().reshape(-1, dim)[ev].mean(1)
    # same elements by their vertex positions (coarsening renumbers vertices)
    oc = old.coords().reshape(-1, dim)
    nc = m.coords().reshape(-1, dim)
    old_sets = {tuple(sorted(map(tuple, oc[v]))): i for i, v in enumerate(ev_old)}
    nprods = checked = 0
    for e in range(len(ev)):
        same = old_sets.get(tuple(sorted(map(tuple, nc[ev[e]]))))
        if same is not None and which != "refine":
            assert np.array_equal(new_vals[e], vals[same]), e
            continue
        nprods += 1
        hosts = containing_elements(old, x[e:e + 1].repeat(old.nents(dim), 0))
        if len(hosts) == 1:
            assert np.array_equal(new_vals[e], vals[hosts[0]]), e
            checked += 1
    assert nprods >= 4 and checked >= 0.7 * nprods
