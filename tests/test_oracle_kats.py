"""Pins the oracle against the known answers in the reference's own unit
tests (src/unit_math.cpp, src/unit_mesh.cpp, src/unit_array_algs.cpp), against
glibc libm for the scalar transcendentals, and against closed forms and brute
force where the reference's tests leave a function open."""
import itertools
import math

import numpy as np
import pytest

import oracle_lib as O


def ulp_diff(a, b):
    if a == b:
        return 0
    ia = np.array([a], dtype=np.float64).view(np.int64)[0]
    ib = np.array([b], dtype=np.float64).view(np.int64)[0]
    return abs(int(ia) - int(ib))


@pytest.mark.parametrize("name,ref,lo,hi", [
    ("om_log", math.log, 1e-300, 1e300),
    ("om_exp", math.exp, -700.0, 700.0),
    ("om_cbrt", lambda x: math.copysign(abs(x) ** (1.0 / 3.0), x), -1e200, 1e200),
    ("om_acos", math.acos, -1.0, 1.0),
    ("om_cos", math.cos, -4.5, 4.5),
])
def test_transcendentals_within_one_ulp_of_libm(name, ref, lo, hi):
    rng = np.random.default_rng(7)
    f = getattr(O.lib(), name)
    if name == "om_log":
        xs = np.exp(rng.uniform(math.log(lo), math.log(hi), 20000))
    elif name == "om_cbrt":
        mag = np.exp(rng.uniform(-460, 460, 20000))
        xs = mag * rng.choice([-1.0, 1.0], 20000)
    else:
        xs = rng.uniform(lo, hi, 20000)
    worst = 0
    exact = 0
    for x in xs:
        x = float(x)
        got = f(x)
        if name == "om_cbrt":
            want = np.cbrt(x)
        else:
            want = ref(x)
        d = ulp_diff(got, want)
        worst = max(worst, d)
        exact += d == 0
    assert worst <= 1, (name, worst)
    assert exact > 0.85 * len(xs), (name, exact)


def test_transcendental_special_values():
    L = O.lib()
    assert L.om_log(1.0) == 0.0
    assert L.om_exp(0.0) == 1.0
    assert L.om_cbrt(27.0) == 3.0
    assert L.om_cbrt(-8.0) == -2.0
    assert L.om_acos(1.0) == 0.0
    assert L.om_cos(0.0) == 1.0
    assert math.isinf(L.om_log(0.0)) and L.om_log(0.0) < 0


def test_anisotropic_edge_length():
    # src/unit_math.cpp:11-15
    L = O.lib()
    assert abs(L.om_kat_aniso_length(1.0, 1.0) - 1.0) < 1e-10
    v = L.om_kat_aniso_length(1.0, 2.0)
    assert 1.0 < v < 1.5
    # closed form of Loseille's integral for h(t) = h0^(1-t) h1^t
    assert abs(v - 1.0 / math.log(2.0)) < 1e-15


def test_cubic_roots():
    # src/unit_math.cpp:37-53
    cases = [
        ([0, 0, 0], 1, [0], [3]),
        ([1.0, -1.5, -1.5], 3, [2, -1, 0.5], [1, 1, 1]),
        ([2.0, -3.0, 0], 2, [-2, 1], [1, 2]),
        ([-8, -6, 3], 3, [2, -4, -1], [1, 1, 1]),
    ]
    for coeffs, n, roots, mults in cases:
        got_n, got_r, got_m = O.kat_roots3(coeffs)
        assert got_n == n
        for i in range(n):
            assert got_m[i] == mults[i]
            assert abs(got_r[i] - roots[i]) <= 1e-10 * max(1.0, abs(roots[i]))


def _rotate(angle, axis):
    axis = np.asarray(axis, dtype=np.float64)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return math.cos(angle) * np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * np.outer(axis, axis)


def test_eigen_quadratic():
    # src/unit_math.cpp:108-117; matrix_2x2(a,b,c,d) is row-first
    m = np.array([[8.67958, -14.0234], [-1.04985, 2.25873]])
    q, l = O.kat_decompose_eigen(m)
    q_expect = np.array([[9.9192948778227130e-01, 8.6289280817702185e-01],
                         [-1.2679073810022995e-01, 5.0538698202107812e-01]])
    l_expect = np.array([1.0472083659357935e+01, 4.6622634064206342e-01])
    assert np.allclose(q, q_expect, rtol=1e-10, atol=1e-10)
    assert np.allclose(l, l_expect, rtol=1e-10, atol=1e-10)
    q, l = O.kat_decompose_eigen(np.eye(2))
    assert np.array_equal(q, np.eye(2)) and np.array_equal(l, [1, 1])
    q, l = O.kat_decompose_eigen(np.zeros((2, 2)))
    assert np.array_equal(q, np.eye(2)) and np.array_equal(l, [0, 0])


@pytest.mark.parametrize("h", [(1e3, 1, 1), (1, 1e3, 1e3), (1e-3, 1, 1),
                               (1, 1e-3, 1e-3), (1e-6, 1e-3, 1e-3)])
def test_eigen_metric_roundtrip(h):
    # src/unit_math.cpp:100-139
    q = _rotate(math.pi / 4, [0, 0, 1]) @ _rotate(math.pi / 4, [0, 1, 0])
    l = 1.0 / np.square(h)
    a = q @ np.diag(l) @ q.T
    qq, ll = O.kat_decompose_eigen(a)
    assert np.allclose(qq.T @ qq, np.eye(3), rtol=1e-8, atol=1e-8)
    rel = np.abs(ll - l) / np.maximum(np.abs(l), 1e-8)
    assert np.all(rel <= 1e-8)
    back = qq @ np.diag(ll) @ qq.T
    assert np.allclose(back, a, rtol=1e-8, atol=1e-8 * np.abs(a).max())


def test_eigen_nonsymmetric_cubic():
    # src/unit_math.cpp:127: eigenvalues of a non-symmetric matrix
    m = np.array([[-1, 3, -1], [-3, 5, -1], [-3, 3, 1]], dtype=np.float64)
    q, l = O.kat_decompose_eigen(m)
    assert np.allclose(sorted(l), [1, 2, 2], rtol=1e-8, atol=1e-8)


def test_log_exp_spd_are_inverse_and_match_scipy():
    import scipy.linalg
    rng = np.random.default_rng(3)
    for dim in (2, 3):
        for _ in range(50):
            a = rng.normal(size=(dim, dim))
            m = a @ a.T + dim * np.eye(dim)
            lm = O.kat_spd("om_kat_log_spd", m)
            assert np.allclose(lm, scipy.linalg.logm(m).real, rtol=1e-9, atol=1e-11)
            back = O.kat_spd("om_kat_exp_spd", lm)
            assert np.allclose(back, m, rtol=1e-9, atol=1e-11)


def test_quality_known_answers():
    # src/unit_mesh.cpp:317-359
    s3 = math.sqrt(3.0)
    r2 = 1.0 / math.sqrt(2.0)
    perfect_tri = [[1, 0], [0, s3], [-1, 0]]
    perfect_tet = [[1, 0, -r2], [-1, 0, -r2], [0, -1, r2], [0, 1, r2]]
    flat_tri = [[1, 0], [0, 0], [-1, 0]]
    flat_tet = [[1, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 1, 0]]
    inv_tri = [[1, 0], [-1, 0], [0, s3]]
    inv_tet = [[1, 0, -r2], [-1, 0, -r2], [0, 1, r2], [0, -1, r2]]
    id2 = [1, 1, 0] * 3
    id3 = [1, 1, 1, 0, 0, 0] * 4
    assert abs(O.kat_quality(2, 2, perfect_tri, id2) - 1.0) < 1e-10
    assert abs(O.kat_quality(3, 3, perfect_tet, id3) - 1.0) < 1e-10
    assert abs(O.kat_quality(2, 2, flat_tri, id2)) < 1e-10
    assert abs(O.kat_quality(3, 3, flat_tet, id3)) < 1e-10
    assert O.kat_quality(2, 2, inv_tri, id2) < 0
    assert O.kat_quality(3, 3, inv_tet, id3) < 0
    # anisotropic metric matched to a squashed element
    x_tri = [[p[0], p[1] / 2] for p in perfect_tri]
    x_tet = [[p[0], p[1], p[2] / 2] for p in perfect_tet]
    xm2 = [1, 4, 0] * 3
    xm3 = [1, 1, 4, 0, 0, 0] * 4
    assert abs(O.kat_quality(2, 2, x_tri, xm2) - 1.0) < 1e-10
    assert abs(O.kat_quality(3, 3, x_tet, xm3) - 1.0) < 1e-10
    # isotropic metric forms agree with the full tensor forms
    assert abs(O.kat_quality(3, 1, perfect_tet, [1] * 4) - 1.0) < 1e-10
    assert abs(O.kat_quality(2, 1, perfect_tri, [1] * 3) - 1.0) < 1e-10


def test_quality_matches_liu_joe_mean_ratio():
    # mean ratio = 12 (3V)^(2/3) / sum(l^2) in 3D, 4 sqrt(3) A / sum(l^2) in 2D
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = rng.normal(size=(4, 3))
        v = np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) / 6.0
        if v <= 0:
            p[[1, 2]] = p[[2, 1]]
            v = -v
        sl = sum(np.sum((p[i] - p[j]) ** 2) for i, j in itertools.combinations(range(4), 2))
        want = 12.0 * (3.0 * v) ** (2.0 / 3.0) / sl
        got = O.kat_quality(3, 3, p, [1, 1, 1, 0, 0, 0] * 4)
        assert abs(got - want) <= 1e-12 * max(1, want)


def test_edge_length_closed_forms():
    p0 = np.array([0.0, 0.0, 0.0])
    p1 = np.array([0.3, -0.2, 0.5])
    m = np.array([4.0, 9.0, 16.0, 0.0, 0.0, 0.0])
    want = math.sqrt(4 * 0.09 + 9 * 0.04 + 16 * 0.25)
    L = O.lib()
    import ctypes as C
    pp0 = p0.ctypes.data_as(C.POINTER(C.c_double))
    pp1 = p1.ctypes.data_as(C.POINTER(C.c_double))
    pm = m.ctypes.data_as(C.POINTER(C.c_double))
    got = L.om_kat_edge_length(3, 3, pp0, pp1, pm, pm)
    assert abs(got - want) < 1e-14


def test_average_metric_endpoints_and_commutes():
    # src/unit_math.cpp:271-280: interpolating between equal metrics returns them
    a = O.symm2vector(np.diag([1e4, 1.0]))
    out = O.kat_average_metric(2, np.concatenate([a, a]))
    assert np.allclose(out, a, rtol=1e-12)
    rng = np.random.default_rng(11)
    import scipy.linalg
    for _ in range(20):
        x = rng.normal(size=(3, 3))
        y = rng.normal(size=(3, 3))
        m1 = x @ x.T + np.eye(3)
        m2 = y @ y.T + np.eye(3)
        got = O.vector2symm(O.kat_average_metric(3, np.concatenate([O.symm2vector(m1), O.symm2vector(m2)])))
        want = scipy.linalg.expm(0.5 * (scipy.linalg.logm(m1) + scipy.linalg.logm(m2))).real
        assert np.allclose(got, want, rtol=1e-9, atol=1e-11)
    # isotropic metrics: exp(mean(log))
    got = O.kat_average_metric(1, np.array([4.0, 9.0]))
    assert abs(got[0] - 6.0) < 1e-14


def test_intersect_ortho_metrics():
    # src/unit_math.cpp:185-238 (non-degenerate cases)
    q = _rotate(math.pi / 4, [0, 0, 1]) @ _rotate(math.pi / 4, [0, 1, 0])
    cases = [((0.5, 1, 1), (1, 0.5, 1), (0.5, 0.5, 1)),
             ((1e-2, 1, 1), (1, 1, 1e-2), (1e-2, 1, 1e-2)),
             ((1e-2, 1e-2, 1), (1, 1, 1e-2), (1e-2, 1e-2, 1e-2)),
             ((1e-5, 1e-3, 1e-3), (1e-3, 1e-3, 1e-5), (1e-5, 1e-3, 1e-5))]
    for h1, h2, hi in cases:
        m1 = q @ np.diag(1.0 / np.square(h1)) @ q.T
        m2 = q @ np.diag(1.0 / np.square(h2)) @ q.T
        mi = O.vector2symm(O.kat_intersect(3, O.symm2vector(m1), O.symm2vector(m2)))
        for i in range(3):
            d = q[:, i]
            h = 1.0 / math.sqrt(d @ mi @ d)
            assert abs(h - hi[i]) <= 1e-3 * max(h, hi[i])
    # subset: the intersection with a contained metric is the larger one
    m1 = np.diag(1.0 / np.square([1.0, 2.0]))
    r = np.array([[math.cos(math.pi / 4), -math.sin(math.pi / 4)],
                  [math.sin(math.pi / 4), math.cos(math.pi / 4)]])
    m2 = r @ np.diag(1.0 / np.square([2.0, 3.0])) @ r.T
    for a, b in ((m1, m2), (m2, m1)):
        got = O.vector2symm(O.kat_intersect(2, O.symm2vector(a), O.symm2vector(b)))
        assert np.allclose(got, m1, rtol=1e-10, atol=1e-12)


def test_hilbert_skilling_example():
    # src/unit_mesh.cpp:211-236: 5 bits, 3 dims, (5,10,20) -> 7865
    X, Y = O.kat_hilbert([5, 10, 20], 5, 3)
    bits = ""
    for b in range(4, -1, -1):
        for i in range(3):
            bits += str((int(X[i]) >> b) & 1)
    assert bits == "001111010111001"
    assert int(Y[0]) * 1024 + int(Y[1]) * 32 + int(Y[2]) == 7865


def test_build_and_find_unique():
    # src/unit_mesh.cpp:197-209, 245-280
    m = O.Mesh.from_elems(2, [0, 1, 2, 2, 3, 0], 4)
    assert list(m.verts(1)) == [0, 1, 0, 2, 3, 0, 1, 2, 2, 3]
    m = O.Mesh.from_elems(2, [0, 1, 2], 3)
    assert list(m.verts(2)) == [0, 1, 2]
    assert list(m.down(2)[0]) == [0, 2, 1]
    assert list(m.verts(1)) == [0, 1, 2, 0, 1, 2]
    m = O.Mesh.from_elems(3, [0, 1, 2, 3], 4)
    assert list(m.verts(3)) == [0, 1, 2, 3]


def test_star_graphs():
    # src/unit_mesh.cpp:282-305
    m = O.Mesh.from_elems(2, [0, 1, 2], 3)
    a, b = m.star(0)
    assert list(a) == [0, 2, 4, 6] and list(b) == [1, 2, 0, 2, 0, 1]
    a, b = m.star(1)
    assert list(a) == [0, 2, 4, 6] and list(b) == [2, 1, 0, 2, 1, 0]
    m = O.Mesh.from_elems(3, [0, 1, 2, 3], 4)
    a, b = m.star(0)
    assert list(a) == [0, 3, 6, 9, 12]
    assert list(b) == [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2]
    a, b = m.star(1)
    assert list(a) == list(range(0, 31, 5))
    assert list(b) == [1, 3, 4, 2, 5, 3, 0, 2, 5, 4, 0, 4, 5, 1, 3, 0, 1, 5,
                       4, 2, 2, 0, 3, 5, 1, 1, 2, 4, 3, 0]


def brute_dual(m):
    """Elements across each element's sides, in side order, found by search."""
    dim = m.dim
    down = m.down(dim)[0].reshape(-1, dim + 1)
    owners = {}
    for e, sides in enumerate(down):
        for s in sides:
            owners.setdefault(int(s), []).append(e)
    a2ab, ab2b = [0], []
    for e, sides in enumerate(down):
        for s in sides:
            ab2b += [x for x in owners[int(s)] if x != e]
        a2ab.append(len(ab2b))
    return a2ab, ab2b


@pytest.mark.parametrize("dim,n", [(2, 1), (2, 4), (3, 1), (3, 3)])
def test_dual_against_search_and_counts(dim, n):
    # Mesh::ask_dual, src/Omega_h_mesh.cpp:260 (src/Omega_h_adj.cpp:591-657)
    m = O.Mesh.box(dim, n)
    a2ab, ab2b = m.dual()
    want_a, want_b = brute_dual(m)
    assert list(a2ab) == want_a and list(ab2b) == want_b
    # every interior side appears twice, boundary sides never
    nbdry = 4 * n if dim == 2 else 12 * n * n
    assert len(ab2b) == 2 * (m.nents(dim - 1) - nbdry)
    pairs = set(zip(np.repeat(np.arange(m.nents(dim)), np.diff(a2ab)), ab2b))
    assert all((b, a) in pairs for a, b in pairs)


def test_dual_single_simplex_is_empty():
    for dim, ev2v in ((2, [0, 1, 2]), (3, [0, 1, 2, 3])):
        a2ab, ab2b = O.Mesh.from_elems(dim, ev2v, dim + 1).dual()
        assert list(a2ab) == [0, 0] and len(ab2b) == 0


@pytest.mark.parametrize("dim,n", [(2, 1), (2, 7), (3, 1), (3, 4)])
def test_sizes_of_box(dim, n):
    # Mesh::ask_sizes, src/Omega_h_mesh.cpp:390-396: the box splits each
    # square into 2 triangles and each cube into 6 tets of equal size
    m = O.Mesh.box(dim, n)
    s = m.sizes()
    unit = 1.0 / (2 * n * n) if dim == 2 else 1.0 / (6 * n ** 3)
    assert np.allclose(s, unit, rtol=1e-12, atol=0)
    assert abs(s.sum() - 1.0) < 1e-12
    # against the determinant of the edge basis
    c = m.coords().reshape(-1, dim)
    ev = m.verts(dim).reshape(-1, dim + 1)
    b = c[ev[:, 1:]] - c[ev[:, :1]]
    fact = 2.0 if dim == 2 else 6.0
    assert np.allclose(s, np.linalg.det(b) / fact, rtol=1e-12, atol=1e-15)


def test_refine_qualities_unit_square():
    # src/unit_mesh.cpp:382-391
    m = O.Mesh.box(2, 1, internal=True)
    m.set_metric(np.tile([1.0, 1.0, 0.0], m.nents(0)), 3)
    q = m.refine_qualities(np.arange(m.nents(1)))
    want = [0.494872, 0.494872, 0.866025, 0.494872, 0.494872]
    assert np.allclose(q, want, atol=1e-4 * 0.866025)


def test_swap3d_find_loop_unit_cube():
    # src/unit_mesh.cpp:424-447
    m = O.Mesh.box(3, 1, internal=True)
    eev, loop = m.find_loop(6)
    assert list(eev) == [7, 0]
    assert list(loop) == [2, 3, 1, 5, 4, 6]


def test_box_counts_and_euler():
    for n in (1, 2, 3, 5):
        m = O.Mesh.box(3, n)
        V, E, F, T = (m.nents(d) for d in range(4))
        assert T == 6 * n ** 3
        assert V == (n + 1) ** 3
        assert V - E + F - T == 1
    m = O.Mesh.box(2, 32)
    assert m.nents(2) == 2048 and m.nents(0) == 33 * 33
    assert m.nents(0) - m.nents(1) + m.nents(2) == 1


def test_box_elements_positive_and_classified():
    m = O.Mesh.box(3, 4)
    q = m.qualities_with_identity() if hasattr(m, "qualities_with_identity") else None
    m.set_metric(np.tile([1.0], m.nents(0)), 1)
    q = m.qualities()
    assert np.all(q > 0)
    ids, dims = m.classification(3)
    assert np.all(dims == 3) and np.all(ids == 13)
    ids, dims = m.classification(0)
    # the 8 corners are classified on model vertices
    assert np.sum(dims == 0) == 8


def _greedy_indset(a2ab, ab2b, q, cands):
    order = sorted(np.nonzero(cands)[0], key=lambda v: (q[v], v), reverse=True)
    state = np.zeros(len(q), dtype=np.int8)
    blocked = np.zeros(len(q), dtype=bool)
    for v in order:
        if blocked[v]:
            continue
        state[v] = 1
        for j in range(a2ab[v], a2ab[v + 1]):
            blocked[ab2b[j]] = True
    return state


def test_indset_equals_lexicographic_greedy():
    m = O.Mesh.box(3, 3)
    rng = np.random.default_rng(1)
    for d in (0, 1):
        a2ab, ab2b = m.star(d)
        n = m.nents(d)
        for trial in range(5):
            q = rng.integers(0, 4, n).astype(np.float64)  # many ties
            c = (rng.random(n) < 0.6).astype(np.int8)
            got = m.indset(d, q, c)
            want = _greedy_indset(a2ab, ab2b, q, c)
            assert np.array_equal(got, want)


def test_swap_tables_are_triangulations():
    """Every table entry is a triangulation of the n-gon with the loop's
    orientation, the triangulations are distinct, there are Catalan(n-2) of
    them, and the listed edges are exactly each one's interior diagonals."""
    import ctypes as C
    L = O.lib()
    sizes = (C.c_int * 8).in_dll(L, "om_swap_mesh_sizes")
    counts = (C.c_int * 8).in_dll(L, "om_swap_mesh_counts")
    nedges = (C.c_int * 8).in_dll(L, "om_swap_nedges")
    L.om_swap_triangle.restype = C.POINTER(C.c_int)
    L.om_swap_unique_edge.restype = C.POINTER(C.c_int)
    catalan = {3: 1, 4: 2, 5: 5, 6: 14, 7: 42}
    for n in range(3, 8):
        assert counts[n] == catalan[n] and sizes[n] == n - 2 and nedges[n] == max(0, n - 3)
        seen = set()
        for mesh in range(counts[n]):
            tris = []
            for t in range(sizes[n]):
                ut = L.om_swap_mesh_tri(n, mesh, t)
                p = L.om_swap_triangle(n, ut)
                tri = (p[0], p[1], p[2])
                # counterclockwise in loop order: a cyclic rotation of an increasing triple
                assert sorted(tri) in ([tri[0], tri[1], tri[2]], [tri[1], tri[2], tri[0]],
                                       [tri[2], tri[0], tri[1]])
                tris.append(frozenset(tri))
            key = frozenset(tris)
            assert key not in seen
            seen.add(key)
            # area check: triangles of a convex n-gon triangulation
            tri_edges = set()
            for t in tris:
                for a, b in itertools.combinations(sorted(t), 2):
                    tri_edges.add((a, b))
            boundary = {(i, i + 1) for i in range(n - 1)} | {(0, n - 1)}
            interior = tri_edges - boundary
            assert len(interior) == n - 3
            listed = set()
            for e in range(nedges[n]):
                ue = L.om_swap_mesh_edge(n, mesh, e)
                p = L.om_swap_unique_edge(n, ue)
                listed.add(tuple(sorted((p[0], p[1]))))
            assert listed == interior
