"""The oracle's conservative transfer (src/Omega_h_conserve.cpp) and its
simplex intersection (src/r3d.hpp), pinned by what the mathematics fixes:

* intersection sizes: a simplex with itself is its size, disjoint simplices
  give zero, the pieces of a simplex cut by a partition of space add up to
  its size (the six tetrahedra of a cube, the two triangles of a square),
  and in 2D the size equals an independent polygon clip
  (Sutherland-Hodgman) of the two triangles;
* conservation, as the reference's own tests check it
  (src/2d_conserve_test.cpp, src/3d_conserve_test.cpp): after adapt() the
  total mass (density times element size) and the total momentum (element
  mean velocity times mass) equal their values before, to round-off;
* a density that is constant stays that constant."""
import numpy as np
import pytest

import oracle_lib as O


def simplex_size(p):
    p = np.asarray(p, dtype=np.float64)
    d = len(p) - 1
    m = (p[1:] - p[0]).T
    return np.linalg.det(m) / (2.0 if d == 2 else 6.0)


def random_simplex(rng, dim, lo=0.0, hi=1.0):
    while True:
        p = rng.uniform(lo, hi, size=(dim + 1, dim))
        if simplex_size(p) > 1e-3:
            return p
        p[[0, 1]] = p[[1, 0]]
        if simplex_size(p) > 1e-3:
            return p


CUBE_TETS = [[0, 1, 3, 7], [0, 1, 5, 7], [0, 4, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7],
             [0, 4, 6, 7]]


def cube_corner(i):
    return [float(i & 1), float((i >> 1) & 1), float((i >> 2) & 1)]


def positive(p):
    p = np.array(p, dtype=np.float64)
    if simplex_size(p) < 0:
        p[[0, 1]] = p[[1, 0]]
    return p


@pytest.mark.parametrize("dim", [2, 3])
def test_intersection_with_itself_and_disjoint(dim):
    rng = np.random.default_rng(dim)
    for _ in range(50):
        a = random_simplex(rng, dim)
        assert abs(O.intersect_size(dim, a, a) - simplex_size(a)) <= 1e-13 * simplex_size(a)
        b = random_simplex(rng, dim, 2.0, 3.0)
        assert O.intersect_size(dim, a, b) == 0.0


def test_intersection_pieces_add_up_3d():
    rng = np.random.default_rng(5)
    parts = [positive([cube_corner(i) for i in t]) for t in CUBE_TETS]
    assert abs(sum(simplex_size(p) for p in parts) - 1.0) < 1e-15
    for _ in range(40):
        a = random_simplex(rng, 3, 0.05, 0.95)
        total = sum(O.intersect_size(3, a, p) for p in parts)
        assert abs(total - simplex_size(a)) <= 1e-13
        # and the other way round: the cube's pieces cut by the simplex
        back = sum(O.intersect_size(3, p, a) for p in parts)
        assert abs(back - simplex_size(a)) <= 1e-13


def test_intersection_pieces_add_up_2d():
    rng = np.random.default_rng(6)
    parts = [positive([[0, 0], [1, 0], [1, 1]]), positive([[0, 0], [1, 1], [0, 1]])]
    for _ in range(40):
        a = random_simplex(rng, 2, 0.05, 0.95)
        total = sum(O.intersect_size(2, a, p) for p in parts)
        assert abs(total - simplex_size(a)) <= 1e-14


def clip_polygon(poly, a, b):
    """Sutherland-Hodgman: the part of the polygon left of the line a -> b."""
    def side(p):
        return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0])
    out = []
    for i in range(len(poly)):
        p, q = poly[i], poly[(i + 1) % len(poly)]
        sp, sq = side(p), side(q)
        if sp >= 0:
            out.append(p)
        if (sp >= 0) != (sq >= 0):
            t = sp / (sp - sq)
            out.append(p + t * (q - p))
    return out


def polygon_area(poly):
    if len(poly) < 3:
        return 0.0
    x = np.array([p[0] for p in poly])
    y = np.array([p[1] for p in poly])
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def test_intersection_matches_polygon_clipping_2d():
    rng = np.random.default_rng(7)
    for _ in range(300):
        a = random_simplex(rng, 2)
        b = random_simplex(rng, 2)
        poly = [np.array(p) for p in a]
        for i in range(3):
            poly = clip_polygon(poly, b[i], b[(i + 1) % 3])
            if not poly:
                break
        assert abs(O.intersect_size(2, a, b) - polygon_area(poly)) <= 1e-14


def conserve_box(dim, nx, diffuse_mass=(1, 0.9, 0.0), diffuse_momentum=(1, 0.02, 1e-6)):
    """The setup of src/2d_conserve_test.cpp and src/3d_conserve_test.cpp:
    implied isotropic metric scaled to length 1.3, density 1e-4 above
    y = 0.5 and 1 below, velocity (sign(x - 0.3) sqrt|x - 0.3|, 0, 0)."""
    m = O.Mesh.box(dim, nx)
    m.set_metric(m.implied_isos() * (1.0 / 1.3 ** 2), 1)
    c = m.coords().reshape(-1, dim)
    ev = m.verts(dim).reshape(-1, dim + 1)
    cen = c[ev].mean(axis=1)
    m.add_elem_tag("density", 1, np.where(cen[:, 1] > 0.5, 1e-4, 1.0))
    x2 = c[:, 0] - 0.3
    vel = np.zeros((len(c), dim))
    vel[:, 0] = np.sign(x2) * np.sqrt(np.abs(x2))
    m.add_vertex_tag("velocity", dim, O.Mesh.LINEAR, vel.ravel())
    m.set_elem_transfer("density", O.Mesh.CONSERVE, diffuse_mass)
    m.set_momentum_velocity("velocity", "density", diffuse_momentum)
    return m


def totals(m):
    dim = m.dim
    mass = m.elem_tag("density") * m.sizes()
    vel = m.vertex_tag("velocity").reshape(-1, dim)
    ev = m.verts(dim).reshape(-1, dim + 1)
    evel = vel[ev].sum(axis=1) / (dim + 1)
    return mass.sum(), (evel * mass[:, None]).sum(axis=0)


@pytest.mark.parametrize("dim,nx", [(2, 10), (3, 6)])
@pytest.mark.parametrize("diffuse", ["diffused", "none"])
def test_adapt_conserves_mass_and_momentum(dim, nx, diffuse):
    if diffuse == "none":
        m = conserve_box(dim, nx, (0, 0.0, 0.0), (0, 0.0, 0.0))
    else:
        m = conserve_box(dim, nx)
    m0, p0 = totals(m)
    assert m0 == pytest.approx(0.5 + 0.5e-4, rel=1e-12)
    ne0 = m.nents(dim)
    o = O.default_opts(dim)
    o.verbosity = 0
    m.adapt(o)
    assert m.nents(dim) != ne0
    m1, p1 = totals(m)
    assert abs(m1 - m0) <= 1e-12 * m0
    assert np.all(np.abs(p1 - p0) <= 1e-12 * np.abs(p0).max())
    d = m.elem_tag("density")
    assert d.min() > 0.0


def test_constant_density_stays_constant():
    m = O.Mesh.box(3, 6)
    m.set_metric(m.implied_isos() * (1.0 / 1.6 ** 2), 1)
    m.add_elem_tag("rho", 1, np.full(m.nents(3), 2.5))
    m.set_elem_transfer("rho", O.Mesh.DENSITY)
    o = O.default_opts(3)
    o.verbosity = 0
    m.adapt(o)
    np.testing.assert_allclose(m.elem_tag("rho"), 2.5, rtol=1e-13)
