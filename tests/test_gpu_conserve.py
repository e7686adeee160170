"""Conservative and momentum transfer on the device (osh_conserve.hip) against
the oracle (oracle/om_conserve.c), bit for bit.

The setup is the reference's own conservation test
(src/2d_conserve_test.cpp, src/3d_conserve_test.cpp): an implied isotropic
metric scaled to length 1.3, a conserved density 1e-4 above y = 0.5 and 1
below, and a momentum velocity (sign(x - 0.3) sqrt|x - 0.3|, 0, 0).  Besides
the plain box, a two-material box (elements left and right of x = 0.5 in
regions 1 and 2, the interface classified on a model face) exercises the
cavities that split by element class, and one case fixes the first velocity
component on the boundary x = 0 ("momentum_velocity_fixed").  After
adapt() the device's mesh, densities and velocities must equal the
oracle's bit for bit, and mass and momentum must equal their values before
to round-off.  A per-pass run with an intersection-transferred density
compares the densities after every pass.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")

from test_gpu_parity import assert_same, compare_meshes  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def two_materials(g):
    """Regions 1 (x < 0.5) and 2 (x > 0.5); entities on the plane x = 0.5 go
    one model dimension down (the interior onto the interface, a box face
    onto the interface's edge, a box edge onto a model vertex) with new ids."""
    dim = g.dim
    x = g.coords().reshape(-1, dim)[:, 0]
    for d in range(dim + 1):
        ids, dims = g.classification(d)
        ev = g.verts(d).reshape(-1, d + 1) if d > 0 else np.arange(g.nents(0))[:, None]
        if d == dim:
            cen = x[ev].mean(axis=1)
            ids = np.where(cen < 0.5, 1, 2).astype(np.int32)
        else:
            on = np.all(x[ev] == 0.5, axis=1)
            newd = np.where(on, np.maximum(dims.astype(np.int32) - 1, d), dims)
            ids = np.where(newd != dims, 100 * (dim - dims.astype(np.int32) + 1) + ids, ids)
            ids = ids.astype(np.int32)
            dims = newd.astype(np.int8)
        g.set_class(d, ids, dims)


def setup(m, iso, dim, diffuse_mass, diffuse_momentum, fixed):
    m.set_metric(iso, 1)
    c = m.coords().reshape(-1, dim)
    ev = m.verts(dim).reshape(-1, dim + 1)
    cen = c[ev].mean(axis=1)
    m.add_elem_tag("density", 1, np.where(cen[:, 1] > 0.5, 1e-4, 1.0))
    x2 = c[:, 0] - 0.3
    vel = np.zeros((len(c), dim))
    vel[:, 0] = np.sign(x2) * np.sqrt(np.abs(x2))
    m.add_vertex_tag("velocity", dim, m.LINEAR, vel.ravel())
    m.set_elem_transfer("density", m.CONSERVE, diffuse_mass)
    m.set_momentum_velocity("velocity", "density", diffuse_momentum)
    if fixed:
        per_dim = [np.where(c[:, 0] == 0.0, 1.0, 0.0)]
        per_dim += [np.zeros(m.nents(d)) for d in range(1, dim + 1)]
        m.add_inherit_tag("momentum_velocity_fixed", 1, per_dim)


def totals(m):
    dim = m.dim
    mass = m.elem_tag("density") * m.sizes()
    vel = m.vertex_tag("velocity").reshape(-1, dim)
    ev = m.verts(dim).reshape(-1, dim + 1)
    evel = vel[ev].sum(axis=1) / (dim + 1)
    return mass.sum(), (evel * mass[:, None]).sum(axis=0)


DIFFUSED = ((1, 0.9, 0.0), (1, 0.02, 1e-6))
ABSOLUTE = ((2, 0.5, 0.0), (2, 0.05, 1e-6))
NONE = ((0, 0.0, 0.0), (0, 0.0, 0.0))

CASES = {
    "2d_diffused": (2, 10, DIFFUSED, False, False),
    "3d_diffused": (3, 6, DIFFUSED, False, False),
    "2d_none": (2, 10, NONE, False, False),
    "3d_none": (3, 6, NONE, False, False),
    "2d_absolute_fixed": (2, 10, ABSOLUTE, False, True),
    "2d_two_materials": (2, 10, DIFFUSED, True, False),
    "3d_two_materials": (3, 6, DIFFUSED, True, True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_adapt_conserves_like_the_oracle(ctx, name):
    dim, n, (dm, dp), materials, fixed = CASES[name]
    g = A.Mesh.box(ctx, dim, n)
    if materials:
        two_materials(g)
    g.set_metric(np.ones(g.nents(0)), 1)  # replaced below; copy_of carries a metric
    o = O.Mesh.copy_of(g)
    iso = o.implied_isos() * (1.0 / 1.3 ** 2)
    for m in (o, g):
        setup(m, iso, dim, dm, dp, fixed)
    m0, p0 = totals(g)
    ne0 = g.nents(dim)
    oo, go = O.default_opts(dim), ctx.default_opts(dim)
    oo.verbosity = go.verbosity = 0
    ro, po = o.adapt(oo)
    rg, pg = g.adapt(go)
    assert ro == rg
    assert [p[:3] for p in po] == [p[:3] for p in pg]
    assert g.nents(dim) != ne0
    compare_meshes(o, g)
    assert_same(o.elem_tag("density"), g.elem_tag("density"), "density")
    assert_same(o.vertex_tag("velocity"), g.vertex_tag("velocity"), "velocity")
    m1, p1 = totals(g)
    assert abs(m1 - m0) <= 1e-12 * m0
    if not fixed:
        assert np.all(np.abs(p1 - p0) <= 1e-12 * np.abs(p0).max())


@pytest.mark.parametrize("dim,n", [(2, 10), (3, 6)])
def test_intersection_densities_pass_by_pass(ctx, dim, n):
    """A two-component density transferred by intersection
    (OMEGA_H_DENSITY) on a jittered two-material box, compared after every
    refine, coarsen, swap and sliver pass (each of which changes the mesh)."""
    g = A.Mesh.box(ctx, dim, n)
    two_materials(g)
    rng = np.random.default_rng(dim)
    c = g.coords().reshape(-1, dim).copy()
    _, vd = g.classification(0)
    inner = vd == dim
    c[inner] += rng.uniform(-0.35 / n, 0.35 / n, size=(inner.sum(), dim))
    g.set_coords(c.ravel())
    g.set_metric(np.full(g.nents(0), float(n * n)), 1)
    o = O.Mesh.copy_of(g)
    rho = rng.uniform(0.5, 2.0, size=(g.nents(dim), 2))
    for m in (o, g):
        m.add_elem_tag("rho", 2, rho.ravel())
        m.set_elem_transfer("rho", m.DENSITY)
    oo, go = O.default_opts(dim), ctx.default_opts(dim)
    oo.verbosity = go.verbosity = 0
    changed = []
    for name in ("refine_by_size", "coarsen_by_size", "swap_edges", "coarsen_slivers",
                 "refine_by_size", "coarsen_by_size", "swap_edges"):
        did = getattr(o, name)(oo)
        assert getattr(g, name)(go) == did, name
        changed.append(did)
        assert_same(o.elem_tag("rho"), g.elem_tag("rho"), f"rho after {name}")
    assert changed[:3] == [1, 1, 1]
    compare_meshes(o, g)


def test_constant_density_stays_constant(ctx):
    g = A.Mesh.box(ctx, 3, 6)
    g.set_metric(g.implied_isos() * (1.0 / 1.6 ** 2), 1)
    g.add_elem_tag("rho", 1, np.full(g.nents(3), 2.5))
    g.set_elem_transfer("rho", g.DENSITY)
    opts = ctx.default_opts(3)
    opts.verbosity = 0
    g.adapt(opts)
    np.testing.assert_allclose(g.elem_tag("rho"), 2.5, rtol=1e-13)
