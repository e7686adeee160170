"""The reference's topology known answers run through the library on the GPU:
invert_adj with codes (src/unit_array_algs.cpp:139-148), the single-simplex
stars (src/unit_mesh.cpp:282-305) and the dual of two triangles (:307-315),
each on a mesh built by osh_build_from_elems."""
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
mc = O.make_code


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def test_invert_adj_with_codes(ctx):
    m = A.Mesh.from_elems(ctx, 2, [0, 1, 2, 2, 3, 0], 4)
    a2ab, ab2b, codes = m.up(0, 2)
    assert a2ab.tolist() == [0, 2, 3, 5, 6]
    assert ab2b.tolist() == [0, 1, 0, 0, 1, 1]
    assert codes.tolist() == [mc(0, 0, 0), mc(0, 0, 2), mc(0, 0, 1), mc(0, 0, 2), mc(0, 0, 0),
                              mc(0, 0, 1)]


def test_single_simplex_stars(ctx):
    m = A.Mesh.from_elems(ctx, 2, [0, 1, 2], 3)
    a2ab, ab2b = m.star(0)
    assert a2ab.tolist() == [0, 2, 4, 6] and ab2b.tolist() == [1, 2, 0, 2, 0, 1]
    a2ab, ab2b = m.star(1)
    assert a2ab.tolist() == [0, 2, 4, 6] and ab2b.tolist() == [2, 1, 0, 2, 1, 0]
    m = A.Mesh.from_elems(ctx, 3, [0, 1, 2, 3], 4)
    a2ab, ab2b = m.star(0)
    assert a2ab.tolist() == [0, 3, 6, 9, 12]
    assert ab2b.tolist() == [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2]
    a2ab, ab2b = m.star(1)
    assert a2ab.tolist() == list(range(0, 35, 5))
    assert ab2b.tolist() == [1, 3, 4, 2, 5, 3, 0, 2, 5, 4, 0, 4, 5, 1, 3, 0, 1, 5, 4, 2, 2,
                             0, 3, 5, 1, 1, 2, 4, 3, 0]


def test_dual_of_two_triangles(ctx):
    m = A.Mesh.from_elems(ctx, 2, [0, 1, 2, 2, 3, 0], 4)
    a2ab, ab2b = m.dual()
    assert a2ab.tolist() == [0, 1, 2] and ab2b.tolist() == [1, 0]
