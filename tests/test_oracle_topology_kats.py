"""The reference's own known answers for the topology side of the oracle.

Every expected value below is the one the reference's unit tests assert:
simplex_down_template (src/unit_mesh.cpp:25-66), the alignment of a
triangle's vertex list under rotation and flip codes (:73-110), form_uses
(:112-127), reflect_down (:129-195), the single-simplex stars (:282-305), the
dual of two triangles (:307-315), mark_up/mark_down (:393-400),
swap2d_topology (:411-422), invert_map_by_atomics/_by_sorting and invert_adj
with codes (src/unit_array_algs.cpp:106-148).  The swap3d triangulation tables
are compared with the reference's data in tests/golden/swap3d_tables.json
(tests/golden/make_swap3d_tables.py), for the oracle and for the library.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

L = O.lib()
mc = O.make_code


def test_simplex_down_template():
    expect = {
        (1, 0): [[0], [1]],
        (2, 0): [[0], [1], [2]],
        (2, 1): [[0, 1], [1, 2], [2, 0]],
        (3, 0): [[0], [1], [2], [3]],
        (3, 1): [[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]],
        (3, 2): [[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]],
    }
    for (ed, bd), rows in expect.items():
        if ed == 1:
            continue  # 1D elements are not on the path; the oracle has no 1D templates
        for wb, verts in enumerate(rows):
            for wv, v in enumerate(verts):
                assert L.om_down_template(ed, bd, wb, wv) == v, (ed, bd, wb, wv)


def align_adj(code, n=3):
    """align_adj<3>(code, ident) (src/Omega_h_align.hpp:180-185) expressed by
    the oracle's index alignment: out[align_vert_index(i)] = in[i]."""
    out = [None] * n
    for i in range(n):
        out[O.align_vert_index(n, i, code)] = i
    return out


def test_triangle_alignment_codes():
    assert align_adj(mc(True, 0, 0)) == [0, 2, 1]
    assert align_adj(mc(False, 1, 0)) == [2, 0, 1]
    assert align_adj(mc(False, 2, 0)) == [1, 2, 0]
    assert align_adj(mc(False, 0, 0)) == [0, 1, 2]


def test_form_uses():
    assert O.form_uses(2, 1, [0, 1, 2]).tolist() == [0, 1, 1, 2, 2, 0]
    assert O.form_uses(3, 1, [0, 1, 2, 3]).tolist() == [0, 1, 1, 2, 2, 0, 0, 3, 1, 3, 2, 3]
    assert O.form_uses(3, 2, [0, 1, 2, 3]).tolist() == [0, 2, 1, 0, 1, 3, 1, 2, 3, 2, 0, 3]


def test_reflect_down():
    for hd, ld in ((2, 1), (3, 1), (3, 2)):
        ab2b, codes = O.reflect_down([], [], 0, hd, ld)
        assert ab2b.size == 0 and codes.size == 0
    a, c = O.reflect_down([0, 1, 2], [0, 1, 1, 2, 2, 0], 3, 2, 1)
    assert a.tolist() == [0, 1, 2] and c.tolist() == [0, 0, 0]
    a, c = O.reflect_down([0, 1, 2, 3], [0, 1, 1, 2, 2, 0, 0, 3, 1, 3, 2, 3], 4, 3, 1)
    assert a.tolist() == [0, 1, 2, 3, 4, 5] and c.tolist() == [0] * 6
    a, c = O.reflect_down([0, 1, 2, 3], [0, 2, 1, 0, 1, 3, 1, 2, 3, 2, 0, 3], 4, 3, 2)
    assert a.tolist() == [0, 1, 2, 3] and c.tolist() == [0] * 4
    a, c = O.reflect_down([0, 1, 2, 3], [0, 1, 2, 0, 3, 1, 1, 3, 2, 2, 3, 0], 4, 3, 2)
    assert a.tolist() == [0, 1, 2, 3] and c.tolist() == [mc(True, 0, 0)] * 4
    a, _ = O.reflect_down([0, 1, 2, 2, 3, 0], [0, 1, 1, 2, 2, 3, 3, 0, 0, 2], 4, 2, 1)
    assert a.tolist() == [0, 1, 4, 2, 3, 4]


def test_single_simplex_stars():
    m = O.Mesh.from_elems(2, [0, 1, 2], 3)
    a2ab, ab2b = m.star(0)
    assert a2ab.tolist() == [0, 2, 4, 6] and ab2b.tolist() == [1, 2, 0, 2, 0, 1]
    a2ab, ab2b = m.star(1)
    assert a2ab.tolist() == [0, 2, 4, 6] and ab2b.tolist() == [2, 1, 0, 2, 1, 0]
    m = O.Mesh.from_elems(3, [0, 1, 2, 3], 4)
    a2ab, ab2b = m.star(0)
    assert a2ab.tolist() == [0, 3, 6, 9, 12]
    assert ab2b.tolist() == [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2]
    a2ab, ab2b = m.star(1)
    assert a2ab.tolist() == list(range(0, 35, 5))
    assert ab2b.tolist() == [1, 3, 4, 2, 5, 3, 0, 2, 5, 4, 0, 4, 5, 1, 3, 0, 1, 5, 4, 2, 2,
                             0, 3, 5, 1, 1, 2, 4, 3, 0]


def test_dual_of_two_triangles():
    m = O.Mesh.from_elems(2, [0, 1, 2, 2, 3, 0], 4)
    a2ab, ab2b = m.dual()
    assert a2ab.tolist() == [0, 1, 2] and ab2b.tolist() == [1, 0]


def test_mark_up_down():
    m = O.Mesh.box(2, 1, internal=True)
    assert m.mark_down(2, 0, [1, 0]).tolist() == [1, 1, 0, 1]
    assert m.mark_up(0, 2, [0, 1, 0, 0]).tolist() == [1, 0]


def test_swap2d_topology():
    m = O.Mesh.box(2, 1, internal=True)
    ev, tv = m.swap2d_topology([2])
    assert ev.tolist() == [2, 1]
    assert tv.tolist() == [3, 2, 1, 0, 1, 2]


def test_invert_map():
    a2ab, ab2b = O.invert_map([], 4)
    assert a2ab.tolist() == [0] * 5 and ab2b.size == 0
    a2ab, ab2b = O.invert_map([0, 1, 2, 3], 4)
    assert a2ab.tolist() == [0, 1, 2, 3, 4] and ab2b.tolist() == [0, 1, 2, 3]
    a2ab, ab2b = O.invert_map([1, 0, 1, 0], 2)
    assert a2ab.tolist() == [0, 2, 4] and ab2b.tolist() == [1, 3, 0, 2]


def test_invert_adj_with_codes():
    # the up adjacency of the two triangles {0,1,2}, {2,3,0}: invert_adj
    # (src/Omega_h_adj.cpp:231-263) with which_down in the codes
    m = O.Mesh.from_elems(2, [0, 1, 2, 2, 3, 0], 4)
    a2ab, ab2b, codes = m.up(0, 2)
    assert a2ab.tolist() == [0, 2, 3, 5, 6]
    assert ab2b.tolist() == [0, 1, 0, 0, 1, 1]
    assert codes.tolist() == [mc(0, 0, 0), mc(0, 0, 2), mc(0, 0, 1), mc(0, 0, 2), mc(0, 0, 0),
                              mc(0, 0, 1)]


def swap_fixture():
    with open(os.path.join(os.path.dirname(__file__), "golden", "swap3d_tables.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_oracle_swap3d_tables_are_the_references(n):
    ref = swap_fixture()[str(n)]
    tris, meshes, ue, me = O.swap3d_tables_of_oracle(n)
    assert tris == ref["triangles"]
    assert meshes == ref["meshes"]
    assert ue == ref["unique_edges"]
    if n >= 4:
        assert me == ref["mesh_edges"]


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_library_swap3d_tables_are_the_references(n):
    A = pytest.importorskip("omega_h_amd")
    ref = swap_fixture()[str(n)]
    tris, meshes, ue, me = A.swap3d_tables(n)
    assert tris == ref["triangles"]
    assert meshes == ref["meshes"]
    assert ue == ref["unique_edges"]
    if n >= 4:
        assert me == ref["mesh_edges"]
