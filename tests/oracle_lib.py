"""ctypes binding of the oracle (oracle/liboracle.so).

The oracle is the CPU restatement of Omega_h's adapt() path that the tests
use as the checker.  Only tests/, the smoke check and bench.py's CPU-baseline
leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


class AdaptOpts(C.Structure):
    _fields_ = [
        ("min_length_desired", C.c_double),
        ("max_length_desired", C.c_double),
        ("max_length_allowed", C.c_double),
        ("min_quality_allowed", C.c_double),
        ("min_quality_desired", C.c_double),
        ("nsliver_layers", C.c_int),
        ("verbosity", C.c_int),
        ("should_refine", C.c_int),
        ("should_coarsen", C.c_int),
        ("should_swap", C.c_int),
        ("should_coarsen_slivers", C.c_int),
        ("should_prevent_coarsen_flip", C.c_int),
    ]


MAX_PASSES = 4096


class AdaptStats(C.Structure):
    _fields_ = [
        ("npasses", C.c_int),
        ("kind", C.c_int * MAX_PASSES),
        ("nelems_in", C.c_int * MAX_PASSES),
        ("nelems_out", C.c_int * MAX_PASSES),
        ("seconds", C.c_double * MAX_PASSES),
    ]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        L.om_build_box.restype = P
        L.om_build_box.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_int]
        L.om_build_box_internal.restype = P
        L.om_build_box_internal.argtypes = L.om_build_box.argtypes
        L.om_build_from_elems.restype = P
        L.om_build_from_elems.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
        L.om_mesh_free.argtypes = [P]
        L.om_mesh_join.restype = P
        L.om_mesh_join.argtypes = [C.c_int, C.POINTER(P)]
        L.om_mesh_from_arrays.restype = P
        L.om_mesh_from_arrays.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(P),
                                          C.POINTER(P), C.POINTER(C.c_double), C.c_int,
                                          C.POINTER(C.c_double), C.POINTER(P), C.POINTER(P)]
        L.om_mesh_dim.argtypes = [P]
        L.om_mesh_nents.argtypes = [P, C.c_int]
        L.om_get_verts.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.om_get_down.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_byte)]
        L.om_get_coords.argtypes = [P, C.POINTER(C.c_double)]
        L.om_set_metric.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
        L.om_metric_ncomps.argtypes = [P]
        L.om_get_metric.argtypes = [P, C.POINTER(C.c_double)]
        L.om_get_class.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_byte)]
        L.om_get_lengths.argtypes = [P, C.POINTER(C.c_double)]
        L.om_get_qualities.argtypes = [P, C.POINTER(C.c_double)]
        L.om_up_size.argtypes = [P, C.c_int, C.c_int]
        L.om_get_up.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.POINTER(C.c_byte)]
        L.om_dual_size.argtypes = [P]
        L.om_get_dual.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.om_get_sizes.argtypes = [P, C.POINTER(C.c_double)]
        L.om_star_size.argtypes = [P, C.c_int]
        L.om_get_star.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for name in ("om_refine_by_size", "om_coarsen_by_size", "om_swap_edges",
                     "om_coarsen_slivers"):
            getattr(L, name).argtypes = [P, C.POINTER(AdaptOpts)]
        L.om_adapt.argtypes = [P, C.POINTER(AdaptOpts), C.POINTER(AdaptStats)]
        L.om_adapt_digests.argtypes = [P, C.POINTER(AdaptOpts), C.POINTER(AdaptStats),
                                       C.POINTER(C.c_uint64)]
        L.om_mesh_digests.argtypes = [P, C.POINTER(C.c_uint64)]
        IPt, BPt = C.POINTER(C.c_int), C.POINTER(C.c_byte)
        L.om_down_template.argtypes = [C.c_int] * 4
        L.om_opposite_template.argtypes = [C.c_int] * 3
        L.om_kat_form_uses.argtypes = [C.c_int, C.c_int, C.c_int, IPt, IPt]
        L.om_kat_reflect_down.argtypes = [C.c_int, C.c_int, C.c_int, IPt, C.c_int, IPt,
                                          C.c_int, IPt, BPt]
        L.om_kat_invert_map.argtypes = [C.c_int, IPt, C.c_int, IPt, IPt]
        L.om_kat_invert_adj.argtypes = [C.c_int, C.c_int, C.c_int, IPt, BPt, IPt, IPt, BPt]
        L.om_kat_mident_metrics.argtypes = [P, C.c_int, IPt, C.POINTER(C.c_double)]
        L.om_kat_swap3d_qualities.argtypes = [P, C.c_double, C.c_int, IPt,
                                              C.POINTER(C.c_double), BPt]
        L.om_kat_align_vert_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.om_kat_mark_down.argtypes = [P, C.c_int, C.c_int, BPt, BPt]
        L.om_kat_mark_up.argtypes = [P, C.c_int, C.c_int, BPt, BPt]
        L.om_kat_swap2d_topology.argtypes = [P, C.c_int, IPt, IPt, IPt]
        L.om_swap_triangle.restype = IPt
        L.om_swap_triangle.argtypes = [C.c_int, C.c_int]
        L.om_swap_unique_edge.restype = IPt
        L.om_swap_unique_edge.argtypes = [C.c_int, C.c_int]
        L.om_swap_mesh_tri.argtypes = [C.c_int, C.c_int, C.c_int]
        L.om_swap_mesh_edge.argtypes = [C.c_int, C.c_int, C.c_int]
        L.om_limit_metric_gradation.argtypes = [P, C.c_double, C.c_double]
        L.om_add_vtag.argtypes = [P, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.om_vtag_ncomps.argtypes = [P, C.c_char_p]
        L.om_get_vtag.argtypes = [P, C.c_char_p, C.POINTER(C.c_double)]
        L.om_remove_vtag.argtypes = [P, C.c_char_p]
        L.om_interpolate_metrics.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.c_double,
                                             C.POINTER(C.c_double)]
        L.om_approach_metric.argtypes = [P, C.POINTER(AdaptOpts), C.c_double]
        L.om_warp_to_limit.argtypes = [P, C.POINTER(AdaptOpts), C.c_int, C.c_int]
        L.om_get_implied_isos.argtypes = [P, C.POINTER(C.c_double)]
        L.om_add_etag.argtypes = [P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]
        L.om_etag_ncomps.argtypes = [P, C.c_char_p]
        L.om_get_etag.argtypes = [P, C.c_char_p, C.POINTER(C.c_double)]
        L.om_set_etag_transfer.argtypes = [P, C.c_char_p, C.c_int, C.c_int, C.c_double,
                                           C.c_double]
        L.om_set_momentum_velocity.argtypes = [P, C.c_char_p, C.c_char_p, C.c_int, C.c_double,
                                               C.c_double]
        L.om_intersect_size.restype = C.c_double
        L.om_intersect_size.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int)]
        L.om_kat_barycentric.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
        L.om_kat_max_eigenvalue.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.om_kat_max_eigenvalue.restype = C.c_double
        L.om_ill_metric.argtypes = [P, C.POINTER(AdaptOpts), C.c_int]
        L.om_fix_for_given_metric.argtypes = [P, C.POINTER(AdaptOpts)]
        L.om_fix.argtypes = [P, C.POINTER(AdaptOpts), C.c_int]
        L.om_grade_fix_adapt.argtypes = [P, C.POINTER(AdaptOpts), C.c_int,
                                         C.POINTER(C.c_double)]
        L.om_set_coords.argtypes = [P, C.POINTER(C.c_double)]
        L.om_get_implied_metrics.argtypes = [P, C.POINTER(C.c_double)]
        L.om_add_itag.argtypes = [P, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.om_itag_ncomps.argtypes = [P, C.c_char_p]
        L.om_get_itag.argtypes = [P, C.c_char_p, C.c_int, C.POINTER(C.c_double)]
        L.om_element_implied_metric.argtypes = [C.c_int, C.POINTER(C.c_void_p),
                                                C.POINTER(C.c_double)]
        L.om_default_opts.argtypes = [C.c_int, C.POINTER(AdaptOpts)]
        for name in ("om_log", "om_exp", "om_cbrt", "om_acos", "om_cos"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.om_kat_hilbert.argtypes = [C.POINTER(C.c_uint64), C.c_int, C.c_int,
                                     C.POINTER(C.c_uint64)]
        L.om_kat_roots3.argtypes = [C.POINTER(C.c_double), C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.om_kat_decompose_eigen.argtypes = [C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.om_kat_log_spd.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.om_kat_exp_spd.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.om_kat_quality.restype = C.c_double
        L.om_kat_quality.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.om_kat_edge_length.restype = C.c_double
        L.om_kat_edge_length.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 4
        L.om_kat_aniso_length.restype = C.c_double
        L.om_kat_aniso_length.argtypes = [C.c_double, C.c_double]
        L.om_kat_average_metric.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
        L.om_kat_intersect.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 3
        L.om_kat_refine_qualities.argtypes = [P, C.c_int, C.POINTER(C.c_int),
                                              C.POINTER(C.c_double)]
        L.om_kat_coarsen_qualities.argtypes = [P, C.c_int, C.POINTER(C.c_int),
                                                C.POINTER(C.c_byte), C.POINTER(C.c_double)]
        L.om_kat_find_loop.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.om_kat_indset.argtypes = [P, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_byte), C.POINTER(C.c_byte)]
        _lib = L
    return _lib


def default_opts(dim):
    o = AdaptOpts()
    lib().om_default_opts(dim, C.byref(o))
    return o


class Mesh:
    """Owning handle on an oracle mesh."""

    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_mesh_free(self.h)
            self.h = None

    @classmethod
    def box(cls, dim, n, internal=False, x=1.0, y=1.0, z=1.0):
        nz = n if dim == 3 else 0
        z = z if dim == 3 else 0.0
        f = lib().om_build_box_internal if internal else lib().om_build_box
        return cls(f(dim, x, y, z, n, n, nz))

    @classmethod
    def from_elems(cls, dim, ev2v, nverts):
        a = np.ascontiguousarray(ev2v, dtype=np.int32)
        return cls(lib().om_build_from_elems(dim, len(a) // (dim + 1), _ptr(a, C.c_int), nverts))

    @classmethod
    def join(cls, parts):
        """One disjoint mesh of the parts (vertex tag "part" = part number)."""
        arr = (C.c_void_p * len(parts))(*[p.h for p in parts])
        return cls(lib().om_mesh_join(len(parts), arr))

    @classmethod
    def copy_of(cls, other):
        """An oracle mesh with the arrays of another mesh object (the library's
        or the oracle's): down adjacencies and codes, coordinates, metric and
        classification."""
        d = other.dim
        nents = (C.c_int * 4)(*[other.nents(k) if k <= d else 0 for k in range(4)])
        keep = []
        downs = (C.c_void_p * 4)()
        codes = (C.c_void_p * 4)()
        for k in range(1, d + 1):
            dn, cd = other.down(k)
            dn = np.ascontiguousarray(dn, dtype=np.int32)
            cd = np.ascontiguousarray(cd, dtype=np.int8)
            keep += [dn, cd]
            downs[k], codes[k] = dn.ctypes.data, cd.ctypes.data
        ids = (C.c_void_p * 4)()
        dims = (C.c_void_p * 4)()
        for k in range(d + 1):
            i, c = other.classification(k)
            i = np.ascontiguousarray(i, dtype=np.int32)
            c = np.ascontiguousarray(c, dtype=np.int8)
            keep += [i, c]
            ids[k], dims[k] = i.ctypes.data, c.ctypes.data
        x = np.ascontiguousarray(other.coords(), dtype=np.float64)
        mt = np.ascontiguousarray(other.metric(), dtype=np.float64)
        nc = len(mt) // max(other.nents(0), 1)
        h = lib().om_mesh_from_arrays(d, nents, downs, codes, _ptr(x, C.c_double), nc,
                                      _ptr(mt, C.c_double), ids, dims)
        return cls(h)

    @property
    def dim(self):
        return lib().om_mesh_dim(self.h)

    def nents(self, d):
        return lib().om_mesh_nents(self.h, d)

    def verts(self, d):
        out = np.zeros(self.nents(d) * (d + 1), dtype=np.int32)
        lib().om_get_verts(self.h, d, _ptr(out, C.c_int))
        return out

    def down(self, d):
        n = self.nents(d) * (d + 1)
        out = np.zeros(n, dtype=np.int32)
        codes = np.zeros(n, dtype=np.int8)
        lib().om_get_down(self.h, d, _ptr(out, C.c_int), _ptr(codes, C.c_byte))
        return out, codes

    def coords(self):
        out = np.zeros(self.nents(0) * self.dim, dtype=np.float64)
        lib().om_get_coords(self.h, _ptr(out, C.c_double))
        return out

    def set_metric(self, metric, ncomps):
        a = np.ascontiguousarray(metric, dtype=np.float64)
        lib().om_set_metric(self.h, ncomps, _ptr(a, C.c_double))

    def metric(self):
        nc = lib().om_metric_ncomps(self.h)
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        lib().om_get_metric(self.h, _ptr(out, C.c_double))
        return out

    def classification(self, d):
        ids = np.zeros(self.nents(d), dtype=np.int32)
        dims = np.zeros(self.nents(d), dtype=np.int8)
        lib().om_get_class(self.h, d, _ptr(ids, C.c_int), _ptr(dims, C.c_byte))
        return ids, dims

    def lengths(self):
        out = np.zeros(self.nents(1), dtype=np.float64)
        lib().om_get_lengths(self.h, _ptr(out, C.c_double))
        return out

    def qualities(self):
        out = np.zeros(self.nents(self.dim), dtype=np.float64)
        lib().om_get_qualities(self.h, _ptr(out, C.c_double))
        return out

    def up(self, low, high):
        n = lib().om_up_size(self.h, low, high)
        a2ab = np.zeros(self.nents(low) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        codes = np.zeros(n, dtype=np.int8)
        lib().om_get_up(self.h, low, high, _ptr(a2ab, C.c_int), _ptr(ab2b, C.c_int),
                        _ptr(codes, C.c_byte))
        return a2ab, ab2b, codes

    def dual(self):
        n = lib().om_dual_size(self.h)
        a2ab = np.zeros(self.nents(self.dim) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        lib().om_get_dual(self.h, _ptr(a2ab, C.c_int), _ptr(ab2b, C.c_int))
        return a2ab, ab2b

    def sizes(self):
        out = np.zeros(self.nents(self.dim), dtype=np.float64)
        lib().om_get_sizes(self.h, _ptr(out, C.c_double))
        return out

    def star(self, d):
        n = lib().om_star_size(self.h, d)
        a2ab = np.zeros(self.nents(d) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        lib().om_get_star(self.h, d, _ptr(a2ab, C.c_int), _ptr(ab2b, C.c_int))
        return a2ab, ab2b

    def mident_metrics(self, edges):
        e = _ints(edges)
        nc = lib().om_metric_ncomps(self.h)
        out = np.zeros(max(len(e) * nc, 1), dtype=np.float64)
        eb = e if len(e) else np.zeros(1, dtype=np.int32)
        lib().om_kat_mident_metrics(self.h, len(e), _ptr(eb, C.c_int), _ptr(out, C.c_double))
        return out[:len(e) * nc]

    def swap3d_qualities(self, max_length, edges):
        """(qualities, configs) of the listed interior edges, as the swap pass
        decides them."""
        e = _ints(edges)
        q = np.zeros(max(len(e), 1), dtype=np.float64)
        cfg = np.zeros(max(len(e), 1), dtype=np.int8)
        eb = e if len(e) else np.zeros(1, dtype=np.int32)
        lib().om_kat_swap3d_qualities(self.h, max_length, len(e), _ptr(eb, C.c_int),
                                      _ptr(q, C.c_double), _ptr(cfg, C.c_byte))
        return q[:len(e)], cfg[:len(e)]

    def refine_qualities(self, cands):
        c = np.ascontiguousarray(cands, dtype=np.int32)
        out = np.zeros(len(c), dtype=np.float64)
        lib().om_kat_refine_qualities(self.h, len(c), _ptr(c, C.c_int), _ptr(out, C.c_double))
        return out

    def coarsen_qualities(self, cands, codes):
        c = _ints(cands)
        k = np.ascontiguousarray(codes, dtype=np.int8)
        out = np.zeros(2 * len(c), dtype=np.float64)
        lib().om_kat_coarsen_qualities(self.h, len(c), _ptr(c, C.c_int), _ptr(k, C.c_byte),
                                       _ptr(out, C.c_double))
        return out

    def find_loop(self, edge):
        eev = np.zeros(2, dtype=np.int32)
        lv = np.zeros(7, dtype=np.int32)
        n = lib().om_kat_find_loop(self.h, edge, _ptr(eev, C.c_int), _ptr(lv, C.c_int))
        return eev, lv[:n]

    def mark_down(self, high, low, marks):
        mk = np.ascontiguousarray(marks, dtype=np.int8)
        out = np.zeros(self.nents(low), dtype=np.int8)
        lib().om_kat_mark_down(self.h, high, low, _ptr(mk, C.c_byte), _ptr(out, C.c_byte))
        return out

    def mark_up(self, low, high, marks):
        mk = np.ascontiguousarray(marks, dtype=np.int8)
        out = np.zeros(self.nents(high), dtype=np.int8)
        lib().om_kat_mark_up(self.h, low, high, _ptr(mk, C.c_byte), _ptr(out, C.c_byte))
        return out

    def swap2d_topology(self, keys2edges):
        k = _ints(keys2edges)
        ev = np.zeros(2 * len(k), dtype=np.int32)
        tv = np.zeros(6 * len(k), dtype=np.int32)
        lib().om_kat_swap2d_topology(self.h, len(k), _ptr(k, C.c_int), _ptr(ev, C.c_int),
                                     _ptr(tv, C.c_int))
        return ev, tv

    def indset(self, d, quality, cands):
        q = np.ascontiguousarray(quality, dtype=np.float64)
        c = np.ascontiguousarray(cands, dtype=np.int8)
        out = np.zeros(self.nents(d), dtype=np.int8)
        lib().om_kat_indset(self.h, d, _ptr(q, C.c_double), _ptr(c, C.c_byte),
                            _ptr(out, C.c_byte))
        return out

    def refine_by_size(self, opts):
        return lib().om_refine_by_size(self.h, C.byref(opts))

    def coarsen_by_size(self, opts):
        return lib().om_coarsen_by_size(self.h, C.byref(opts))

    def swap_edges(self, opts):
        return lib().om_swap_edges(self.h, C.byref(opts))

    def coarsen_slivers(self, opts):
        return lib().om_coarsen_slivers(self.h, C.byref(opts))

    def limit_metric_gradation(self, max_rate, tol=1e-2):
        return lib().om_limit_metric_gradation(self.h, max_rate, tol)

    LINEAR, METRIC = 0, 1

    def add_vertex_tag(self, name, ncomps, kind, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert lib().om_add_vtag(self.h, name.encode(), ncomps, kind, _ptr(a, C.c_double)) == 0

    def vertex_tag(self, name):
        nc = lib().om_vtag_ncomps(self.h, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        lib().om_get_vtag(self.h, name.encode(), _ptr(out, C.c_double))
        return out

    def approach_metric(self, opts, min_step=1e-4):
        return lib().om_approach_metric(self.h, C.byref(opts), min_step)

    def warp_to_limit(self, opts, exit_on_stall=False, max_niters=40):
        return lib().om_warp_to_limit(self.h, C.byref(opts), int(exit_on_stall), max_niters)

    def set_coords(self, coords):
        a = np.ascontiguousarray(coords, dtype=np.float64)
        assert a.size == self.nents(0) * self.dim
        lib().om_set_coords(self.h, _ptr(a, C.c_double))

    @property
    def metric_ncomps(self):
        return lib().om_metric_ncomps(self.h)

    def add_elem_tag(self, name, ncomps, data):
        """An element tag transferred pointwise (OMEGA_H_POINTWISE)."""
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.size == self.nents(self.dim) * ncomps
        assert lib().om_add_etag(self.h, name.encode(), ncomps, _ptr(a, C.c_double)) == 0

    POINTWISE, CONSERVE, DENSITY = 0, 1, 2
    DIFFUSE_NONE, DIFFUSE_RELATIVE, DIFFUSE_ABSOLUTE = 0, 1, 2

    def set_elem_transfer(self, name, kind, diffuse=(0, 0.0, 0.0)):
        """The transfer of an element tag (POINTWISE, CONSERVE or DENSITY) and,
        for CONSERVE, the diffusion (type, tolerance, floor) of its integral's
        error."""
        t, tol, fl = diffuse
        assert lib().om_set_etag_transfer(self.h, name.encode(), kind, t, tol, fl) == 0

    def set_momentum_velocity(self, name, density, diffuse=(0, 0.0, 0.0)):
        t, tol, fl = diffuse
        assert lib().om_set_momentum_velocity(self.h, name.encode(), density.encode(), t, tol,
                                              fl) == 0

    def elem_tag(self, name):
        nc = lib().om_etag_ncomps(self.h, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(self.dim) * nc, dtype=np.float64)
        lib().om_get_etag(self.h, name.encode(), _ptr(out, C.c_double))
        return out

    def implied_isos(self):
        out = np.zeros(self.nents(0), dtype=np.float64)
        lib().om_get_implied_isos(self.h, _ptr(out, C.c_double))
        return out

    def ill_metric(self, opts, isotropy):
        lib().om_ill_metric(self.h, C.byref(opts), isotropy)

    def fix(self, opts, isotropy):
        return lib().om_fix(self.h, C.byref(opts), isotropy)

    def grade_fix_adapt(self, opts, target, ncomps):
        t = np.ascontiguousarray(target, dtype=np.float64)
        return lib().om_grade_fix_adapt(self.h, C.byref(opts), ncomps, _ptr(t, C.c_double))

    def add_inherit_tag(self, name, ncomps, per_dim):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in per_dim]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        assert lib().om_add_itag(self.h, name.encode(), ncomps, ptrs) == 0

    def inherit_tag(self, name, d):
        nc = lib().om_itag_ncomps(self.h, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(d) * nc, dtype=np.float64)
        lib().om_get_itag(self.h, name.encode(), d, _ptr(out, C.c_double))
        return out

    def implied_metrics(self):
        nc = self.dim * (self.dim + 1) // 2
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        lib().om_get_implied_metrics(self.h, _ptr(out, C.c_double))
        return out

    def adapt_digests(self, opts):
        """adapt() with four digests of the mesh after every pass (topology and
        classification, coordinates, metric, lengths and qualities)."""
        st = AdaptStats()
        dg = (C.c_uint64 * (4 * MAX_PASSES))()
        r = lib().om_adapt_digests(self.h, C.byref(opts), C.byref(st), dg)
        passes = [(st.kind[i], st.nelems_in[i], st.nelems_out[i], st.seconds[i])
                  for i in range(st.npasses)]
        digests = [tuple(dg[4 * i + k] for k in range(4)) for i in range(st.npasses)]
        return r, passes, digests

    def digests(self):
        dg = (C.c_uint64 * 4)()
        lib().om_mesh_digests(self.h, dg)
        return tuple(dg)

    def adapt(self, opts):
        st = AdaptStats()
        r = lib().om_adapt(self.h, C.byref(opts), C.byref(st))
        passes = [(st.kind[i], st.nelems_in[i], st.nelems_out[i], st.seconds[i])
                  for i in range(st.npasses)]
        return r, passes


def _ints(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def form_uses(high_dim, low_dim, hv2v):
    """form_uses(hv2v, SIMPLEX, high_dim, low_dim), src/Omega_h_adj.cpp:155-176."""
    h = _ints(hv2v)
    nhigh = len(h) // (high_dim + 1)
    deg = lib().om_simplex_degree(high_dim, low_dim)
    out = np.zeros(max(nhigh * deg * (low_dim + 1), 1), dtype=np.int32)
    lib().om_kat_form_uses(nhigh, high_dim, low_dim, _ptr(h, C.c_int), _ptr(out, C.c_int))
    return out[:nhigh * deg * (low_dim + 1)]


def reflect_down(hv2v, lv2v, nverts, high_dim, low_dim):
    """reflect_down(hv2v, lv2v, SIMPLEX, nverts, high_dim, low_dim),
    src/Omega_h_adj.cpp:424-441: (ab2b, codes)."""
    h, lo = _ints(hv2v), _ints(lv2v)
    nhigh = len(h) // (high_dim + 1)
    nlow = len(lo) // (low_dim + 1)
    deg = lib().om_simplex_degree(high_dim, low_dim)
    ab2b = np.zeros(max(nhigh * deg, 1), dtype=np.int32)
    codes = np.zeros(max(nhigh * deg, 1), dtype=np.int8)
    hb = h if len(h) else np.zeros(1, dtype=np.int32)
    lb = lo if len(lo) else np.zeros(1, dtype=np.int32)
    lib().om_kat_reflect_down(high_dim, low_dim, nhigh, _ptr(hb, C.c_int), nlow,
                              _ptr(lb, C.c_int), nverts, _ptr(ab2b, C.c_int),
                              _ptr(codes, C.c_byte))
    return ab2b[:nhigh * deg], codes[:nhigh * deg]


def invert_adj(hl2l, deg, nlow, down_codes=None):
    """invert_adj, src/Omega_h_adj.cpp:231-263: (a2ab, ab2b, codes)."""
    h = _ints(hl2l)
    nhigh = len(h) // deg
    a2ab = np.zeros(nlow + 1, dtype=np.int32)
    ab2b = np.zeros(max(len(h), 1), dtype=np.int32)
    codes = np.zeros(max(len(h), 1), dtype=np.int8)
    hb = h if len(h) else np.zeros(1, dtype=np.int32)
    dc = None if down_codes is None else np.ascontiguousarray(down_codes, dtype=np.int8)
    lib().om_kat_invert_adj(nhigh, deg, nlow, _ptr(hb, C.c_int),
                            None if dc is None else _ptr(dc, C.c_byte), _ptr(a2ab, C.c_int),
                            _ptr(ab2b, C.c_int), _ptr(codes, C.c_byte))
    return a2ab, ab2b[:len(h)], codes[:len(h)]


def invert_map(hl2l, nlow):
    """invert_map_by_atomics / _by_sorting, src/Omega_h_map.cpp:238-280."""
    h = _ints(hl2l)
    a2ab = np.zeros(nlow + 1, dtype=np.int32)
    ab2b = np.zeros(max(len(h), 1), dtype=np.int32)
    hb = h if len(h) else np.zeros(1, dtype=np.int32)
    lib().om_kat_invert_map(len(h), _ptr(hb, C.c_int), nlow, _ptr(a2ab, C.c_int),
                            _ptr(ab2b, C.c_int))
    return a2ab, ab2b[:len(h)]


def align_vert_index(n, index, code):
    return lib().om_kat_align_vert_index(n, index, code)


def make_code(flipped, rotation, which_down):
    """make_code, src/Omega_h_align.hpp:50-53."""
    return (which_down << 3) | (rotation << 1) | int(bool(flipped))


def swap3d_tables_of_oracle(n):
    L = lib()
    ntris = {3: 1, 4: 4, 5: 10, 6: 20, 7: 35}[n]
    nue = {3: 0, 4: 2, 5: 7, 6: 15, 7: 26}[n]
    nm = (C.c_int * 8).in_dll(L, "om_swap_mesh_counts")[n]
    tris = [[L.om_swap_triangle(n, t)[k] for k in range(3)] for t in range(ntris)]
    meshes = [[L.om_swap_mesh_tri(n, m, k) for k in range(n - 2)] for m in range(nm)]
    ue = [[L.om_swap_unique_edge(n, e)[k] for k in range(2)] for e in range(nue)]
    me = [[L.om_swap_mesh_edge(n, m, k) for k in range(n - 3)] for m in range(nm)]
    return tris, meshes, ue, me


def kat_quality(dim, mdim, pts, ms):
    p = np.ascontiguousarray(pts, dtype=np.float64).ravel()
    m = np.ascontiguousarray(ms, dtype=np.float64).ravel()
    return lib().om_kat_quality(dim, mdim, _ptr(p, C.c_double), _ptr(m, C.c_double))


def kat_decompose_eigen(m):
    """m: dim x dim numpy array (row i, col j); returns (q as columns, l)."""
    dim = m.shape[0]
    a = np.ascontiguousarray(m.T.ravel(), dtype=np.float64)  # column-major
    q = np.zeros(dim * dim)
    l = np.zeros(dim)
    lib().om_kat_decompose_eigen(dim, _ptr(a, C.c_double), _ptr(q, C.c_double),
                                 _ptr(l, C.c_double))
    return q.reshape(dim, dim).T, l


def kat_spd(fn, m):
    dim = m.shape[0]
    a = np.ascontiguousarray(m.T.ravel(), dtype=np.float64)
    out = np.zeros(dim * dim)
    getattr(lib(), fn)(dim, _ptr(a, C.c_double), _ptr(out, C.c_double))
    return out.reshape(dim, dim).T


def kat_roots3(coeffs, eps=1e-6):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    r = np.zeros(3)
    mm = np.zeros(3, dtype=np.int32)
    n = lib().om_kat_roots3(_ptr(c, C.c_double), eps, _ptr(r, C.c_double), _ptr(mm, C.c_int))
    return n, r, mm


def kat_average_metric(mdim, ms):
    m = np.ascontiguousarray(ms, dtype=np.float64).ravel()
    nc = mdim * (mdim + 1) // 2
    out = np.zeros(nc)
    lib().om_kat_average_metric(mdim, len(m) // nc, _ptr(m, C.c_double), _ptr(out, C.c_double))
    return out


def kat_intersect(dim, m1, m2):
    a = np.ascontiguousarray(m1, dtype=np.float64)
    b = np.ascontiguousarray(m2, dtype=np.float64)
    out = np.zeros(dim * (dim + 1) // 2)
    lib().om_kat_intersect(dim, _ptr(a, C.c_double), _ptr(b, C.c_double), _ptr(out, C.c_double))
    return out


def element_implied_metric(pts):
    """pts: (dim+1) x dim vertex coordinates; the symmetric-storage metric."""
    pts = [np.ascontiguousarray(p, dtype=np.float64) for p in pts]
    arr = (C.c_void_p * len(pts))(*[p.ctypes.data for p in pts])
    dim = len(pts) - 1
    out = np.zeros(dim * (dim + 1) // 2)
    lib().om_element_implied_metric(dim, arr, _ptr(out, C.c_double))
    return out


def interpolate_metrics(mdim, a, b, t):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    nc = mdim * (mdim + 1) // 2
    out = np.zeros_like(a)
    lib().om_interpolate_metrics(mdim, len(a) // nc, _ptr(a, C.c_double), _ptr(b, C.c_double),
                                 t, _ptr(out, C.c_double))
    return out


def kat_hilbert(X, b, n):
    x = np.ascontiguousarray(X, dtype=np.uint64)
    y = np.zeros(n, dtype=np.uint64)
    lib().om_kat_hilbert(_ptr(x, C.c_uint64), b, n, _ptr(y, C.c_uint64))
    return x, y


def symm2vector(m):
    """Omega_h's symmetric storage order (src/Omega_h_matrix.hpp:391-431)."""
    if m.shape[0] == 1:
        return np.array([m[0, 0]])
    if m.shape[0] == 2:
        return np.array([m[0, 0], m[1, 1], m[0, 1]])
    return np.array([m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[1, 2], m[0, 2]])


def vector2symm(v):
    if len(v) == 1:
        return np.array([[v[0]]])
    if len(v) == 3:
        return np.array([[v[0], v[2]], [v[2], v[1]]])
    return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])


def max_eigenvalue(dim, symm):
    """get_max_eigenvalues of one metric (src/Omega_h_eigen.cpp:7-40)."""
    a = np.ascontiguousarray(symm, dtype=np.float64)
    return lib().om_kat_max_eigenvalue(dim, _ptr(a, C.c_double))


def barycentric(dim, x, simplex):
    """barycentric_from_global (src/Omega_h_shape.hpp:46-63)."""
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ps = np.ascontiguousarray(simplex, dtype=np.float64)
    out = np.zeros(dim + 1, dtype=np.float64)
    lib().om_kat_barycentric(dim, _ptr(xs, C.c_double), _ptr(ps, C.c_double),
                             _ptr(out, C.c_double))
    return out


def intersect_size(dim, a, b):
    """The size of the intersection of simplex a with simplex b (points
    (dim + 1) x dim), as the oracle's restatement of r3d computes it."""
    pa = np.ascontiguousarray(a, dtype=np.float64).ravel()
    pb = np.ascontiguousarray(b, dtype=np.float64).ravel()
    bad = C.c_int(0)
    r = lib().om_intersect_size(dim, _ptr(pa, C.c_double), _ptr(pb, C.c_double), C.byref(bad))
    assert bad.value == 0
    return r
