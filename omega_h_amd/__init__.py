"""MI355X-native Omega_h adapt() path: Python binding of libomega_h_amd.so.

The library is the product: HIP kernels for gfx950 behind the C ABI declared
in include/omega_h_amd.h.  This module only binds that ABI; it raises when the
built library is missing instead of computing anything itself.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libomega_h_amd.so")

MAX_PASSES = 4096
PASS_NAMES = ("refine", "coarsen", "swap", "coarsen_slivers")


class AdaptOpts(C.Structure):
    """Omega_h::AdaptOpts (src/Omega_h_adapt.hpp:50-77)."""
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
        ("length_histogram_min", C.c_double),
        ("length_histogram_max", C.c_double),
        ("nlength_histogram_bins", C.c_int),
        ("nquality_histogram_bins", C.c_int),
        ("should_prevent_coarsen_flip", C.c_int),
        ("precision", C.c_int),
    ]


class AdaptStats(C.Structure):
    _fields_ = [
        ("npasses", C.c_int),
        ("kind", C.c_int * MAX_PASSES),
        ("nelems_in", C.c_int * MAX_PASSES),
        ("nelems_out", C.c_int * MAX_PASSES),
        ("seconds", C.c_double * MAX_PASSES),
        ("total_passes", C.c_longlong),
        ("total_nelems_in", C.c_longlong),
        ("total_seconds", C.c_double),
    ]


class ServeResult(C.Structure):
    """osh_serve_result of include/omega_h_amd.h."""
    _fields_ = [
        ("status", C.c_int),
        ("ticket", C.c_int),
        ("nelems_in", C.c_int),
        ("nelems_out", C.c_int),
        ("npasses", C.c_int),
        ("pass_elems", C.c_longlong),
        ("pass_seconds", C.c_double),
        ("queue_seconds", C.c_double),
        ("run_seconds", C.c_double),
        ("latency_seconds", C.c_double),
        ("digest", C.c_ulonglong),
        ("error", C.c_char * 200),
    ]


class OshError(RuntimeError):
    pass


_lib = None


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def load(path=None):
    """Load the HIP library (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OshError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    P = C.c_void_p
    I = C.c_int
    IP = C.POINTER(C.c_int)
    DP = C.POINTER(C.c_double)
    BP = C.POINTER(C.c_byte)
    sig = {
        "osh_ctx_create": [I, C.POINTER(P)],
        "osh_ctx_destroy": [P],
        "osh_ctx_synchronize": [P],
        "osh_ctx_stream": [P, C.POINTER(P)],
        "osh_ctx_set_tuning": [P, C.c_char_p, I],
        "osh_ctx_get_tuning": [P, C.c_char_p, IP],
        "osh_ctx_set_profile": [P, I],
        "osh_sort_by_keys": [P, I, I, C.POINTER(C.c_longlong), IP],
        "osh_invert_adj": [P, I, I, I, IP, BP, IP, IP, BP],
        "osh_collect_marked": [P, I, BP, IP, IP],
        "osh_mident_metrics": [P, I, IP, DP],
        "osh_swap3d_qualities": [P, C.POINTER(AdaptOpts), I, IP, DP, BP],
        "osh_reflect_down": [P, I, I, IP, IP, BP],
        "osh_ctx_profile_clear": [P],
        "osh_ctx_profile_nframes": [P, IP],
        "osh_ctx_profile_frame": [P, I, IP, DP, C.POINTER(C.c_longlong), C.c_char_p, I],
        "osh_ctx_profile_report": [P, C.c_double, I, C.c_char_p, C.c_longlong,
                                   C.POINTER(C.c_longlong)],
        "osh_default_opts": [I, C.POINTER(AdaptOpts)],
        "osh_build_box": [P, I, C.c_double, C.c_double, C.c_double, I, I, I, C.POINTER(P)],
        "osh_build_box_internal": [P, I, C.c_double, C.c_double, C.c_double, I, I, I,
                                   C.POINTER(P)],
        "osh_build_from_elems": [P, I, I, IP, I, C.POINTER(P)],
        "osh_mesh_destroy": [P],
        "osh_mesh_copy": [P, C.POINTER(P)],
        "osh_mesh_dim": [P, IP],
        "osh_mesh_join": [P, I, C.POINTER(P), C.POINTER(P)],
        "osh_mesh_split": [P, I, C.POINTER(P)],
        "osh_mesh_part_counts": [P, I, IP],
        "osh_mesh_nents": [P, I, IP],
        "osh_mesh_get_verts": [P, I, IP],
        "osh_mesh_get_down": [P, I, IP, BP],
        "osh_mesh_up_size": [P, I, I, IP],
        "osh_mesh_get_up": [P, I, I, IP, IP, BP],
        "osh_mesh_star_size": [P, I, IP],
        "osh_mesh_get_star": [P, I, IP, IP],
        "osh_mesh_dual_size": [P, IP],
        "osh_mesh_get_dual": [P, IP, IP],
        "osh_mesh_get_sizes": [P, DP],
        "osh_mesh_get_coords": [P, DP],
        "osh_mesh_set_coords": [P, DP],
        "osh_warp_to_limit": [P, C.POINTER(AdaptOpts), I, I, IP],
        "osh_get_implied_isos": [P, I, DP],
        "osh_mesh_add_elem_tag": [P, C.c_char_p, I, DP],
        "osh_mesh_elem_tag_ncomps": [P, C.c_char_p, IP],
        "osh_mesh_get_elem_tag": [P, C.c_char_p, DP],
        "osh_mesh_remove_elem_tag": [P, C.c_char_p],
        "osh_mesh_set_class": [P, I, IP, C.POINTER(C.c_byte)],
        "osh_mesh_set_elem_transfer": [P, C.c_char_p, I, I, C.c_double, C.c_double],
        "osh_mesh_set_momentum_velocity": [P, C.c_char_p, C.c_char_p, I, C.c_double,
                                           C.c_double],
        "osh_ill_metric": [P, C.POINTER(AdaptOpts), I],
        "osh_fix": [P, C.POINTER(AdaptOpts), I, IP],
        "osh_fix_for_given_metric": [P, C.POINTER(AdaptOpts), IP],
        "osh_grade_fix_adapt": [P, C.POINTER(AdaptOpts), I, DP, IP],
        "osh_mesh_set_metric": [P, I, DP],
        "osh_mesh_metric_ncomps": [P, IP],
        "osh_mesh_get_metric": [P, DP],
        "osh_mesh_get_class": [P, I, IP, BP],
        "osh_mesh_get_lengths": [P, DP],
        "osh_mesh_get_qualities": [P, DP],
        "osh_refine_by_size": [P, C.POINTER(AdaptOpts), IP],
        "osh_coarsen_by_size": [P, C.POINTER(AdaptOpts), IP],
        "osh_swap_edges": [P, C.POINTER(AdaptOpts), IP],
        "osh_coarsen_slivers": [P, C.POINTER(AdaptOpts), IP],
        "osh_adapt": [P, C.POINTER(AdaptOpts), C.POINTER(AdaptStats), IP],
        "osh_adapt_batch": [I, C.POINTER(P), C.POINTER(AdaptOpts), C.POINTER(AdaptStats), IP],
        "osh_adapt_digests": [P, C.POINTER(AdaptOpts), C.POINTER(AdaptStats),
                              C.POINTER(C.c_ulonglong), IP],
        "osh_mesh_digests": [P, C.POINTER(C.c_ulonglong)],
        "osh_mesh_screen_estimates": [P, DP, DP, BP],
        "osh_measure_edges_metric": [P, I, I, I, P, P, P, P, P],
        "osh_measure_qualities": [P, I, I, I, P, P, P, P, P],
        "osh_offset_scans": [P, I, P, IP, P],
        "osh_offset_scans_noalloc": [P, I, P, IP, P],
        "osh_refine_qualities": [P, I, IP, DP],
        "osh_find_loop": [P, I, IP, IP, IP],
        "osh_coarsen_qualities": [P, I, P, P, P, I, C.c_double],
        "osh_find_indset": [P, I, DP, BP, BP],
        "osh_eval_math": [P, I, I, DP, DP],
        "osh_swap3d_tables": [I, IP, IP, IP, IP, IP],
        "osh_limit_metric_gradation": [P, C.c_double, C.c_double, IP],
        "osh_intersect_metrics": [P, I, I, DP, DP, DP],
        "osh_mesh_digest": [P, C.POINTER(C.c_ulonglong)],
        "osh_mesh_add_vertex_tag": [P, C.c_char_p, I, I, DP],
        "osh_mesh_vertex_tag_ncomps": [P, C.c_char_p, IP],
        "osh_mesh_get_vertex_tag": [P, C.c_char_p, DP],
        "osh_mesh_remove_vertex_tag": [P, C.c_char_p],
        "osh_approach_metric": [P, C.POINTER(AdaptOpts), C.c_double, IP],
        "osh_interpolate_metrics": [P, I, I, DP, DP, C.c_double, DP],
        "osh_get_implied_metrics": [P, I, DP],
        "osh_write_osh": [P, C.c_char_p],
        "osh_mesh_add_inherit_tag": [P, C.c_char_p, I, C.POINTER(P)],
        "osh_mesh_inherit_tag_ncomps": [P, C.c_char_p, IP],
        "osh_mesh_get_inherit_tag": [P, C.c_char_p, I, DP],
        "osh_read_osh": [P, C.c_char_p, C.POINTER(P)],
        "osh_server_create": [I, I, C.c_double, C.POINTER(P)],
        "osh_server_destroy": [P],
        "osh_server_submit": [P, I, I, I, I, I, DP, C.POINTER(AdaptOpts), I, IP],
        "osh_server_wait": [P, I, C.POINTER(ServeResult)],
        "osh_server_set_timeout": [P, C.c_double],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.osh_last_error.restype = C.c_char_p
    L.osh_last_error.argtypes = []
    if path == LIB_PATH or _lib is None:
        _lib = L
    return L


def swap3d_tables(loop_size, lib=None):
    """The library's edge-swap tables for one loop size (host data, no GPU):
    (triangles, triangulations, unique interior edges, interior edges of each
    triangulation) as lists."""
    L = lib or load()
    cnt = np.zeros(4, dtype=np.int32)
    check(L.osh_swap3d_tables(loop_size, _p(cnt, C.c_int), None, None, None, None), L)
    ntris, nmeshes, nue, per = (int(x) for x in cnt)
    epm = max(loop_size - 3, 0)
    tris = np.zeros(3 * ntris, dtype=np.int32)
    meshes = np.zeros(max(nmeshes * per, 1), dtype=np.int32)
    ue = np.zeros(max(2 * nue, 1), dtype=np.int32)
    me = np.zeros(max(nmeshes * epm, 1), dtype=np.int32)
    check(L.osh_swap3d_tables(loop_size, None, _p(tris, C.c_int), _p(meshes, C.c_int),
                              _p(ue, C.c_int), _p(me, C.c_int)), L)
    return (tris.reshape(-1, 3).tolist(), meshes[:nmeshes * per].reshape(nmeshes, per).tolist(),
            ue[:2 * nue].reshape(-1, 2).tolist(),
            me[:nmeshes * epm].reshape(nmeshes, epm).tolist())


def check(status, lib=None):
    if status != 0:
        lib = lib or load()
        raise OshError(lib.osh_last_error().decode())


class Context:
    """One HIP stream on one device."""

    def __init__(self, device=0, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.osh_ctx_create(device, C.byref(h)), self.lib)
        self.h = h

    @property
    def stream(self):
        """The hipStream_t (as an int) that all work of this context runs on."""
        s = C.c_void_p()
        check(self.lib.osh_ctx_stream(self.h, C.byref(s)), self.lib)
        return s.value or 0

    TUNING = ("coarsen_lanes", "side_lanes", "choose_lanes", "first_lanes")

    def set_tuning(self, name, value):
        """One of the context's lane counts (TUNING; a power of two, 1 to 64)."""
        check(self.lib.osh_ctx_set_tuning(self.h, name.encode(), int(value)), self.lib)

    def tuning(self):
        """The context's lane counts, by name."""
        out = {}
        for name in self.TUNING:
            v = C.c_int()
            check(self.lib.osh_ctx_get_tuning(self.h, name.encode(), C.byref(v)), self.lib)
            out[name] = v.value
        return out

    def apply_tuning(self, settings):
        for name, value in settings.items():
            self.set_tuning(name, value)

    def sort_by_keys(self, keys, width=1):
        """The stable lexical sort permutation of len(keys) / width tuples
        of 64-bit keys (sort_by_keys)."""
        k = np.ascontiguousarray(keys, dtype=np.int64)
        n = len(k) // width
        perm = np.empty(n, dtype=np.int32)
        check(self.lib.osh_sort_by_keys(self.h, n, width, _p(k, C.c_longlong),
                                        _p(perm, C.c_int)), self.lib)
        return perm

    def invert_adj(self, hl2l, deg, nlow, down_codes=None):
        """invert_adj of a high -> low map: (a2ab, ab2b, codes)."""
        h = np.ascontiguousarray(hl2l, dtype=np.int32)
        nhigh = len(h) // deg
        a2ab = np.empty(nlow + 1, dtype=np.int32)
        ab2b = np.empty(max(len(h), 1), dtype=np.int32)
        codes = np.empty(max(len(h), 1), dtype=np.int8)
        dc = None if down_codes is None else np.ascontiguousarray(down_codes, dtype=np.int8)
        check(self.lib.osh_invert_adj(self.h, nhigh, deg, nlow, _p(h, C.c_int),
                                      None if dc is None else _p(dc, C.c_byte),
                                      _p(a2ab, C.c_int), _p(ab2b, C.c_int),
                                      _p(codes, C.c_byte)), self.lib)
        return a2ab, ab2b[:len(h)], codes[:len(h)]

    def collect_marked(self, marks):
        """Indices of the nonzero marks, in order."""
        m = np.ascontiguousarray(marks, dtype=np.int8)
        out = np.empty(max(len(m), 1), dtype=np.int32)
        n = C.c_int()
        check(self.lib.osh_collect_marked(self.h, len(m), _p(m, C.c_byte), _p(out, C.c_int),
                                          C.byref(n)), self.lib)
        return out[:n.value]

    PROFILE_TREE, PROFILE_MARKERS, PROFILE_SYNC = 1, 2, 4

    def set_profile(self, flags):
        """Named ranges (Omega_h::profile): PROFILE_TREE keeps the timer tree,
        PROFILE_MARKERS emits roctx ranges, PROFILE_SYNC makes each range wait
        for the stream at its end so that it is charged its device work."""
        check(self.lib.osh_ctx_set_profile(self.h, int(flags)), self.lib)

    def profile_clear(self):
        check(self.lib.osh_ctx_profile_clear(self.h), self.lib)

    def profile_frames(self):
        """The timer tree as a list of (name, parent index or -1, seconds,
        calls), parents before their children."""
        n = C.c_int()
        check(self.lib.osh_ctx_profile_nframes(self.h, C.byref(n)), self.lib)
        out = []
        name = C.create_string_buffer(256)
        parent, secs, calls = C.c_int(), C.c_double(), C.c_longlong()
        for i in range(n.value):
            check(self.lib.osh_ctx_profile_frame(self.h, i, C.byref(parent), C.byref(secs),
                                                 C.byref(calls), name, len(name)), self.lib)
            out.append((name.value.decode(), parent.value, secs.value, calls.value))
        return out

    def profile_report(self, chop=0.0, percent=False):
        """The TOP-DOWN and BOTTOM-UP listings of the timer tree."""
        n = C.c_longlong()
        check(self.lib.osh_ctx_profile_report(self.h, float(chop), int(percent), None, 0,
                                              C.byref(n)), self.lib)
        buf = C.create_string_buffer(n.value + 1)
        check(self.lib.osh_ctx_profile_report(self.h, float(chop), int(percent), buf,
                                              len(buf), C.byref(n)), self.lib)
        return buf.value.decode()

    def measure_qualities(self, dim, metric_ncomps, n, ev2v_ptr, coords_ptr, metric_ptr,
                          out_ptr, elems_ptr=0):
        """measure_qualities on device pointers (ints), on this context's stream."""
        check(self.lib.osh_measure_qualities(self.h, dim, metric_ncomps, n, elems_ptr or None,
                                             ev2v_ptr, coords_ptr, metric_ptr, out_ptr), self.lib)

    def measure_edges_metric(self, dim, metric_ncomps, n, ev2v_ptr, coords_ptr, metric_ptr,
                             out_ptr, edges_ptr=0):
        """measure_edges_metric on device pointers (ints), on this context's stream."""
        check(self.lib.osh_measure_edges_metric(self.h, dim, metric_ncomps, n, edges_ptr or None,
                                                ev2v_ptr, coords_ptr, metric_ptr, out_ptr),
              self.lib)

    def offset_scans(self, in_ptrs, sizes, out_ptrs):
        """offset_scan of up to twelve int arrays on device pointers, in one launch on
        this context's stream: out[j] gets n[j] + 1 exclusive offsets."""
        k = len(in_ptrs)
        ins = (C.c_void_p * k)(*[p or None for p in in_ptrs])
        outs = (C.c_void_p * k)(*out_ptrs)
        ns = (C.c_int * k)(*sizes)
        check(self.lib.osh_offset_scans(self.h, k, ins, ns, outs), self.lib)

    def offset_scans_noalloc(self, in_ptrs, sizes, out_ptrs):
        """offset_scans without any allocation (a fixed per-GPU scratch; at most
        16384 tiles of 4096 values per call)."""
        k = len(in_ptrs)
        ins = (C.c_void_p * k)(*[p or None for p in in_ptrs])
        outs = (C.c_void_p * k)(*out_ptrs)
        ns = (C.c_int * k)(*sizes)
        check(self.lib.osh_offset_scans_noalloc(self.h, k, ins, ns, outs), self.lib)

    def synchronize(self):
        check(self.lib.osh_ctx_synchronize(self.h), self.lib)

    def close(self):
        if getattr(self, "h", None):
            self.lib.osh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def default_opts(self, dim):
        o = AdaptOpts()
        check(self.lib.osh_default_opts(dim, C.byref(o)), self.lib)
        return o

    def intersect_metrics(self, dim, m1, m2):
        """Elementwise intersection of two arrays of symmetric metrics."""
        a = np.ascontiguousarray(m1, dtype=np.float64)
        b = np.ascontiguousarray(m2, dtype=np.float64)
        nc = dim * (dim + 1) // 2
        n = a.size // nc
        out = np.zeros_like(a)
        check(self.lib.osh_intersect_metrics(self.h, dim, n, _p(a, C.c_double), _p(b, C.c_double),
                                             _p(out, C.c_double)), self.lib)
        return out

    def interpolate_metrics(self, dim, a, b, t):
        """interpolate_between_metrics of two arrays of symmetric metrics."""
        x = np.ascontiguousarray(a, dtype=np.float64)
        y = np.ascontiguousarray(b, dtype=np.float64)
        n = x.size // (dim * (dim + 1) // 2)
        out = np.zeros_like(x)
        check(self.lib.osh_interpolate_metrics(self.h, dim, n, _p(x, C.c_double),
                                               _p(y, C.c_double), t, _p(out, C.c_double)),
              self.lib)
        return out

    def eval_math(self, which, xs):
        x = np.ascontiguousarray(xs, dtype=np.float64)
        out = np.zeros_like(x)
        check(self.lib.osh_eval_math(self.h, which, len(x), _p(x, C.c_double),
                                     _p(out, C.c_double)), self.lib)
        return out


def adapt_batch(meshes, opts):
    """adapt() of many meshes of one context at once (the batched form):
    each mesh ends where its own adapt() would leave it.  Returns the
    per-mesh results, the first passes of the joined meshes as (kind,
    elements in, elements out, seconds) and the totals over every pass
    (passes, elements in, seconds)."""
    lib = meshes[0].lib
    arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
    st = AdaptStats()
    did = (C.c_int * len(meshes))()
    check(lib.osh_adapt_batch(len(meshes), arr, C.byref(opts), C.byref(st), did), lib)
    passes = [(st.kind[i], st.nelems_in[i], st.nelems_out[i], st.seconds[i])
              for i in range(st.npasses)]
    return [int(x) for x in did], passes, (st.total_passes, st.total_nelems_in, st.total_seconds)


class Mesh:
    """A device-resident simplex mesh (the Omega_h::Mesh of the adapt path)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.osh_mesh_destroy(self.h)
            self.h = None

    @classmethod
    def box(cls, ctx, dim, n, internal=False, x=1.0, y=1.0, z=1.0):
        h = C.c_void_p()
        f = ctx.lib.osh_build_box_internal if internal else ctx.lib.osh_build_box
        nz = n if dim == 3 else 0
        check(f(ctx.h, dim, x, y, z if dim == 3 else 0.0, n, n, nz, C.byref(h)), ctx.lib)
        return cls(ctx, h)

    @classmethod
    def from_elems(cls, ctx, dim, ev2v, nverts):
        a = np.ascontiguousarray(ev2v, dtype=np.int32)
        h = C.c_void_p()
        check(ctx.lib.osh_build_from_elems(ctx.h, dim, len(a) // (dim + 1), _p(a, C.c_int),
                                           nverts, C.byref(h)), ctx.lib)
        return cls(ctx, h)

    @classmethod
    def read_osh(cls, ctx, path):
        """Read a one-part .osh directory (Omega_h's binary format)."""
        h = C.c_void_p()
        check(ctx.lib.osh_read_osh(ctx.h, os.fsencode(path), C.byref(h)), ctx.lib)
        return cls(ctx, h)

    @classmethod
    def join(cls, ctx, parts):
        """One disjoint mesh of the parts (the batched form of adapt())."""
        arr = (C.c_void_p * len(parts))(*[p.h for p in parts])
        h = C.c_void_p()
        check(ctx.lib.osh_mesh_join(ctx.h, len(parts), arr, C.byref(h)), ctx.lib)
        return cls(ctx, h)

    def split(self, nparts):
        """The parts of a joined mesh, as new meshes."""
        arr = (C.c_void_p * nparts)()
        check(self.lib.osh_mesh_split(self.h, nparts, arr), self.lib)
        return [Mesh(self.ctx, C.c_void_p(arr[k])) for k in range(nparts)]

    def part_counts(self, nparts):
        """Entity counts of each part of a joined mesh: nparts x (dim + 1)."""
        out = np.zeros(nparts * (self.dim + 1), dtype=np.int32)
        check(self.lib.osh_mesh_part_counts(self.h, nparts, _p(out, C.c_int)), self.lib)
        return out.reshape(nparts, self.dim + 1)

    def write_osh(self, path):
        check(self.lib.osh_write_osh(self.h, os.fsencode(path)), self.lib)

    def copy(self):
        h = C.c_void_p()
        check(self.lib.osh_mesh_copy(self.h, C.byref(h)), self.lib)
        return Mesh(self.ctx, h)

    def _int(self, f, *args):
        out = C.c_int()
        check(f(self.h, *args, C.byref(out)), self.lib)
        return out.value

    @property
    def dim(self):
        return self._int(self.lib.osh_mesh_dim)

    def nents(self, d):
        return self._int(self.lib.osh_mesh_nents, d)

    def verts(self, d):
        out = np.zeros(self.nents(d) * (d + 1), dtype=np.int32)
        check(self.lib.osh_mesh_get_verts(self.h, d, _p(out, C.c_int)), self.lib)
        return out

    def down(self, d):
        n = self.nents(d) * (d + 1)
        out = np.zeros(n, dtype=np.int32)
        codes = np.zeros(n, dtype=np.int8)
        check(self.lib.osh_mesh_get_down(self.h, d, _p(out, C.c_int), _p(codes, C.c_byte)),
              self.lib)
        return out, codes

    def up(self, low, high):
        n = self._int(self.lib.osh_mesh_up_size, low, high)
        a2ab = np.zeros(self.nents(low) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        codes = np.zeros(n, dtype=np.int8)
        check(self.lib.osh_mesh_get_up(self.h, low, high, _p(a2ab, C.c_int), _p(ab2b, C.c_int),
                                       _p(codes, C.c_byte)), self.lib)
        return a2ab, ab2b, codes

    def star(self, d):
        n = self._int(self.lib.osh_mesh_star_size, d)
        a2ab = np.zeros(self.nents(d) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        check(self.lib.osh_mesh_get_star(self.h, d, _p(a2ab, C.c_int), _p(ab2b, C.c_int)),
              self.lib)
        return a2ab, ab2b

    def dual(self):
        """Mesh::ask_dual (src/Omega_h_mesh.cpp:260): elements across sides."""
        n = self._int(self.lib.osh_mesh_dual_size)
        a2ab = np.zeros(self.nents(self.dim) + 1, dtype=np.int32)
        ab2b = np.zeros(n, dtype=np.int32)
        check(self.lib.osh_mesh_get_dual(self.h, _p(a2ab, C.c_int), _p(ab2b, C.c_int)),
              self.lib)
        return a2ab, ab2b

    def sizes(self):
        """Mesh::ask_sizes (src/Omega_h_mesh.cpp:390-396): signed element sizes."""
        out = np.zeros(self.nents(self.dim), dtype=np.float64)
        check(self.lib.osh_mesh_get_sizes(self.h, _p(out, C.c_double)), self.lib)
        return out

    def coords(self):
        out = np.zeros(self.nents(0) * self.dim, dtype=np.float64)
        check(self.lib.osh_mesh_get_coords(self.h, _p(out, C.c_double)), self.lib)
        return out

    def set_metric(self, metric, ncomps):
        a = np.ascontiguousarray(metric, dtype=np.float64)
        check(self.lib.osh_mesh_set_metric(self.h, ncomps, _p(a, C.c_double)), self.lib)

    @property
    def metric_ncomps(self):
        return self._int(self.lib.osh_mesh_metric_ncomps)

    def metric(self):
        nc = self._int(self.lib.osh_mesh_metric_ncomps)
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        check(self.lib.osh_mesh_get_metric(self.h, _p(out, C.c_double)), self.lib)
        return out

    def classification(self, d):
        ids = np.zeros(self.nents(d), dtype=np.int32)
        dims = np.zeros(self.nents(d), dtype=np.int8)
        check(self.lib.osh_mesh_get_class(self.h, d, _p(ids, C.c_int), _p(dims, C.c_byte)),
              self.lib)
        return ids, dims

    def set_class(self, d, ids, dims):
        """The classification of dimension d (class_id, class_dim)."""
        i = np.ascontiguousarray(ids, dtype=np.int32)
        c = np.ascontiguousarray(dims, dtype=np.int8)
        assert len(i) == len(c) == self.nents(d)
        check(self.lib.osh_mesh_set_class(self.h, d, _p(i, C.c_int), _p(c, C.c_byte)), self.lib)

    def lengths(self):
        out = np.zeros(self.nents(1), dtype=np.float64)
        check(self.lib.osh_mesh_get_lengths(self.h, _p(out, C.c_double)), self.lib)
        return out

    def qualities(self):
        out = np.zeros(self.nents(self.dim), dtype=np.float64)
        check(self.lib.osh_mesh_get_qualities(self.h, _p(out, C.c_double)), self.lib)
        return out

    def mident_metrics(self, edges):
        """Log-Euclidean midpoint metrics of the listed edges."""
        e = np.ascontiguousarray(edges, dtype=np.int32)
        out = np.empty(max(len(e) * self.metric_ncomps, 1), dtype=np.float64)
        check(self.lib.osh_mident_metrics(self.h, len(e), _p(e, C.c_int),
                                          _p(out, C.c_double)), self.lib)
        return out[:len(e) * self.metric_ncomps]

    def swap3d_qualities(self, opts, edges):
        """(qualities, configs) of the listed interior edges of a 3D mesh."""
        e = np.ascontiguousarray(edges, dtype=np.int32)
        q = np.empty(max(len(e), 1), dtype=np.float64)
        cfg = np.empty(max(len(e), 1), dtype=np.int8)
        check(self.lib.osh_swap3d_qualities(self.h, C.byref(opts), len(e), _p(e, C.c_int),
                                            _p(q, C.c_double), _p(cfg, C.c_byte)), self.lib)
        return q[:len(e)], cfg[:len(e)]

    def reflect_down(self, high_dim, prod_verts):
        """The sides (among the mesh's entities one dimension down) and codes
        of simplices given by vertex lists."""
        pv = np.ascontiguousarray(prod_verts, dtype=np.int32)
        n = len(pv) // (high_dim + 1)
        down = np.empty(max(n * (high_dim + 1), 1), dtype=np.int32)
        codes = np.empty(max(n * (high_dim + 1), 1), dtype=np.int8)
        check(self.lib.osh_reflect_down(self.h, high_dim, n, _p(pv, C.c_int),
                                        _p(down, C.c_int), _p(codes, C.c_byte)), self.lib)
        return down[:n * (high_dim + 1)], codes[:n * (high_dim + 1)]

    def refine_qualities(self, cands):
        c = np.ascontiguousarray(cands, dtype=np.int32)
        out = np.zeros(len(c), dtype=np.float64)
        check(self.lib.osh_refine_qualities(self.h, len(c), _p(c, C.c_int),
                                            _p(out, C.c_double)), self.lib)
        return out

    def coarsen_qualities_ptr(self, ncands, cands_ptr, codes_ptr, out_ptr, precision=0,
                              min_quality=0.0):
        """coarsen_qualities on device pointers, ordered on the context's stream."""
        check(self.lib.osh_coarsen_qualities(self.h, ncands, cands_ptr, codes_ptr, out_ptr,
                                             precision, min_quality), self.lib)

    def coarsen_qualities(self, cands, codes, precision=0, min_quality=0.0):
        """coarsen_qualities (src/Omega_h_coarsen_qualities.cpp:65-88) of host
        arrays: two values per candidate edge."""
        from .hip import DeviceArray
        c = DeviceArray(np.ascontiguousarray(cands, dtype=np.int32))
        k = DeviceArray(np.ascontiguousarray(codes, dtype=np.int8))
        out = DeviceArray(dtype=np.float64, n=2 * len(cands))
        self.coarsen_qualities_ptr(len(cands), c.ptr, k.ptr, out.ptr, precision, min_quality)
        self.ctx.synchronize()
        return out.numpy()

    def find_loop(self, edge):
        eev = np.zeros(2, dtype=np.int32)
        lv = np.zeros(7, dtype=np.int32)
        n = C.c_int()
        check(self.lib.osh_find_loop(self.h, edge, _p(eev, C.c_int), _p(lv, C.c_int),
                                     C.byref(n)), self.lib)
        return eev, lv[:n.value]

    def indset(self, d, quality, cands):
        q = np.ascontiguousarray(quality, dtype=np.float64)
        c = np.ascontiguousarray(cands, dtype=np.int8)
        out = np.zeros(self.nents(d), dtype=np.int8)
        check(self.lib.osh_find_indset(self.h, d, _p(q, C.c_double), _p(c, C.c_byte),
                                       _p(out, C.c_byte)), self.lib)
        return out

    LINEAR, METRIC = 0, 1

    def add_vertex_tag(self, name, ncomps, kind, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        check(self.lib.osh_mesh_add_vertex_tag(self.h, name.encode(), ncomps, kind,
                                               _p(a, C.c_double)), self.lib)

    def vertex_tag(self, name):
        nc = self._int(self.lib.osh_mesh_vertex_tag_ncomps, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        check(self.lib.osh_mesh_get_vertex_tag(self.h, name.encode(), _p(out, C.c_double)),
              self.lib)
        return out

    def add_inherit_tag(self, name, ncomps, per_dim):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in per_dim]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        check(self.lib.osh_mesh_add_inherit_tag(self.h, name.encode(), ncomps, ptrs), self.lib)

    def inherit_tag(self, name, d):
        nc = self._int(self.lib.osh_mesh_inherit_tag_ncomps, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(d) * nc, dtype=np.float64)
        check(self.lib.osh_mesh_get_inherit_tag(self.h, name.encode(), d, _p(out, C.c_double)),
              self.lib)
        return out

    def remove_vertex_tag(self, name):
        check(self.lib.osh_mesh_remove_vertex_tag(self.h, name.encode()), self.lib)

    def implied_metrics(self, set_as_metric=False):
        """get_implied_metrics; optionally make them the mesh's metric."""
        nc = self.dim * (self.dim + 1) // 2
        out = np.zeros(self.nents(0) * nc, dtype=np.float64)
        check(self.lib.osh_get_implied_metrics(self.h, int(set_as_metric), _p(out, C.c_double)),
              self.lib)
        return out

    def approach_metric(self, opts, min_step=1e-4):
        did = C.c_int()
        check(self.lib.osh_approach_metric(self.h, C.byref(opts), min_step, C.byref(did)),
              self.lib)
        return did.value

    def warp_to_limit(self, opts, exit_on_stall=False, max_niters=40):
        """warp_to_limit (src/Omega_h_approach.cpp:36-81); 1 if it moved."""
        did = C.c_int()
        check(self.lib.osh_warp_to_limit(self.h, C.byref(opts), int(exit_on_stall),
                                         max_niters, C.byref(did)), self.lib)
        return did.value

    def set_coords(self, coords):
        a = np.ascontiguousarray(coords, dtype=np.float64)
        assert a.size == self.nents(0) * self.dim
        check(self.lib.osh_mesh_set_coords(self.h, _p(a, C.c_double)), self.lib)

    def add_elem_tag(self, name, ncomps, data):
        """An element tag carried through adapt(), pointwise (OMEGA_H_POINTWISE)
        unless set_elem_transfer names another transfer."""
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.size == self.nents(self.dim) * ncomps
        check(self.lib.osh_mesh_add_elem_tag(self.h, name.encode(), ncomps, _p(a, C.c_double)),
              self.lib)

    POINTWISE, CONSERVE, DENSITY = 0, 1, 2
    DIFFUSE_NONE, DIFFUSE_RELATIVE, DIFFUSE_ABSOLUTE = 0, 1, 2

    def set_elem_transfer(self, name, kind, diffuse=(0, 0.0, 0.0)):
        """The transfer of an element tag (POINTWISE, CONSERVE or DENSITY) and,
        for CONSERVE, the diffusion (type, tolerance, floor) of its integral's
        error (TransferOpts::type_map and integral_diffuse_map)."""
        t, tol, fl = diffuse
        check(self.lib.osh_mesh_set_elem_transfer(self.h, name.encode(), kind, t, tol, fl),
              self.lib)

    def set_momentum_velocity(self, name, density, diffuse=(0, 0.0, 0.0)):
        """A vertex tag as a momentum velocity whose momentum with the conserved
        element density `density` adapt() conserves (OMEGA_H_MOMENTUM_VELOCITY)."""
        t, tol, fl = diffuse
        check(self.lib.osh_mesh_set_momentum_velocity(self.h, name.encode(), density.encode(),
                                                      t, tol, fl), self.lib)

    def elem_tag(self, name):
        nc = self._int(self.lib.osh_mesh_elem_tag_ncomps, name.encode())
        if nc < 0:
            return None
        out = np.zeros(self.nents(self.dim) * nc, dtype=np.float64)
        check(self.lib.osh_mesh_get_elem_tag(self.h, name.encode(), _p(out, C.c_double)),
              self.lib)
        return out

    def implied_isos(self, set_as_metric=False):
        """get_implied_isos (src/Omega_h_metric.cpp:316-321)."""
        out = np.zeros(self.nents(0), dtype=np.float64)
        check(self.lib.osh_get_implied_isos(self.h, int(set_as_metric), _p(out, C.c_double)),
              self.lib)
        return out

    def ill_metric(self, opts, isotropy):
        """compute_ill_metric (src/Omega_h_fix.cpp:16-79)."""
        check(self.lib.osh_ill_metric(self.h, C.byref(opts), isotropy), self.lib)

    def fix(self, opts, isotropy):
        """fix (src/Omega_h_fix.cpp:120-134); the number of adapt() calls."""
        n = C.c_int()
        check(self.lib.osh_fix(self.h, C.byref(opts), isotropy, C.byref(n)), self.lib)
        return n.value

    def fix_for_given_metric(self, opts):
        n = C.c_int()
        check(self.lib.osh_fix_for_given_metric(self.h, C.byref(opts), C.byref(n)), self.lib)
        return n.value

    def grade_fix_adapt(self, opts, target, ncomps):
        """grade_fix_adapt (src/Omega_h_fix.cpp:136-176); adapt() calls after repair."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        assert t.size == self.nents(0) * ncomps
        n = C.c_int()
        check(self.lib.osh_grade_fix_adapt(self.h, C.byref(opts), ncomps, _p(t, C.c_double),
                                           C.byref(n)), self.lib)
        return n.value

    def digest(self):
        """FNV-1a digest of the element vertex lists and coordinate bits."""
        d = C.c_ulonglong()
        check(self.lib.osh_mesh_digest(self.h, C.byref(d)), self.lib)
        return d.value

    def limit_metric_gradation(self, max_rate, tol=1e-2):
        """Gradation limiting of the mesh's metric; returns the sweep count."""
        n = C.c_int()
        check(self.lib.osh_limit_metric_gradation(self.h, max_rate, tol, C.byref(n)), self.lib)
        return n.value

    def _pass(self, f, opts):
        did = C.c_int()
        check(f(self.h, C.byref(opts), C.byref(did)), self.lib)
        return did.value

    def refine_by_size(self, opts):
        return self._pass(self.lib.osh_refine_by_size, opts)

    def coarsen_by_size(self, opts):
        return self._pass(self.lib.osh_coarsen_by_size, opts)

    def swap_edges(self, opts):
        return self._pass(self.lib.osh_swap_edges, opts)

    def coarsen_slivers(self, opts):
        return self._pass(self.lib.osh_coarsen_slivers, opts)

    def adapt(self, opts):
        """adapt(); returns (did, passes) with the first OSH_MAX_PASSES
        passes as (kind, elements in, elements out, seconds) and keeps the
        totals over all passes as last_adapt_totals = (passes, elements in,
        seconds)."""
        st = AdaptStats()
        did = C.c_int()
        check(self.lib.osh_adapt(self.h, C.byref(opts), C.byref(st), C.byref(did)), self.lib)
        passes = [(st.kind[i], st.nelems_in[i], st.nelems_out[i], st.seconds[i])
                  for i in range(st.npasses)]
        self.last_adapt_totals = (st.total_passes, st.total_nelems_in, st.total_seconds)
        return did.value, passes


    def adapt_digests(self, opts):
        """adapt() recording four digests of the mesh after every pass:
        returns (did, passes, digests) with digests[i] = (topology and
        classification, coordinates, metric, lengths and qualities)."""
        st = AdaptStats()
        did = C.c_int()
        dg = (C.c_ulonglong * (4 * MAX_PASSES))()
        check(self.lib.osh_adapt_digests(self.h, C.byref(opts), C.byref(st), dg, C.byref(did)),
              self.lib)
        passes = [(st.kind[i], st.nelems_in[i], st.nelems_out[i], st.seconds[i])
                  for i in range(st.npasses)]
        digests = [tuple(dg[4 * i + k] for k in range(4)) for i in range(st.npasses)]
        return did.value, passes, digests

    def screen_estimates(self):
        """FP32 estimates of the edge lengths and element qualities (the
        mixed-precision screens) and the elements marked unsure."""
        l32 = np.zeros(self.nents(1), dtype=np.float64)
        q32 = np.zeros(self.nents(self.dim), dtype=np.float64)
        u = np.zeros(self.nents(self.dim), dtype=np.int8)
        check(self.lib.osh_mesh_screen_estimates(self.h, _p(l32, C.c_double),
                                                 _p(q32, C.c_double), _p(u, C.c_byte)),
              self.lib)
        return l32, q32, u

    def digests(self):
        """The same four digests of the mesh as it is."""
        dg = (C.c_ulonglong * 4)()
        check(self.lib.osh_mesh_digests(self.h, dg), self.lib)
        return tuple(dg)


class Server:
    """Continuous-batching adapt() server on one GPU (osh_server_*): nslots
    worker contexts, admission by predicted HBM footprint."""

    def __init__(self, nslots=4, hbm_budget_bytes=200e9, device=0, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.osh_server_create(device, nslots, float(hbm_budget_bytes), C.byref(h)),
              self.lib)
        self.h = h

    def submit(self, dim, n, metric, opts, want_digest=False):
        """Adapt an n-cells-per-side unit box under a constant metric
        (1, 3 or 6 values); returns a ticket."""
        m = np.ascontiguousarray(metric, dtype=np.float64)
        t = C.c_int()
        check(self.lib.osh_server_submit(self.h, dim, n, n, n if dim == 3 else 0, len(m),
                                         _p(m, C.c_double), C.byref(opts), int(want_digest),
                                         C.byref(t)), self.lib)
        return t.value

    def set_timeout(self, seconds):
        """Per-request time limit for later submissions (0: none)."""
        check(self.lib.osh_server_set_timeout(self.h, float(seconds)), self.lib)

    def wait(self, ticket):
        r = ServeResult()
        check(self.lib.osh_server_wait(self.h, ticket, C.byref(r)), self.lib)
        if r.status != 0:
            raise OshError(r.error.decode())
        return r

    def close(self):
        if getattr(self, "h", None):
            self.lib.osh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
