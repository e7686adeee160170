/* C ABI of the MI355X adapt() path (libomega_h_amd.so).
 *
 * Omega_h has no plugin ABI: callers link the library and call
 * Omega_h::adapt(Mesh*, AdaptOpts const&) (src/Omega_h_adapt.hpp:82) and the
 * pass entry points.  This header is the thin C layer those C++ entry
 * points (include/Omega_h_amd.hpp) and foreign-language callers bind to.
 * Every function returns 0 on success and a nonzero status otherwise, never
 * aborts, and leaves a message for osh_last_error().  All work is ordered on
 * the context's HIP stream; functions that hand data to the host return
 * after it has arrived.  Array arguments of the mesh, metric and pass
 * functions may point to host memory or to device memory (osh_device_alloc,
 * any hipMalloc'd block): copies let the HIP runtime tell the two apart, so
 * data that already lives on the GPU never passes through the host.
 */
#ifndef OMEGA_H_AMD_H
#define OMEGA_H_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osh_ctx osh_ctx;   /* one HIP stream on one device */
typedef struct osh_mesh osh_mesh; /* a device-resident simplex mesh */

/* Omega_h::AdaptOpts, src/Omega_h_adapt.hpp:50-77 (defaults
 * src/Omega_h_adapt.cpp:52-85) */
typedef struct {
  double min_length_desired;
  double max_length_desired;
  double max_length_allowed;
  double min_quality_allowed;
  double min_quality_desired;
  int nsliver_layers;
  int verbosity;
  int should_refine;
  int should_coarsen;
  int should_swap;
  int should_coarsen_slivers;
  /* the histograms printed at verbosity EXTRA_STATS (3),
   * print_adapt_histograms src/Omega_h_adapt.cpp:122-138 */
  double length_histogram_min;
  double length_histogram_max;
  int nlength_histogram_bins;
  int nquality_histogram_bins;
  /* vetoes collapses that would turn an exposed side against its
   * vertices' normals, prevent_coarsen_flip src/Omega_h_coarsen_flip.cpp:68-150
   * (called from src/Omega_h_coarsen.cpp:73-77) */
  int should_prevent_coarsen_flip;
  /* arithmetic of the cavity quality and length screens (no counterpart in
   * the reference): OSH_FP64, or OSH_MIXED, which screens cavities in FP32
   * and re-decides in FP64 every cavity not rejected by a margin, so the
   * adapted mesh is the same bit for bit */
  int precision;
} osh_adapt_opts;

#define OSH_FP64 0
#define OSH_MIXED 1

#define OSH_MAX_PASSES 4096
/* per-pass record of an adapt() run: kind 0 refine, 1 coarsen, 2 swap,
 * 3 coarsen slivers; the first OSH_MAX_PASSES passes are listed, and the
 * totals cover every pass */
typedef struct {
  int npasses;
  int kind[OSH_MAX_PASSES];
  int nelems_in[OSH_MAX_PASSES];
  int nelems_out[OSH_MAX_PASSES];
  double seconds[OSH_MAX_PASSES];
  long long total_passes;
  long long total_nelems_in;
  double total_seconds;
} osh_adapt_stats;

/* message of the last failed call on this thread */
const char* osh_last_error(void);

int osh_ctx_create(int device, osh_ctx** out);
int osh_ctx_destroy(osh_ctx* ctx);
int osh_ctx_synchronize(osh_ctx* ctx);
/* the hipStream_t the context orders its work on (for events and interop) */
int osh_ctx_stream(osh_ctx* ctx, void** out);
/* Device memory a caller keeps on the GPU: the storage of the C++ facade's
 * Read<T>/Write<T>, which live on the device as the reference's do in a GPU
 * build (src/Omega_h_array.hpp:21-204, allocated through Kokkos views).
 * Blocks come from the context's caching pool and are used in its stream
 * order; osh_device_free takes the size given to osh_device_alloc. */
int osh_device_alloc(osh_ctx* ctx, long long bytes, void** out);
int osh_device_free(osh_ctx* ctx, void* p, long long bytes);
/* copy `bytes` between any host or device buffers after the context's
 * earlier work; returns when the copy is done (Omega_h's deep_copy) */
int osh_copy(osh_ctx* ctx, void* dst, const void* src, long long bytes);
/* set n items of item_bytes (1, 4 or 8) bytes at dst (device memory) to
 * *value; returns when done (the fill of Write<T>(size, value)) */
int osh_device_fill(osh_ctx* ctx, void* dst, long long n, int item_bytes, const void* value);
/* *out = 1 when p is device memory */
int osh_pointer_is_device(const void* p, int* out);
/* the context's lane counts (no counterpart in the reference): "coarsen_lanes",
 * "side_lanes", "choose_lanes" and "first_lanes", each a power of two from 1
 * to 64, the threads that share one collapse side, one swap product side,
 * one swap candidate and one edge of the first independent-set round.  They
 * change no result, only speed; omega_h_amd/autotune.py picks them for a
 * workload. */
int osh_ctx_set_tuning(osh_ctx* ctx, const char* name, int value);
int osh_ctx_get_tuning(osh_ctx* ctx, const char* name, int* out);

/* Named ranges of a context, the timer tree of Omega_h::profile::History
 * with ScopedTimer/OMEGA_H_TIME_FUNCTION (src/Omega_h_profile.hpp:47-232),
 * per context instead of the reference's process-wide singleton.  flags: 1
 * keeps the tree, 2 emits roctx ranges (rocprofv3 --marker-trace), 4 makes
 * each range wait for the context's stream at its end so that its time
 * covers the device work it enqueued (profiling runs only: it serialises the
 * host and the device).  Ranges are named after the reference's timed
 * functions (adapt, satisfy_lengths, satisfy_quality, refine_by_size,
 * coarsen_by_size, swap_edges, coarsen_slivers, ...).  Flags change only
 * while no range is open.  The environment variables OSH_TIME, OSH_TIME_PERCENT
 * and OSH_TIME_CHOP stand for Omega_h::Library's --osh-time options
 * (src/Omega_h_library.cpp:99-131): a context opened with OSH_TIME=1 profiles
 * and prints its listings when it closes. */
int osh_ctx_set_profile(osh_ctx* ctx, int flags);
/* forget every range recorded so far */
int osh_ctx_profile_clear(osh_ctx* ctx);
/* the tree's frames: *parent is -1 for a root; seconds and calls sum over
 * every time the range opened under the same chain of ranges */
int osh_ctx_profile_nframes(osh_ctx* ctx, int* n);
int osh_ctx_profile_frame(osh_ctx* ctx, int i, int* parent, double* seconds, long long* calls,
    char* name, int name_cap);
/* the TOP-DOWN and BOTTOM-UP listings of print_top_down_and_bottom_up
 * (src/Omega_h_profile.cpp:151-214) as text; *len receives the full length
 * (write at most cap bytes, the last a NUL).  chop: leave out lines under
 * this percentage of the roots' total; percent: times as percentages. */
int osh_ctx_profile_report(osh_ctx* ctx, double chop, int percent, char* buf, long long cap,
    long long* len);

/* AdaptOpts::AdaptOpts(Int dim), src/Omega_h_adapt.cpp:52-85 */
int osh_default_opts(int dim, osh_adapt_opts* out);

/* build_box(comm, SIMPLEX, x, y, z, nx, ny, nz), src/Omega_h_build.cpp:136-149:
 * box, simplex split, Hilbert reordering and box classification; nz = 0
 * builds a 2D box */
int osh_build_box(osh_ctx* ctx, int dim, double x, double y, double z, int nx,
    int ny, int nz, osh_mesh** out);
/* build_box_internal (no reordering, no classification),
 * src/Omega_h_build.cpp:107-134 */
int osh_build_box_internal(osh_ctx* ctx, int dim, double x, double y, double z,
    int nx, int ny, int nz, osh_mesh** out);
/* build_from_elems2verts, src/Omega_h_build.cpp:84-97 (host vertex lists) */
int osh_build_from_elems(osh_ctx* ctx, int dim, int nelems, const int* ev2v,
    int nverts, osh_mesh** out);
int osh_mesh_destroy(osh_mesh* mesh);
int osh_mesh_copy(osh_mesh* mesh, osh_mesh** out);

/* The batched form of adapt() (no counterpart in the reference, whose
 * adapt() takes one mesh): osh_mesh_join makes one disjoint mesh of nparts
 * meshes of the context (same dimension and metric layout, classified, no
 * inheritance or element tags), part k's entities of every dimension
 * numbered after those of parts < k.  Every pass and osh_adapt on the joined
 * mesh does to each part what it does to that part alone (all decisions are
 * cavity-local and new entities are numbered in old order), with one kernel
 * launch for all parts.  osh_mesh_split gives back nparts new meshes (the
 * joined one stays valid); osh_mesh_part_counts writes nparts * (dim + 1)
 * entity counts.  Both fail unless every part is still one block. */
int osh_mesh_join(osh_ctx* ctx, int nparts, osh_mesh* const* parts, osh_mesh** out);
int osh_mesh_split(osh_mesh* joined, int nparts, osh_mesh** out);
int osh_mesh_part_counts(osh_mesh* joined, int nparts, int* counts);
/* adapt() of nmeshes meshes of one context at once (the batched form, no
 * counterpart in the reference): they are joined as above and adapted as
 * one, and whenever a quarter of the parts can no longer change in the
 * current phase (unchanged over a run of passes holding every enabled pass
 * kind, or in the quality phase already at min_quality_desired) the joined
 * mesh is split and the rest joined again.  Each mesh ends exactly where
 * osh_adapt alone leaves it; did[i] is osh_adapt's result for mesh i, and
 * stats lists the passes of the joined meshes.  Same requirements as
 * osh_mesh_join; the options' verbosity is ignored (silent). */
int osh_adapt_batch(int nmeshes, osh_mesh* const* meshes, const osh_adapt_opts* opts,
    osh_adapt_stats* stats, int* did);

/* Mesh::dim / Mesh::nents, src/Omega_h_mesh.hpp */
int osh_mesh_dim(osh_mesh* mesh, int* out);
int osh_mesh_nents(osh_mesh* mesh, int ent_dim, int* out);
/* Mesh::ask_verts_of(dim), host copy of nents*(dim+1) */
int osh_mesh_get_verts(osh_mesh* mesh, int ent_dim, int* out);
/* Mesh::ask_down(dim, dim-1): entities and alignment codes */
int osh_mesh_get_down(osh_mesh* mesh, int ent_dim, int* out, signed char* codes);
/* Mesh::ask_up(low, high): CSR offsets, entities, codes */
int osh_mesh_up_size(osh_mesh* mesh, int low, int high, int* out);
int osh_mesh_get_up(osh_mesh* mesh, int low, int high, int* a2ab, int* ab2b,
    signed char* codes);
/* Mesh::ask_star(dim) */
int osh_mesh_star_size(osh_mesh* mesh, int ent_dim, int* out);
int osh_mesh_get_star(osh_mesh* mesh, int ent_dim, int* a2ab, int* ab2b);
/* Mesh::ask_dual() (src/Omega_h_mesh.cpp:260, elements_across_sides in
 * src/Omega_h_adj.cpp:591-657): CSR of the elements across each element's
 * sides, in side order; a side shared by more than two elements is an error */
int osh_mesh_dual_size(osh_mesh* mesh, int* out);
int osh_mesh_get_dual(osh_mesh* mesh, int* a2ab, int* ab2b);
/* Mesh::ask_sizes() (src/Omega_h_mesh.cpp:390-396, measure_elements_real in
 * src/Omega_h_shape.cpp:55-89): signed element areas (2D) or volumes (3D) */
int osh_mesh_get_sizes(osh_mesh* mesh, double* out);
/* tags "coordinates", "metric", "class_id"/"class_dim", "length", "quality" */
int osh_mesh_get_coords(osh_mesh* mesh, double* out);
int osh_mesh_set_metric(osh_mesh* mesh, int ncomps, const double* metric);
int osh_mesh_metric_ncomps(osh_mesh* mesh, int* out);
int osh_mesh_get_metric(osh_mesh* mesh, double* out);
int osh_mesh_get_class(osh_mesh* mesh, int ent_dim, int* ids, signed char* dims);
/* Mesh::set_tag(ent_dim, "class_id" / "class_dim", ...)
 * (src/Omega_h_mesh.cpp:167-177): a caller's own classification of one
 * dimension (dims[i] between ent_dim and the mesh dimension) */
int osh_mesh_set_class(osh_mesh* mesh, int ent_dim, const int* ids, const signed char* dims);
/* Mesh::ask_lengths / ask_qualities, src/Omega_h_mesh.cpp:374-388 */
int osh_mesh_get_lengths(osh_mesh* mesh, double* out);
int osh_mesh_get_qualities(osh_mesh* mesh, double* out);
/* sort_by_keys (src/Omega_h_sort.hpp:8-22, src/Omega_h_sort.cpp:43-92):
 * perm[i] is the index of the i-th smallest of n tuples of `width` (1 to 3)
 * signed 64-bit keys in lexical order; the sort is stable (equal tuples keep
 * their order).  The mesh builders use it (find_unique, Hilbert order). */
int osh_sort_by_keys(osh_ctx* ctx, int n, int width, const long long* keys, int* perm);
/* Mesh::min_quality / max_length (src/Omega_h_mesh.cpp:698-707): reduced on
 * the device */
int osh_mesh_min_quality(osh_mesh* mesh, double* out);
int osh_mesh_max_length(osh_mesh* mesh, double* out);

/* passes: refine_by_size src/Omega_h_refine.cpp:92, coarsen_by_size
 * src/Omega_h_coarsen.cpp:198, swap_edges src/Omega_h_swap.cpp:45,
 * coarsen_slivers src/Omega_h_coarsen.cpp:210; *did = 1 if the mesh changed */
int osh_refine_by_size(osh_mesh* mesh, const osh_adapt_opts* opts, int* did);
int osh_coarsen_by_size(osh_mesh* mesh, const osh_adapt_opts* opts, int* did);
int osh_swap_edges(osh_mesh* mesh, const osh_adapt_opts* opts, int* did);
int osh_coarsen_slivers(osh_mesh* mesh, const osh_adapt_opts* opts, int* did);
/* adapt(Mesh*, AdaptOpts const&), src/Omega_h_adapt.cpp:272-288;
 * *did = 0 when the mesh already met the goals; stats may be NULL */
int osh_adapt(osh_mesh* mesh, const osh_adapt_opts* opts, osh_adapt_stats* stats,
    int* did);

/* osh_adapt that also records the mesh after every pass as four FNV-1a
 * digests (no counterpart in the reference; for comparing runs pass by
 * pass): digests[4*i + k] after pass i, k = 0 the vertex lists of every
 * dimension 1..dim (int32) then class_id (int32) and class_dim (int8) of
 * every dimension 0..dim, 1 the coordinate bits, 2 the metric bits, 3 the
 * edge length bits then the element quality bits.  digests holds
 * 4 * OSH_MAX_PASSES values; stats is required. */
int osh_adapt_digests(osh_mesh* mesh, const osh_adapt_opts* opts, osh_adapt_stats* stats,
    unsigned long long* digests, int* did);
/* the FP32 estimates that the mixed-precision form (precision OSH_MIXED)
 * screens cavities with, for every edge length (nedges values) and element
 * quality (nelems values), and the elements whose max-determinant vertex
 * metric is not clear in FP32 (unsure = 1; the screen re-decides those in
 * FP64); no counterpart in the reference, for reporting the estimates'
 * error against osh_mesh_get_lengths / osh_mesh_get_qualities */
int osh_mesh_screen_estimates(osh_mesh* mesh, double* lengths32, double* qualities32,
    signed char* unsure);
/* the same four digests of the mesh as it is */
int osh_mesh_digests(osh_mesh* mesh, unsigned long long* out);

/* measure_edges_metric, src/Omega_h_shape.cpp:7-48, on device pointers:
 * out[i] = metric length of edge edges[i] (or i when edges is NULL).
 * Asynchronous: ordered on the context's stream (osh_ctx_stream), so the
 * caller synchronises before reading out.  ev2v must be 8-byte aligned and a
 * 6-component metric 16-byte aligned (any hipMalloc pointer is). */
int osh_measure_edges_metric(osh_ctx* ctx, int dim, int metric_ncomps, int n,
    const int* edges, const int* ev2v, const double* coords,
    const double* metric, double* out);
/* measure_qualities, src/Omega_h_quality.cpp:7-41, on device pointers;
 * asynchronous like osh_measure_edges_metric; a tet vertex list must be
 * 16-byte aligned */
int osh_measure_qualities(osh_ctx* ctx, int dim, int metric_ncomps, int n,
    const int* elems, const int* ev2v, const double* coords,
    const double* metric, double* out);

/* offset_scan(Read<I32>), src/Omega_h_int_scan.cpp:10-20, for njobs (1 to 12)
 * arrays at once on device pointers: out[j] (n[j] + 1 ints) receives 0
 * followed by the running sums of in[j] (n[j] ints).  Asynchronous like
 * osh_measure_edges_metric. */
int osh_offset_scans(osh_ctx* ctx, int njobs, const int* const* in, const int* n,
    int* const* out);
/* osh_offset_scans without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The look-back
 * state lives in a fixed 128 KiB scratch array per GPU, set up when a context
 * is created on that GPU; one call holds at most 16384 tiles of 4096 values
 * (67,108,864 values over all its jobs), and a larger call is refused with a
 * nonzero status before anything runs.  Calls that overlap on one GPU (two
 * streams, two threads) race for the scratch; the next sync of a context then
 * reports a scan error.  Same arguments, results and stream ordering as
 * osh_offset_scans; errors found on the device surface at the next
 * osh_ctx_synchronize. */
int osh_offset_scans_noalloc(osh_ctx* ctx, int njobs, const int* const* in, const int* n,
    int* const* out);

/* limit_metric_gradation(mesh, metric, max_rate, tol),
 * src/Omega_h_metric.cpp:200-220, applied to the mesh's metric in place
 * (each sweep intersects every vertex metric with its neighbours' scaled
 * metrics, src/Omega_h_metric.cpp:156-180, until two sweeps are close within
 * tol); *nsteps = sweeps taken */
int osh_limit_metric_gradation(osh_mesh* mesh, double max_rate, double tol, int* nsteps);
/* intersect_metrics, src/Omega_h_metric_intersect.hpp:215-269 (with the
 * degenerate cases :18-213), for n pairs of symmetric metrics of dimension
 * dim (1, 2 or 3) in host arrays */
int osh_intersect_metrics(osh_ctx* ctx, int dim, int n, const double* m1, const double* m2,
    double* out);

/* vertex tags carried through adapt(): Mesh::add_tag(VERT, name, ncomps,
 * array) with the transfer type TransferOpts gives it
 * (src/Omega_h_adapt.hpp:30-42, src/Omega_h_transfer.cpp:391-428):
 * kind 0 = OMEGA_H_LINEAR_INTERP (midpoint average), 1 = OMEGA_H_METRIC
 * (log-Euclidean midpoint); surviving vertices keep their values.  The tag
 * "target_metric" drives osh_approach_metric. */
int osh_mesh_add_vertex_tag(osh_mesh* mesh, const char* name, int ncomps, int kind,
    const double* data);
int osh_mesh_vertex_tag_ncomps(osh_mesh* mesh, const char* name, int* out); /* -1: absent */
int osh_mesh_get_vertex_tag(osh_mesh* mesh, const char* name, double* out);
int osh_mesh_remove_vertex_tag(osh_mesh* mesh, const char* name);
/* tags on every dimension transferred by inheritance (OMEGA_H_INHERIT,
 * should_inherit src/Omega_h_transfer.cpp:20-34): products take the value of
 * the entity they came from, as class_id does; per_dim[d] holds
 * nents(d) * ncomps values for d = 0..dim */
int osh_mesh_add_inherit_tag(osh_mesh* mesh, const char* name, int ncomps,
    const double* const* per_dim);
int osh_mesh_inherit_tag_ncomps(osh_mesh* mesh, const char* name, int* out); /* -1: absent */
int osh_mesh_get_inherit_tag(osh_mesh* mesh, const char* name, int ent_dim, double* out);
/* element tags carried through adapt() pointwise: Mesh::add_tag(dim, name,
 * ncomps, array) with OMEGA_H_POINTWISE as its transfer type
 * (src/Omega_h_transfer.cpp:47-54,515-583): products of refinement take
 * their parent element's value, products of coarsening and swapping the
 * value of the old element of their cavity in which their centroid lies
 * deepest; surviving elements keep theirs.  data: nelems * ncomps values */
int osh_mesh_add_elem_tag(osh_mesh* mesh, const char* name, int ncomps, const double* data);
int osh_mesh_elem_tag_ncomps(osh_mesh* mesh, const char* name, int* out); /* -1: absent */
int osh_mesh_get_elem_tag(osh_mesh* mesh, const char* name, double* out);
int osh_mesh_remove_elem_tag(osh_mesh* mesh, const char* name);
/* the transfer of an element tag through adapt() (TransferOpts::type_map,
 * src/Omega_h_adapt.hpp:30-42): kind 0 = OMEGA_H_POINTWISE (the default),
 * 1 = OMEGA_H_CONSERVE (a density whose integral adapt() conserves:
 * src/Omega_h_conserve.cpp:372-537 through the passes and
 * correct_integral_errors :799-829 at the end), 2 = OMEGA_H_DENSITY (a
 * density transferred by intersection, transfer_by_intersection
 * :391-455).  For kind 1, how the integral's error is diffused over the
 * elements before the correction (TransferOpts::integral_diffuse_map, a
 * VarCompareOpts src/Omega_h_compare.hpp:14-21): diffuse_type 0 none,
 * 1 relative, 2 absolute, with its tolerance and floor.  Re-adding the
 * tag's values keeps its transfer. */
int osh_mesh_set_elem_transfer(osh_mesh* mesh, const char* name, int kind, int diffuse_type,
    double tol, double floor);
/* a vertex tag of dim components as a momentum velocity
 * (OMEGA_H_MOMENTUM_VELOCITY): interpolated like kind 0, and its momentum
 * with the conserved element density `density` (TransferOpts::
 * velocity_density_map) is conserved through adapt(), its error diffused as
 * above (integral_diffuse_map of its velocity_momentum_map name) */
int osh_mesh_set_momentum_velocity(osh_mesh* mesh, const char* name, const char* density,
    int diffuse_type, double tol, double floor);
/* approach_metric(mesh, opts, min_step), src/Omega_h_approach.cpp:83-126:
 * set the metric to the "target_metric" tag, or halve the log-Euclidean
 * step toward it until the mesh's quality and lengths are acceptable;
 * *did = 0 when there is no target.  The reference's idiom
 * (src/aniso_test.cpp:15-27) is: while (approach_metric) adapt. */
int osh_approach_metric(osh_mesh* mesh, const osh_adapt_opts* opts, double min_step, int* did);
/* warp_to_limit(mesh, opts, exit_on_stall, max_niters),
 * src/Omega_h_approach.cpp:36-81: move the vertices by the "warp" vertex tag
 * (dim components; add it as a linear tag so adapt() carries it), halving
 * the move until quality and lengths are acceptable; the part not yet moved
 * stays in the tag, which is removed once the whole warp is done.  *did = 0
 * when there is no warp tag.  An unacceptable starting mesh is an error, and
 * so are more than max_niters halvings unless exit_on_stall is set (then the
 * tag is dropped).  The reference's idiom (src/bend_test.cpp:55-65) is:
 * do { adapt } while (warp_to_limit). */
int osh_warp_to_limit(osh_mesh* mesh, const osh_adapt_opts* opts, int exit_on_stall,
    int max_niters, int* did);
/* Mesh::set_coords (src/Omega_h_mesh.cpp:365-368): nverts * dim values;
 * drops the cached lengths, qualities and sizes */
int osh_mesh_set_coords(osh_mesh* mesh, const double* data);
/* get_implied_metrics(mesh), src/Omega_h_metric.cpp:280-309: the metric each
 * element implies (all its edges of unit length, src/Omega_h_shape.hpp:240-273),
 * scaled to a typical unit element and projected to the vertices by
 * log-Euclidean averaging; out holds nverts * (3 in 2D, 6 in 3D) values, and
 * with set_as_metric != 0 it also becomes the mesh's metric */
int osh_get_implied_metrics(osh_mesh* mesh, int set_as_metric, double* out);
/* get_implied_isos(mesh), src/Omega_h_metric.cpp:316-321 (the implied metric
 * made isotropic with the same volume; add_implied_isos_tag,
 * src/Omega_h_metric_input.cpp:276-280, when set_as_metric != 0): nverts values */
int osh_get_implied_isos(osh_mesh* mesh, int set_as_metric, double* out);
/* compute_ill_metric(mesh, opts, isotropy), src/Omega_h_fix.cpp:16-79: the
 * metric implied by the elements (isotropy 0 OMEGA_H_ANISOTROPIC, smoothed
 * five times; 1 OMEGA_H_ISO_LENGTH, the largest eigenvalue; 2
 * OMEGA_H_ISO_SIZE), without the elements below opts->min_quality_allowed,
 * graded at rate 1, becomes the mesh's metric (a stage of osh_fix, static in
 * the reference) */
int osh_ill_metric(osh_mesh* mesh, const osh_adapt_opts* opts, int isotropy);
/* fix_for_given_metric(mesh, opts, verbose), src/Omega_h_fix.cpp:90-118:
 * adapt() with the allowed quality raised to the mesh's own minimum until
 * that minimum stops improving by 1e-3; *ncalls = adapt() calls made */
int osh_fix_for_given_metric(osh_mesh* mesh, const osh_adapt_opts* opts, int* ncalls);
/* fix(mesh, opts, isotropy, verbose), src/Omega_h_fix.cpp:120-134:
 * osh_ill_metric then osh_fix_for_given_metric */
int osh_fix(osh_mesh* mesh, const osh_adapt_opts* opts, int isotropy, int* ncalls);
/* grade_fix_adapt(mesh, opts, target_metric, verbose),
 * src/Omega_h_fix.cpp:136-176: the target (nverts * ncomps values, ncomps 1
 * or full) graded at rate 0.5 becomes "target_metric"; a mesh without a
 * metric gets the implied one (isotropic for a scalar target); a mesh below
 * the allowed quality is repaired by osh_fix; then
 * while (approach_metric) adapt.  *nsteps = adapt() calls of that loop */
int osh_grade_fix_adapt(osh_mesh* mesh, const osh_adapt_opts* opts, int ncomps,
    const double* target, int* nsteps);
/* interpolate_between_metrics, src/Omega_h_metric.cpp:106-111, for n pairs of
 * metrics of dimension dim in host arrays */
int osh_interpolate_metrics(osh_ctx* ctx, int dim, int n, const double* a, const double* b,
    double t, double* out);

/* binary::write(path, mesh) / binary::read(path, ...), src/Omega_h_file.cpp:
 * 515-600: the .osh directory format (version 9, zlib-compressed arrays),
 * one part.  Written: down adjacencies with codes, coordinates, metric,
 * vertex tags, class_id, class_dim, global.  Read: the same; other float64
 * vertex tags become linear vertex tags, other tags are skipped. */
int osh_write_osh(osh_mesh* mesh, const char* path);
int osh_read_osh(osh_ctx* ctx, const char* path, osh_mesh** out);

/* FNV-1a digest of the element vertex lists and coordinate bits (for
 * comparing adapted meshes) */
int osh_mesh_digest(osh_mesh* mesh, unsigned long long* out);

/* Request server (no counterpart in the reference, whose adapt() serves one
 * mesh per call): adapt() requests on unit boxes with a constant metric,
 * run by nslots worker contexts (one HIP stream each) so concurrent
 * requests overlap on the GPU; admitted in arrival order when a slot is free
 * and the predicted footprint (1 KiB per element of the larger of the input
 * and the metric's expected element count) fits the HBM budget beside the
 * running requests.  osh_server_wait blocks until the ticket is done. */
typedef struct osh_server osh_server;
typedef struct {
  int status;                 /* 0 ok, 1 failed (see error) */
  int ticket;
  int nelems_in, nelems_out;  /* elements of the box and of the result */
  int npasses;                /* mesh-changing passes */
  long long pass_elems;       /* elements entering those passes */
  double pass_seconds;        /* their summed wall time */
  double queue_seconds;       /* submit to start */
  double run_seconds;         /* start to done */
  double latency_seconds;     /* submit to done */
  unsigned long long digest;  /* osh_mesh_digest of the result, if asked */
  char error[200];
} osh_serve_result;
int osh_server_create(int device, int nslots, double hbm_budget_bytes, osh_server** out);
int osh_server_destroy(osh_server* server);
/* metric: metric_ncomps values of the constant metric (1, or 3 in 2D / 6 in
 * 3D in Omega_h's symmetric order); nz = 0 for a 2D box */
int osh_server_submit(osh_server* server, int dim, int nx, int ny, int nz, int metric_ncomps,
    const double* metric, const osh_adapt_opts* opts, int want_digest, int* ticket);
int osh_server_wait(osh_server* server, int ticket, osh_serve_result* out);
/* per-request time limit for requests submitted afterwards (0: none): a
 * request fails with status 1 and frees its slot once `seconds` have passed
 * since its submission, checked before it starts and after every pass (a
 * pass is not interrupted) */
int osh_server_set_timeout(osh_server* server, double seconds);

/* stage kernels on caller arrays (host or device), for callers that drive
 * the stages themselves and for tests:
 * get_mident_metrics (src/Omega_h_metric.cpp:57-90): the log-Euclidean
 * midpoint metric of each listed edge (metric_ncomps values per edge) */
int osh_mident_metrics(osh_mesh* mesh, int n, const int* edges, double* out);
/* swap3d_qualities (src/Omega_h_swap3d_qualities.cpp:11-75) with the choice
 * of src/Omega_h_swap3d_choice.hpp: for each listed interior edge of a 3D
 * mesh (no edge twice) the chosen triangulation of its loop (-1 for none)
 * and the quality of that triangulation when it improves on the edge's
 * elements, else -1 (what swap_edges keeps).  Honours opts->precision: in
 * the mixed form a candidate the FP32 screen rejects reports -1 for both. */
int osh_swap3d_qualities(osh_mesh* mesh, const osh_adapt_opts* opts, int n, const int* edges,
    double* quals, signed char* configs);
/* invert_adj (src/Omega_h_adj.cpp:231-263) of a high -> low map with deg
 * entries per high entity: a2ab (nlow + 1 offsets), ab2b and codes (nhigh *
 * deg uses); rows are sorted by high index and a code carries the down
 * code's flip and rotation (down_codes may be null) and the use's position */
int osh_invert_adj(osh_ctx* ctx, int nhigh, int deg, int nlow, const int* hl2l,
    const signed char* down_codes, int* a2ab, int* ab2b, signed char* codes);
/* reflect_down (src/Omega_h_adj.cpp:424-441): for simplices of dimension
 * high_dim given by vertex lists, their sides among the mesh's entities of
 * dimension high_dim - 1 and the alignment codes; every side must exist */
int osh_reflect_down(osh_mesh* mesh, int high_dim, int nprods, const int* prod_verts,
    int* down, signed char* codes);
/* collect_marked (src/Omega_h_map.cpp:174-185): the indices of
 * the nonzero marks in order; out has room for n, *count receives how many */
int osh_collect_marked(osh_ctx* ctx, int n, const signed char* marks, int* out, int* count);
/* refine_qualities src/Omega_h_refine_qualities.cpp:88-108 */
int osh_refine_qualities(osh_mesh* mesh, int n, const int* cands, double* out);
/* coarsen_qualities(mesh, cands2edges, cand_codes),
 * src/Omega_h_coarsen_qualities.cpp:65-88, on device pointers: out[2i + col]
 * is the minimum quality of the elements around the collapsing end of edge
 * cands[i] (direction col: bit col of cand_codes[i], as in
 * src/Omega_h_collapse.hpp:11-28) after the collapse, -1 where that
 * direction is not asked for.  Asynchronous like osh_measure_edges_metric
 * (the first call on a mesh builds its vertex -> element adjacency and
 * waits).  No two candidates may name the same edge.  precision OSH_MIXED
 * screens in FP32 first and leaves -1 where the estimate is below
 * min_quality by more than the screening band (1e-3 relative); every other
 * value is the FP64 one. */
int osh_coarsen_qualities(osh_mesh* mesh, int ncands, const int* cands,
    const signed char* cand_codes, double* out, int precision, double min_quality);
/* swap3d::find_loop src/Omega_h_swap3d_loop.hpp:24-78 (loop_verts holds 7) */
int osh_find_loop(osh_mesh* mesh, int edge, int* eev2v, int* loop_verts,
    int* size);
/* find_indset src/Omega_h_indset.cpp:27-35 */
int osh_find_indset(osh_mesh* mesh, int ent_dim, const double* quality,
    const signed char* cands, signed char* out);
/* the edge-swap triangulation tables in use (the data of
 * src/Omega_h_swap3d_tables.hpp:8-300) for one loop size 3..7:
 * counts = {unique triangles, triangulations, unique interior edges,
 * triangles per triangulation}; tris holds 3 loop vertices per unique
 * triangle, meshes (size-2) unique-triangle indices per triangulation, uedges
 * 2 loop vertices per unique interior edge, medges (size-3) unique-edge
 * indices per triangulation; any output may be NULL.  Host data only: needs
 * no GPU. */
int osh_swap3d_tables(int loop_size, int* counts, int* tris, int* meshes, int* uedges,
    int* medges);
/* device evaluation of the scalar transcendentals: which = 0 log, 1 exp,
 * 2 cbrt, 3 acos, 4 cos */
int osh_eval_math(osh_ctx* ctx, int which, int n, const double* in, double* out);

#ifdef __cplusplus
}
#endif

#endif
