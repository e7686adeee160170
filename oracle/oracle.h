/* Oracle (test infrastructure only): C API of the CPU restatement of
 * Omega_h's adapt() hot path, loaded by tests/ through ctypes.  The product
 * (omega_h_amd) never includes, links or calls this. */
#ifndef OMEGA_H_ORACLE_H
#define OMEGA_H_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct om_mesh om_mesh;

/* mirrors Omega_h::AdaptOpts (src/Omega_h_adapt.hpp:50-77) */
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
  int should_prevent_coarsen_flip;
} om_adapt_opts;

#define OM_MAX_PASSES 4096
typedef struct {
  int npasses;
  int kind[OM_MAX_PASSES]; /* 0 refine, 1 coarsen, 2 swap, 3 coarsen slivers */
  int nelems_in[OM_MAX_PASSES];
  int nelems_out[OM_MAX_PASSES];
  double seconds[OM_MAX_PASSES];
} om_adapt_stats;

void om_default_opts(int dim, om_adapt_opts* o);
om_mesh* om_build_box(int dim, double x, double y, double z, int nx, int ny,
    int nz);
om_mesh* om_build_box_internal(int dim, double x, double y, double z, int nx,
    int ny, int nz);
om_mesh* om_build_from_elems(int dim, int nelems, int const* ev2v, int nverts);
void om_mesh_free(om_mesh* m);
/* one disjoint mesh of nparts meshes, part k's entities after those of
 * parts < k, the vertex tag "part" holding each vertex's part */
om_mesh* om_mesh_join(int nparts, om_mesh* const* parts);
/* a mesh from its arrays (down adjacencies and codes per dimension 1..dim,
 * coordinates, metric, classification per dimension 0..dim) */
om_mesh* om_mesh_from_arrays(int dim, int const* nents, int const* const* down,
    signed char const* const* codes, double const* coords, int metric_ncomps,
    double const* metric, int const* const* class_id, signed char const* const* class_dim);
int om_mesh_dim(om_mesh* m);
int om_mesh_nents(om_mesh* m, int d);
void om_get_verts(om_mesh* m, int d, int* out);
void om_get_down(om_mesh* m, int d, int* out, signed char* codes);
void om_get_coords(om_mesh* m, double* out);
void om_set_metric(om_mesh* m, int ncomps, double const* metric);
int om_metric_ncomps(om_mesh* m);
void om_get_metric(om_mesh* m, double* out);
void om_get_class(om_mesh* m, int d, int* ids, signed char* dims);
void om_get_lengths(om_mesh* m, double* out);
void om_get_qualities(om_mesh* m, double* out);
int om_up_size(om_mesh* m, int low, int high);
void om_get_up(om_mesh* m, int low, int high, int* a2ab, int* ab2b,
    signed char* codes);
int om_dual_size(om_mesh* m);
void om_get_dual(om_mesh* m, int* a2ab, int* ab2b);
void om_get_sizes(om_mesh* m, double* out);
int om_star_size(om_mesh* m, int d);
void om_get_star(om_mesh* m, int d, int* a2ab, int* ab2b);

int om_refine_by_size(om_mesh* m, om_adapt_opts const* o);
int om_coarsen_by_size(om_mesh* m, om_adapt_opts const* o);
int om_swap_edges(om_mesh* m, om_adapt_opts const* o);
int om_coarsen_slivers(om_mesh* m, om_adapt_opts const* o);
/* conservative transfer: kind 0 pointwise, 1 conserve, 2 density;
 * diffuse_type 0 none, 1 relative, 2 absolute */
int om_set_etag_transfer(om_mesh* m, char const* name, int kind, int diffuse_type, double tol,
    double floor);
int om_set_momentum_velocity(om_mesh* m, char const* name, char const* density,
    int diffuse_type, double tol, double floor);
int om_adapt(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* stats);
/* om_adapt that also records four FNV-1a digests of the mesh after every
 * pass (digests holds 4 * OM_MAX_PASSES values): 0 vertex lists of every
 * dimension and classification, 1 coordinate bits, 2 metric bits, 3 edge
 * length and element quality bits */
int om_adapt_digests(om_mesh* m, om_adapt_opts const* o, om_adapt_stats* stats,
    uint64_t* digests);
void om_mesh_digests(om_mesh* m, uint64_t* out);

/* known-answer hooks for the reference's unit tests */
double om_log(double x);
double om_exp(double x);
double om_cbrt(double x);
double om_acos(double x);
double om_cos(double x);
void om_kat_hilbert(uint64_t* X, int b, int n, uint64_t* Y);
int om_kat_roots3(double const* coeffs, double eps, double* roots, int* mults);
void om_kat_decompose_eigen(int dim, double const* m, double* q, double* l);
void om_kat_log_spd(int dim, double const* m, double* out);
void om_kat_exp_spd(int dim, double const* m, double* out);
double om_kat_quality(int dim, int mdim, double const* pts, double const* ms);
double om_kat_edge_length(int sdim, int mdim, double const* p0,
    double const* p1, double const* m0, double const* m1);
double om_kat_aniso_length(double la, double lb);
void om_kat_average_metric(int mdim, int n, double const* ms, double* out);
double om_kat_max_eigenvalue(int dim, double const* symm);
void om_kat_intersect(int dim, double const* m1, double const* m2, double* out);
/* vertex tags carried through adapt(): kind 0 linear interpolation, 1 metric */
int om_add_vtag(om_mesh* m, char const* name, int ncomps, int kind, double const* data);
int om_vtag_ncomps(om_mesh* m, char const* name);
void om_get_vtag(om_mesh* m, char const* name, double* out);
void om_remove_vtag(om_mesh* m, char const* name);
/* inheritance tags: per_dim[d] holds nents(d) * ncomps values, d = 0..dim */
int om_add_itag(om_mesh* m, char const* name, int ncomps, double const* const* per_dim);
int om_itag_ncomps(om_mesh* m, char const* name);
void om_get_itag(om_mesh* m, char const* name, int d, double* out);
void om_interpolate_metrics(int mdim, int n, double const* a, double const* b, double t,
    double* out);
int om_approach_metric(om_mesh* m, om_adapt_opts const* o, double min_step);
void om_set_coords(om_mesh* m, double const* coords);
/* element tags transferred pointwise (OMEGA_H_POINTWISE) */
int om_add_etag(om_mesh* m, char const* name, int ncomps, double const* data);
int om_etag_ncomps(om_mesh* m, char const* name);
void om_get_etag(om_mesh* m, char const* name, double* out);
void om_kat_barycentric(int dim, double const* x, double const* simplex, double* xi);
int om_warp_to_limit(om_mesh* m, om_adapt_opts const* o, int exit_on_stall, int max_niters);
/* get_implied_metrics: nverts * (3 or 6) values */
void om_get_implied_metrics(om_mesh* m, double* out);
/* get_implied_isos: nverts values */
void om_get_implied_isos(om_mesh* m, double* out);
/* compute_ill_metric / fix_for_given_metric / fix / grade_fix_adapt
 * (isotropy 0 anisotropic, 1 iso length, 2 iso size) */
void om_ill_metric(om_mesh* m, om_adapt_opts const* o, int isotropy);
int om_fix_for_given_metric(om_mesh* m, om_adapt_opts const* o);
int om_fix(om_mesh* m, om_adapt_opts const* o, int isotropy);
int om_grade_fix_adapt(om_mesh* m, om_adapt_opts const* o, int ncomps, double const* target);
void om_element_implied_metric(int dim, double const* const* p, double* out);
/* limit_metric_gradation on the mesh's metric; returns the sweep count */
int om_limit_metric_gradation(om_mesh* m, double max_rate, double tol);
void om_kat_refine_qualities(om_mesh* m, int n, int const* cands, double* out);
/* coarsen_qualities(mesh, cands2edges, cand_codes): two values per candidate */
void om_kat_coarsen_qualities(om_mesh* m, int n, int const* cands, signed char const* codes,
    double* out);
int om_kat_find_loop(om_mesh* m, int edge, int* eev2v, int* loop_verts);
void om_kat_indset(om_mesh* m, int d, double const* q, signed char const* c,
    signed char* out);
/* topology known answers (src/unit_mesh.cpp, src/unit_array_algs.cpp) */
void om_kat_form_uses(int nhigh, int high_dim, int low_dim, int const* hv2v, int* out);
void om_kat_reflect_down(int high_dim, int low_dim, int nhigh, int const* hv2v, int nlow,
    int const* lv2v, int nverts, int* ab2b, signed char* codes);
void om_kat_invert_map(int n, int const* hl2l, int nlow, int* a2ab, int* ab2b);
void om_kat_invert_adj(int nhigh, int deg, int nlow, int const* hl2l,
    signed char const* down_codes, int* a2ab, int* ab2b, signed char* codes);
void om_kat_mident_metrics(om_mesh* m, int n, int const* edges, double* out);
void om_kat_swap3d_qualities(om_mesh* m, double max_length, int n, int const* edges,
    double* quals, signed char* configs);
int om_kat_align_vert_index(int n, int index, int code);
void om_kat_mark_down(om_mesh* m, int high, int low, signed char const* marks,
    signed char* out);
void om_kat_mark_up(om_mesh* m, int low, int high, signed char const* marks,
    signed char* out);
void om_kat_swap2d_topology(om_mesh* m, int nkeys, int const* keys2edges, int* edge_verts,
    int* tri_verts);

#ifdef __cplusplus
}
#endif

#endif
