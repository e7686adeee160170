/* Oracle (test infrastructure only): internal declarations of the CPU
 * restatement of Omega_h's adapt() hot path.  Only tests/, the smoke check
 * and bench.py's CPU-baseline leg load this code; the product never does. */
#ifndef OM_INTERNAL_H
#define OM_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* scalar math (om_math.c) */
double om_log(double x);
double om_exp(double x);
double om_cbrt(double x);
double om_acos(double x);
double om_cos(double x);
double om_sqrt(double x);

/* small dense algebra, column-major like src/Omega_h_matrix.hpp:14
 * (m[col][row]) */
typedef struct {
  double m[3][3];
} om_mat;

double om_det(int dim, om_mat const* m);
void om_decompose_eigen(int dim, om_mat const* m, om_mat* q, double* l);
double om_max_eigenvalue(int dim, om_mat const* m);
void om_element_implied_metrics(om_mesh* m, double scalar, double* e2m);
void om_project_metrics(om_mesh* m, int mdim, double const* e2m, double* out);
void om_element_mident_metrics(om_mesh* m, int mdim, double const* v2m, double* out);
void om_compose_ortho(int dim, om_mat const* q, double const* l, om_mat* out);
void om_log_spd(int dim, om_mat const* m, om_mat* out);
void om_exp_spd(int dim, om_mat const* m, om_mat* out);
void om_get_symm(int mdim, double const* v, om_mat* m);
void om_set_symm(int mdim, om_mat const* m, double* v);
int om_symm_ncomps(int mdim);
int om_roots3(double const coeffs[3], double eps, double roots[3], int mults[3]);
int om_roots2(double const coeffs[2], double eps, double roots[2], int mults[2]);

/* geometry */
double om_metric_edge_length(int sdim, int mdim, double const* p0,
    double const* p1, double const* m0, double const* m1);
double om_element_size(int dim, double const* const* p);
double om_element_quality(int dim, int mdim, double const* const* p,
    double const* const* ms);
void om_average_metric(int mdim, int n, double const* const* ms, double* out);
void om_barycentric(int dim, double const* x, double const* const* p, double* xi);
void om_intersect_metrics(int dim, double const* m1, double const* m2, double* out);
void om_limit_gradation_once(int sdim, int mdim, int nverts, int const* a2ab, int const* ab2b,
    double const* coords, double const* in, double max_rate, double* out);
int om_are_close(int n, double const* a, double const* b, double tol, double floor);
double om_min_quality(om_mesh* m);
void om_side_vector(int dim, double const* const* p, double* out);
int om_most_normal(int dim, int n, double const* N, double* c);
double om_max_length(om_mesh* m);

/* growable int arrays */
typedef struct {
  int* a2ab;        /* n+1 offsets */
  int* ab2b;        /* entries */
  signed char* codes; /* may be NULL */
  int n;
} om_graph;

void om_graph_free(om_graph* g);

/* simplex templates, src/Omega_h_simplex.hpp */
int om_down_template(int elem_dim, int bdry_dim, int which_bdry, int which_vert);
int om_opposite_template(int elem_dim, int bdry_dim, int which_bdry);
int om_simplex_degree(int from_dim, int to_dim);

/* codes, src/Omega_h_align.hpp:43-54 */
static inline int om_code_is_flipped(signed char c) { return c & 1; }
static inline int om_code_rotation(signed char c) { return (c >> 1) & 3; }
static inline int om_code_which_down(signed char c) { return c >> 3; }
static inline signed char om_make_code(int flipped, int rot, int which_down) {
  return (signed char)((which_down << 3) | (rot << 1) | (flipped ? 1 : 0));
}

/* a vertex tag carried through adapt(): OM_LINEAR values are averaged at
 * refinement midpoints (average_field, src/Omega_h_mesh.cpp:822-844), OM_METRIC
 * values are log-Euclidean midpoint metrics (get_mident_metrics,
 * src/Omega_h_metric.cpp:56-104); both are copied for surviving vertices
 * (src/Omega_h_transfer.cpp:391-428,585-619,701-730) */
#define OM_MAX_VTAGS 8
enum { OM_LINEAR = 0, OM_METRIC = 1, OM_MOMENTUM = 2 };

/* how an integral's error is diffused before it is corrected
 * (VarCompareOpts src/Omega_h_compare.hpp:14-21, as used by
 * diffuse_integrals_weighted src/Omega_h_conserve.cpp:708-734) */
enum { OM_DIFFUSE_NONE = 0, OM_DIFFUSE_RELATIVE = 1, OM_DIFFUSE_ABSOLUTE = 2 };
typedef struct {
  int type;
  double tol;
  double floor;
} om_diffuse;

/* OM_MOMENTUM: a velocity interpolated like OM_LINEAR whose momentum (with
 * the conserved element density named `density`) is conserved
 * (OMEGA_H_MOMENTUM_VELOCITY); err holds the elements' momentum error
 * during adapt() */
typedef struct {
  char name[64];
  int ncomps;
  int kind;
  double* data;
  char density[64];
  om_diffuse diffuse;
  double* err;
} om_vtag;

/* a tag on every dimension transferred by inheritance (OMEGA_H_INHERIT,
 * should_inherit src/Omega_h_transfer.cpp:20-34): surviving entities keep
 * their values, a product takes the value of the entity it came from, as
 * class_id does (src/Omega_h_transfer.cpp:213-267,431-470,656-700) */
#define OM_MAX_ITAGS 8
typedef struct {
  char name[64];
  int ncomps;
  double* data[4];
} om_itag;

/* an element tag transferred pointwise (OMEGA_H_POINTWISE, should_fit
 * src/Omega_h_transfer.cpp:47-54): refinement products take their parent
 * element's value (src/Omega_h_transfer.cpp:324-335); coarsening and swap
 * products take the value of the old element of the cavity nearest to
 * their centroid (transfer_pointwise_tmpl, src/Omega_h_transfer.cpp:515-583) */
#define OM_MAX_ETAGS 8
/* element tag kinds: OM_POINTWISE (OMEGA_H_POINTWISE), OM_CONSERVE (a
 * density whose integral is conserved, OMEGA_H_CONSERVE), OM_DENSITY (a
 * density transferred by intersection, OMEGA_H_DENSITY); err holds the
 * integral error of a conserved density during adapt(), old its values
 * before the final correction */
enum { OM_POINTWISE = 0, OM_CONSERVE = 1, OM_DENSITY = 2 };
typedef struct {
  char name[64];
  int ncomps;
  double* data;
  int kind;
  om_diffuse diffuse;
  double* err;
  double* old;
} om_etag;

struct om_mesh {
  int dim;
  int nvtags;
  om_vtag vtags[OM_MAX_VTAGS];
  int nitags;
  om_itag itags[OM_MAX_ITAGS];
  int netags;
  om_etag etags[OM_MAX_ETAGS];
  int nents[4];
  /* down[d] for d >= 1: nents[d] * (d+1) lower entities of dim d-1 */
  int* down[4];
  signed char* down_codes[4]; /* NULL for d == 1 */
  double* coords;
  double* metric;
  int metric_ncomps;
  int* class_id[4];
  signed char* class_dim[4];
  double* length;  /* cached, may be NULL */
  double* quality; /* cached, may be NULL */
  /* lazily derived caches */
  int* verts_of[4];       /* entity -> vertices, derived by transit */
  om_graph* up[4][4];     /* up[low][high] */
  om_graph* star[4];      /* star graphs of verts/edges */
  om_graph* dual;
  int* tet_edges_cache;              /* tet -> edge by transit */
  signed char* tet_edge_codes_cache;
};

/* mesh helpers (om_mesh.c) */
om_mesh* om_mesh_new(int dim);
void om_mesh_clear_caches(om_mesh* m);
int const* om_verts_of(om_mesh* m, int d);
om_graph const* om_ask_up(om_mesh* m, int low, int high);
int const* om_ask_down(om_mesh* m, int high, int low, signed char const** codes);
om_graph const* om_ask_star(om_mesh* m, int d);
om_graph const* om_ask_dual(om_mesh* m);
void om_reflect_down(om_mesh* m_new, int high_dim, int nprods,
    int const* prod_verts, int* out_down, signed char* out_codes);
double const* om_ask_lengths(om_mesh* m);
double const* om_ask_qualities(om_mesh* m);
void om_measure_edges(om_mesh* m, int n, int const* edges, double* out);
void om_measure_elems(om_mesh* m, int n, int const* elems, double* out);
int* om_find_unique(int nhigh, int high_dim, int low_dim, int const* hv2v,
    int* nlow_out);

/* indset (om_adapt.c) */
void om_find_indset(om_mesh* m, int d, double const* quality,
    signed char const* cands, signed char* out);

/* utilities */
void* om_malloc(size_t n);
void* om_calloc(size_t n, size_t sz);
int* om_offset_scan(int n, int const* counts); /* n+1 */
void om_stable_sort_keys(int n, int width, int64_t const* keys, int* perm);


/* intersection of two simplices (om_r3d.c) */
double om_intersect_size(int dim, double const* a, double const* b, int* bad);

/* conservative and momentum transfer (om_conserve.c) */
int om_conserve_any(om_mesh* m);
void om_conserve_setup(om_mesh* m);
void om_conserve_pass(om_mesh* old, om_mesh* nw, int op, int key_dim, int nkeys,
    int const* keys2kds, int const* k2p, int const* prods2new, int nsame,
    int const* same2old, int const* same2new);
int om_conserve_correct(om_mesh* m, int max_iters);
/* diffusion sweeps after which correct_integral_errors gives up (the
 * reference's loop has no bound) */
#define OM_MAX_DIFFUSION_SWEEPS 100000

/* swap3d tables (om_swap3d_tables.c) */
#define OM_MAX_EDGE_SWAP 7
#define OM_MAX_UNIQUE_TRIS 35
#define OM_MAX_UNIQUE_EDGES 26
extern int const om_swap_mesh_sizes[8];
extern int const om_swap_mesh_counts[8];
extern int const om_swap_nedges[8];
int const* om_swap_triangle(int loop_size, int uniq_tri);   /* 3 loop verts */
int om_swap_mesh_tri(int loop_size, int mesh, int mesh_tri); /* uniq tri */
int const* om_swap_unique_edge(int loop_size, int uniq_edge); /* 2 loop verts */
int om_swap_mesh_edge(int loop_size, int mesh, int mesh_edge); /* uniq edge */

#ifdef __cplusplus
}
#endif

#endif
