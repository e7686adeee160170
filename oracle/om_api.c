/* Oracle (test infrastructure only): accessors behind the ctypes API. */
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

int om_mesh_dim(om_mesh* m) { return m->dim; }
int om_mesh_nents(om_mesh* m, int d) { return m->nents[d]; }

void om_get_verts(om_mesh* m, int d, int* out) {
  int n = m->nents[d];
  if (d == 0) {
    for (int i = 0; i < n; ++i) out[i] = i;
    return;
  }
  memcpy(out, om_verts_of(m, d), sizeof(int) * (size_t)n * (d + 1));
}

void om_get_down(om_mesh* m, int d, int* out, signed char* codes) {
  int n = m->nents[d] * (d + 1);
  memcpy(out, m->down[d], sizeof(int) * (size_t)n);
  if (codes) {
    if (m->down_codes[d])
      memcpy(codes, m->down_codes[d], (size_t)n);
    else
      memset(codes, 0, (size_t)n);
  }
}

void om_get_coords(om_mesh* m, double* out) {
  memcpy(out, m->coords, sizeof(double) * (size_t)m->nents[0] * m->dim);
}

void om_set_metric(om_mesh* m, int ncomps, double const* metric) {
  free(m->metric);
  m->metric_ncomps = ncomps;
  size_t n = (size_t)m->nents[0] * ncomps;
  m->metric = (double*)om_malloc(sizeof(double) * n);
  memcpy(m->metric, metric, sizeof(double) * n);
  /* setting the metric invalidates cached lengths and qualities,
     src/Omega_h_mesh.cpp:167-177 */
  free(m->length);
  free(m->quality);
  m->length = NULL;
  m->quality = NULL;
}

int om_metric_ncomps(om_mesh* m) { return m->metric_ncomps; }

void om_get_metric(om_mesh* m, double* out) {
  memcpy(out, m->metric, sizeof(double) * (size_t)m->nents[0] * m->metric_ncomps);
}

void om_get_class(om_mesh* m, int d, int* ids, signed char* dims) {
  memcpy(ids, m->class_id[d], sizeof(int) * (size_t)m->nents[d]);
  memcpy(dims, m->class_dim[d], (size_t)m->nents[d]);
}

void om_get_lengths(om_mesh* m, double* out) {
  memcpy(out, om_ask_lengths(m), sizeof(double) * (size_t)m->nents[1]);
}

void om_get_qualities(om_mesh* m, double* out) {
  memcpy(out, om_ask_qualities(m), sizeof(double) * (size_t)m->nents[m->dim]);
}

int om_up_size(om_mesh* m, int low, int high) {
  om_graph const* g = om_ask_up(m, low, high);
  return g->a2ab[g->n];
}

void om_get_up(om_mesh* m, int low, int high, int* a2ab, int* ab2b,
    signed char* codes) {
  om_graph const* g = om_ask_up(m, low, high);
  memcpy(a2ab, g->a2ab, sizeof(int) * (size_t)(g->n + 1));
  memcpy(ab2b, g->ab2b, sizeof(int) * (size_t)g->a2ab[g->n]);
  memcpy(codes, g->codes, (size_t)g->a2ab[g->n]);
}

/* Mesh::ask_dual, src/Omega_h_mesh.cpp:260 (elements_across_sides,
 * src/Omega_h_adj.cpp:591-657) */
int om_dual_size(om_mesh* m) {
  om_graph const* g = om_ask_dual(m);
  return g->a2ab[g->n];
}

void om_get_dual(om_mesh* m, int* a2ab, int* ab2b) {
  om_graph const* g = om_ask_dual(m);
  memcpy(a2ab, g->a2ab, sizeof(int) * (size_t)(g->n + 1));
  memcpy(ab2b, g->ab2b, sizeof(int) * (size_t)g->a2ab[g->n]);
}

/* Mesh::ask_sizes, src/Omega_h_mesh.cpp:390-396 (measure_elements_real,
 * src/Omega_h_shape.cpp:55-89) */
void om_get_sizes(om_mesh* m, double* out) {
  int dim = m->dim;
  int const* ev2v = om_verts_of(m, dim);
  for (int e = 0; e < m->nents[dim]; ++e) {
    double const* p[4];
    for (int k = 0; k < dim + 1; ++k) p[k] = m->coords + (size_t)ev2v[e * (dim + 1) + k] * dim;
    out[e] = om_element_size(dim, p);
  }
}

int om_star_size(om_mesh* m, int d) {
  om_graph const* g = om_ask_star(m, d);
  return g->a2ab[g->n];
}

void om_get_star(om_mesh* m, int d, int* a2ab, int* ab2b) {
  om_graph const* g = om_ask_star(m, d);
  memcpy(a2ab, g->a2ab, sizeof(int) * (size_t)(g->n + 1));
  memcpy(ab2b, g->ab2b, sizeof(int) * (size_t)g->a2ab[g->n]);
}

/* limit_metric_gradation(mesh, metric, max_rate, tol), src/Omega_h_metric.cpp:200-220,
 * applied to the mesh's own metric; returns the number of sweeps */
int om_limit_metric_gradation(om_mesh* m, double max_rate, double tol) {
  int nc = m->metric_ncomps;
  int mdim = (nc == 1) ? 1 : m->dim;
  int nv = m->nents[0];
  size_t n = (size_t)nv * nc;
  om_graph const* g = om_ask_star(m, 0);
  double* values = (double*)om_malloc(sizeof(double) * n);
  double* values2 = (double*)om_malloc(sizeof(double) * n);
  memcpy(values2, m->metric, sizeof(double) * n);
  int steps = 0;
  do {
    memcpy(values, values2, sizeof(double) * n);
    om_limit_gradation_once(m->dim, mdim, nv, g->a2ab, g->ab2b, m->coords, values, max_rate,
        values2);
    ++steps;
  } while (!om_are_close((int)n, values, values2, tol, 1e-10) && steps < 10000);
  om_set_metric(m, nc, values2);
  free(values);
  free(values2);
  return steps;
}

/* vertex tags carried through adapt() (add_tag(VERT, name, ncomps, data)
 * with a LINEAR_INTERP or METRIC transfer type) */
static om_vtag* find_vtag(om_mesh* m, char const* name) {
  for (int t = 0; t < m->nvtags; ++t)
    if (strncmp(m->vtags[t].name, name, sizeof(m->vtags[t].name)) == 0) return &m->vtags[t];
  return NULL;
}

int om_add_vtag(om_mesh* m, char const* name, int ncomps, int kind, double const* data) {
  om_vtag* t = find_vtag(m, name);
  if (!t) {
    if (m->nvtags >= OM_MAX_VTAGS) return -1;
    t = &m->vtags[m->nvtags++];
    memset(t, 0, sizeof(*t));
    strncpy(t->name, name, sizeof(t->name) - 1);
  }
  free(t->data);
  t->ncomps = ncomps;
  t->kind = kind;
  size_t n = (size_t)m->nents[0] * ncomps;
  t->data = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(t->data, data, sizeof(double) * n);
  return 0;
}

int om_vtag_ncomps(om_mesh* m, char const* name) {
  om_vtag* t = find_vtag(m, name);
  return t ? t->ncomps : -1;
}

void om_get_vtag(om_mesh* m, char const* name, double* out) {
  om_vtag* t = find_vtag(m, name);
  if (t) memcpy(out, t->data, sizeof(double) * (size_t)m->nents[0] * t->ncomps);
}

void om_remove_vtag(om_mesh* m, char const* name) {
  om_vtag* t = find_vtag(m, name);
  if (!t) return;
  free(t->data);
  int i = (int)(t - m->vtags);
  for (int j = i; j + 1 < m->nvtags; ++j) m->vtags[j] = m->vtags[j + 1];
  --m->nvtags;
}

/* interpolate_between_metrics, src/Omega_h_metric.cpp:106-111: linearize
 * (log), a (1 - t) + b t per stored component
 * (src/Omega_h_array_ops.cpp:567-574), delinearize (exp) */
void om_interpolate_metrics(int mdim, int n, double const* a, double const* b, double t,
    double* out) {
  int nc = om_symm_ncomps(mdim);
  for (int i = 0; i < n; ++i) {
    om_mat ma, mb, la, lb, r;
    double va[6], vb[6], vc[6];
    om_get_symm(mdim, a + (size_t)i * nc, &ma);
    om_get_symm(mdim, b + (size_t)i * nc, &mb);
    om_log_spd(mdim, &ma, &la);
    om_log_spd(mdim, &mb, &lb);
    om_set_symm(mdim, &la, va);
    om_set_symm(mdim, &lb, vb);
    for (int k = 0; k < nc; ++k) vc[k] = va[k] * (1.0 - t) + vb[k] * t;
    om_mat lc;
    om_get_symm(mdim, vc, &lc);
    om_exp_spd(mdim, &lc, &r);
    om_set_symm(mdim, &r, out + (size_t)i * nc);
  }
}

double om_min_quality(om_mesh* m) {
  int ne = m->nents[m->dim];
  double* q = (double*)om_malloc(sizeof(double) * (size_t)(ne + 1));
  om_get_qualities(m, q);
  double r = 1.79769313486231570815e+308;
  for (int i = 0; i < ne; ++i) r = (q[i] < r) ? q[i] : r;
  free(q);
  return r;
}

double om_max_length(om_mesh* m) {
  int ne = m->nents[1];
  double* l = (double*)om_malloc(sizeof(double) * (size_t)(ne + 1));
  om_get_lengths(m, l);
  double r = -1.79769313486231570815e+308;
  for (int i = 0; i < ne; ++i) r = (r < l[i]) ? l[i] : r;
  free(l);
  return r;
}

static int okay(om_mesh* m, om_adapt_opts const* o) {
  return om_min_quality(m) >= o->min_quality_allowed && !(om_max_length(m) > o->max_length_allowed);
}

/* approach_metric, src/Omega_h_approach.cpp:83-126: move "metric" toward the
 * "target_metric" vertex tag by halving steps of log-Euclidean
 * interpolation until the mesh is acceptable.  Returns 1 if it moved, 0
 * without a target, -1 if the mesh is unacceptable to begin with and -2 if
 * the step fell below min_step (the reference fails in both cases). */
int om_approach_metric(om_mesh* m, om_adapt_opts const* o, double min_step) {
  om_vtag* tg = find_vtag(m, "target_metric");
  if (!tg) return 0;
  if (om_min_quality(m) < o->min_quality_allowed) return -1;
  if (om_max_length(m) > o->max_length_allowed) return -1;
  int nc = m->metric_ncomps;
  if (tg->ncomps != nc) return -1;
  int nv = m->nents[0];
  size_t n = (size_t)nv * nc;
  double* orig = (double*)om_malloc(sizeof(double) * (n + 1));
  double* target = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(orig, m->metric, sizeof(double) * n);
  memcpy(target, tg->data, sizeof(double) * n);
  om_set_metric(m, nc, target);
  int r = 1;
  if (okay(m, o)) {
    om_remove_vtag(m, "target_metric");
  } else {
    double factor = 1.0;
    double* cur = (double*)om_malloc(sizeof(double) * (n + 1));
    int mdim = (nc == 1) ? 1 : m->dim;
    do {
      factor /= 2.0;
      if (factor < min_step) {
        r = -2;
        break;
      }
      om_interpolate_metrics(mdim, nv, orig, target, factor, cur);
      om_set_metric(m, nc, cur);
    } while (!okay(m, o));
    free(cur);
  }
  free(orig);
  free(target);
  return r;
}

/* Mesh::set_coords, src/Omega_h_mesh.cpp:365-368: new coordinates drop the
 * cached lengths and qualities (src/Omega_h_mesh.cpp:167-177) */
void om_set_coords(om_mesh* m, double const* coords) {
  size_t n = (size_t)m->nents[0] * m->dim;
  if (!m->coords) m->coords = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(m->coords, coords, sizeof(double) * n);
  free(m->length);
  free(m->quality);
  m->length = NULL;
  m->quality = NULL;
}

/* warp_to_limit, src/Omega_h_approach.cpp:36-81: move the vertices by the
 * "warp" vertex tag, halving the move until the mesh's quality and lengths
 * are acceptable; what is left of the warp stays in the tag.  Returns 1 if
 * it moved, 0 without a warp tag, -1 if the mesh is unacceptable to begin
 * with (or the tag is not one vector per vertex) and -2 when more than
 * max_niters halvings were needed and exit_on_stall is 0 (the reference
 * fails in both cases).  On a stall with exit_on_stall the tag is dropped
 * and the last tried coordinates stay, as in the reference. */
int om_warp_to_limit(om_mesh* m, om_adapt_opts const* o, int exit_on_stall, int max_niters) {
  om_vtag* tg = find_vtag(m, "warp");
  if (!tg) return 0;
  if (tg->ncomps != m->dim) return -1;
  if (!okay(m, o)) return -1;
  size_t n = (size_t)m->nents[0] * m->dim;
  double* coords = (double*)om_malloc(sizeof(double) * (n + 1));
  double* warp = (double*)om_malloc(sizeof(double) * (n + 1));
  double* cur = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(coords, m->coords, sizeof(double) * n);
  memcpy(warp, tg->data, sizeof(double) * n);
  for (size_t k = 0; k < n; ++k) cur[k] = coords[k] + warp[k];
  om_set_coords(m, cur);
  int r = 1;
  if (okay(m, o)) {
    om_remove_vtag(m, "warp");
  } else {
    double* rem = (double*)om_malloc(sizeof(double) * (n + 1));
    for (size_t k = 0; k < n; ++k) rem[k] = 0.0;
    int i = 0;
    int stalled = 0;
    do {
      ++i;
      if (i > max_niters) {
        stalled = 1;
        break;
      }
      for (size_t k = 0; k < n; ++k) {
        double h = warp[k] / 2.0;
        warp[k] = h;
        rem[k] = rem[k] + h;
        cur[k] = coords[k] + h;
      }
      om_set_coords(m, cur);
    } while (!okay(m, o));
    if (!stalled)
      memcpy(tg->data, rem, sizeof(double) * n);
    else if (exit_on_stall)
      om_remove_vtag(m, "warp");
    else
      r = -2;
    free(rem);
  }
  free(coords);
  free(warp);
  free(cur);
  return r;
}

/* element tags transferred pointwise (OMEGA_H_POINTWISE) */
static om_etag* find_etag(om_mesh* m, char const* name) {
  for (int t = 0; t < m->netags; ++t)
    if (strncmp(m->etags[t].name, name, sizeof(m->etags[t].name)) == 0) return &m->etags[t];
  return NULL;
}

int om_add_etag(om_mesh* m, char const* name, int ncomps, double const* data) {
  om_etag* t = find_etag(m, name);
  if (!t) {
    if (m->netags >= OM_MAX_ETAGS || strlen(name) >= sizeof(t->name)) return -1;
    t = &m->etags[m->netags++];
    memset(t, 0, sizeof(*t));
    strcpy(t->name, name);
  }
  free(t->data);
  size_t n = (size_t)m->nents[m->dim] * ncomps;
  t->ncomps = ncomps;
  t->data = (double*)om_malloc(sizeof(double) * (n + 1));
  memcpy(t->data, data, sizeof(double) * n);
  return 0;
}

/* the transfer of an element tag: OM_POINTWISE, OM_CONSERVE or OM_DENSITY
 * (TransferOpts::type_map); for OM_CONSERVE the diffusion of its integral's
 * error (integral_diffuse_map[integral_map[name]]) */
int om_set_etag_transfer(om_mesh* m, char const* name, int kind, int diffuse_type, double tol,
    double floor) {
  om_etag* t = find_etag(m, name);
  if (!t || kind < OM_POINTWISE || kind > OM_DENSITY) return -1;
  t->kind = kind;
  t->diffuse.type = diffuse_type;
  t->diffuse.tol = tol;
  t->diffuse.floor = floor;
  return 0;
}

/* a vertex tag as a momentum velocity (OMEGA_H_MOMENTUM_VELOCITY) with its
 * conserved element density (velocity_density_map) and the diffusion of
 * its momentum's error (integral_diffuse_map[velocity_momentum_map[name]]) */
int om_set_momentum_velocity(om_mesh* m, char const* name, char const* density,
    int diffuse_type, double tol, double floor) {
  om_vtag* t = find_vtag(m, name);
  om_etag* d = find_etag(m, density);
  if (!t || !d || d->ncomps != 1 || d->kind != OM_CONSERVE || t->ncomps != m->dim) return -1;
  t->kind = OM_MOMENTUM;
  memset(t->density, 0, sizeof(t->density));
  strncpy(t->density, density, sizeof(t->density) - 1);
  t->diffuse.type = diffuse_type;
  t->diffuse.tol = tol;
  t->diffuse.floor = floor;
  return 0;
}

int om_etag_ncomps(om_mesh* m, char const* name) {
  om_etag* t = find_etag(m, name);
  return t ? t->ncomps : -1;
}

void om_get_etag(om_mesh* m, char const* name, double* out) {
  om_etag* t = find_etag(m, name);
  if (t) memcpy(out, t->data, sizeof(double) * (size_t)m->nents[m->dim] * t->ncomps);
}

/* inheritance tags (OMEGA_H_INHERIT): one array per dimension 0..dim */
static om_itag* find_itag(om_mesh* m, char const* name) {
  for (int t = 0; t < m->nitags; ++t)
    if (strncmp(m->itags[t].name, name, sizeof(m->itags[t].name)) == 0) return &m->itags[t];
  return NULL;
}

int om_add_itag(om_mesh* m, char const* name, int ncomps, double const* const* per_dim) {
  om_itag* t = find_itag(m, name);
  if (!t) {
    if (m->nitags >= OM_MAX_ITAGS) return -1;
    t = &m->itags[m->nitags++];
    memset(t, 0, sizeof(*t));
    strncpy(t->name, name, sizeof(t->name) - 1);
  }
  t->ncomps = ncomps;
  for (int d = 0; d <= m->dim; ++d) {
    free(t->data[d]);
    size_t n = (size_t)m->nents[d] * ncomps;
    t->data[d] = (double*)om_malloc(sizeof(double) * (n + 1));
    memcpy(t->data[d], per_dim[d], sizeof(double) * n);
  }
  return 0;
}

int om_itag_ncomps(om_mesh* m, char const* name) {
  om_itag* t = find_itag(m, name);
  return t ? t->ncomps : -1;
}

void om_get_itag(om_mesh* m, char const* name, int d, double* out) {
  om_itag* t = find_itag(m, name);
  if (t) memcpy(out, t->data[d], sizeof(double) * (size_t)m->nents[d] * t->ncomps);
}

/* one disjoint mesh of nparts meshes (same dimension and metric layout,
 * classified, no vertex tags): part k's entities of every dimension follow
 * those of parts < k, and the vertex tag "part" holds each vertex's part */
om_mesh* om_mesh_join(int nparts, om_mesh* const* parts) {
  int dim = parts[0]->dim;
  int nc = parts[0]->metric_ncomps;
  om_mesh* r = om_mesh_new(dim);
  r->metric_ncomps = nc;
  for (int d = 0; d <= dim; ++d) {
    r->nents[d] = 0;
    for (int k = 0; k < nparts; ++k) r->nents[d] += parts[k]->nents[d];
  }
  for (int d = 1; d <= dim; ++d) {
    r->down[d] = (int*)om_malloc(sizeof(int) * ((size_t)r->nents[d] * (d + 1) + 1));
    if (d >= 2) r->down_codes[d] = (signed char*)om_malloc((size_t)r->nents[d] * (d + 1) + 1);
  }
  r->coords = (double*)om_malloc(sizeof(double) * ((size_t)r->nents[0] * dim + 1));
  r->metric = (double*)om_malloc(sizeof(double) * ((size_t)r->nents[0] * nc + 1));
  for (int d = 0; d <= dim; ++d) {
    r->class_id[d] = (int*)om_malloc(sizeof(int) * ((size_t)r->nents[d] + 1));
    r->class_dim[d] = (signed char*)om_malloc((size_t)r->nents[d] + 1);
  }
  double* part = (double*)om_malloc(sizeof(double) * ((size_t)r->nents[0] + 1));
  int at[4] = {0, 0, 0, 0};
  for (int k = 0; k < nparts; ++k) {
    om_mesh const* p = parts[k];
    for (int d = 1; d <= dim; ++d) {
      size_t n = (size_t)p->nents[d] * (d + 1);
      for (size_t i = 0; i < n; ++i)
        r->down[d][(size_t)at[d] * (d + 1) + i] = p->down[d][i] + at[d - 1];
      if (d >= 2) memcpy(r->down_codes[d] + (size_t)at[d] * (d + 1), p->down_codes[d], n);
    }
    memcpy(r->coords + (size_t)at[0] * dim, p->coords, sizeof(double) * (size_t)p->nents[0] * dim);
    memcpy(r->metric + (size_t)at[0] * nc, p->metric, sizeof(double) * (size_t)p->nents[0] * nc);
    for (int d = 0; d <= dim; ++d) {
      memcpy(r->class_id[d] + at[d], p->class_id[d], sizeof(int) * (size_t)p->nents[d]);
      memcpy(r->class_dim[d] + at[d], p->class_dim[d], (size_t)p->nents[d]);
    }
    for (int v = 0; v < p->nents[0]; ++v) part[at[0] + v] = (double)k;
    for (int d = 0; d <= dim; ++d) at[d] += p->nents[d];
  }
  om_add_vtag(r, "part", 1, OM_LINEAR, part);
  free(part);
  return r;
}

/* a mesh from its arrays: nents[0..dim], the down adjacencies (and alignment
 * codes for dim >= 2) of every dimension 1..dim, coordinates, metric and the
 * classification of every dimension (copied) */
om_mesh* om_mesh_from_arrays(int dim, int const* nents, int const* const* down,
    signed char const* const* codes, double const* coords, int metric_ncomps,
    double const* metric, int const* const* class_id, signed char const* const* class_dim) {
  om_mesh* m = om_mesh_new(dim);
  for (int d = 0; d <= dim; ++d) m->nents[d] = nents[d];
  for (int d = 1; d <= dim; ++d) {
    size_t n = (size_t)nents[d] * (d + 1);
    m->down[d] = (int*)om_malloc(sizeof(int) * (n + 1));
    memcpy(m->down[d], down[d], sizeof(int) * n);
    if (d >= 2) {
      m->down_codes[d] = (signed char*)om_malloc(n + 1);
      memcpy(m->down_codes[d], codes[d], n);
    }
  }
  m->coords = (double*)om_malloc(sizeof(double) * ((size_t)nents[0] * dim + 1));
  memcpy(m->coords, coords, sizeof(double) * (size_t)nents[0] * dim);
  m->metric_ncomps = metric_ncomps;
  m->metric = (double*)om_malloc(sizeof(double) * ((size_t)nents[0] * metric_ncomps + 1));
  memcpy(m->metric, metric, sizeof(double) * (size_t)nents[0] * metric_ncomps);
  for (int d = 0; d <= dim; ++d) {
    m->class_id[d] = (int*)om_malloc(sizeof(int) * ((size_t)nents[d] + 1));
    memcpy(m->class_id[d], class_id[d], sizeof(int) * (size_t)nents[d]);
    m->class_dim[d] = (signed char*)om_malloc((size_t)nents[d] + 1);
    memcpy(m->class_dim[d], class_dim[d], (size_t)nents[d]);
  }
  return m;
}
