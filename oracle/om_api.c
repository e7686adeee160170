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
