/* Oracle (test infrastructure only): metrics implied by a mesh's elements.
 *
 *   element_implied_metric (2D, 3D)   src/Omega_h_shape.hpp:240-273
 *   solve_using_qr (Householder)      src/Omega_h_qr.hpp:18-149
 *   get_element_implied_length_metrics src/Omega_h_metric.cpp:257-276
 *   get_implied_metrics (size scaling) src/Omega_h_metric.cpp:280-309
 *   project_metrics                   src/Omega_h_metric.cpp:221-245,
 *     average_metric_contrib/_finish  src/Omega_h_metric.hpp:125-139
 * Column-major matrices as in src/Omega_h_matrix.hpp (a[col][row]). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

#define OM_EPSILON 1e-10

static double sign_of(double x) { return (x < 0.0) ? -1.0 : 1.0; }

/* least squares / square solve of a (m x n, n <= m <= 6) by Householder QR */
static void solve_using_qr(int m, int n, double a[6][6], double const* b_in, double* x) {
  double v[6][6];
  for (int k = 0; k < n; ++k) {
    double norm_x = 0.0;
    for (int i = k; i < m; ++i) norm_x += a[k][i] * a[k][i];
    norm_x = om_sqrt(norm_x);
    for (int i = k; i < m; ++i) v[k][i] = a[k][i];
    v[k][k] += sign_of(a[k][k]) * norm_x;
    double norm_v = 0.0;
    for (int i = k; i < m; ++i) norm_v += v[k][i] * v[k][i];
    norm_v = om_sqrt(norm_v);
    for (int i = k; i < m; ++i) v[k][i] /= norm_v;
    for (int j = k; j < n; ++j) {
      double dot = 0.0;
      for (int i = k; i < m; ++i) dot += a[j][i] * v[k][i];
      for (int i = k; i < m; ++i) a[j][i] -= 2.0 * dot * v[k][i];
    }
  }
  double b[6];
  for (int i = 0; i < m; ++i) b[i] = b_in[i];
  for (int k = 0; k < n; ++k) {
    double dot = 0.0;
    for (int i = k; i < m; ++i) dot += v[k][i] * b[i];
    for (int i = k; i < m; ++i) b[i] -= 2.0 * dot * v[k][i];
  }
  for (int ii = 0; ii < n; ++ii) {
    int i = n - ii - 1;
    x[i] = b[i];
    for (int j = i + 1; j < n; ++j) x[i] -= a[j][i] * x[j];
    x[i] /= a[i][i];
  }
}

static void cross3(double const* a, double const* b, double* c) {
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  c[0] = t0;
  c[1] = t1;
  c[2] = t2;
}

/* element_implied_metric: the metric in which every edge of the element has
 * unit length, in symmetric storage (xx, yy, [zz,] xy[, yz, xz]) */
void om_element_implied_metric(int dim, double const* const* p, double* out) {
  double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, ev[6][3];
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) b[i][j] = p[i + 1][j] - p[0][j];
  if (dim == 3) {
    for (int j = 0; j < 3; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
      ev[3][j] = b[2][j];
      ev[4][j] = p[3][j] - p[1][j];
      ev[5][j] = p[3][j] - p[2][j];
    }
    double a[6][6], rhs[6];
    for (int i = 0; i < 6; ++i) {
      a[0][i] = ev[i][0] * ev[i][0];
      a[1][i] = ev[i][1] * ev[i][1];
      a[2][i] = ev[i][2] * ev[i][2];
      a[3][i] = 2 * ev[i][0] * ev[i][1];
      a[4][i] = 2 * ev[i][1] * ev[i][2];
      a[5][i] = 2 * ev[i][0] * ev[i][2];
      rhs[i] = 1.0;
    }
    solve_using_qr(6, 6, a, rhs, out);
  } else {
    for (int j = 0; j < 2; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
    }
    /* x = invert(a) * rhs with the cofactor inverse of src/Omega_h_matrix.hpp:344-350 */
    double a[3][3];
    for (int i = 0; i < 3; ++i) {
      a[0][i] = ev[i][0] * ev[i][0];
      a[1][i] = ev[i][1] * ev[i][1];
      a[2][i] = 2 * ev[i][0] * ev[i][1];
    }
    om_mat am, inv;
    memcpy(am.m, a, sizeof(a));
    double cb[3][3];
    cross3(am.m[1], am.m[2], cb[0]);
    cross3(am.m[2], am.m[0], cb[1]);
    cross3(am.m[0], am.m[1], cb[2]);
    double det = om_det(3, &am);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) inv.m[j][i] = cb[i][j] / det;
    double rhs[3] = {1.0, 1.0, 1.0};
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = inv.m[0][i] * rhs[0];
    for (int j = 1; j < 3; ++j)
      for (int i = 0; i < 3; ++i) x[i] = x[i] + inv.m[j][i] * rhs[j];
    out[0] = x[0];
    out[1] = x[1];
    out[2] = x[2];
  }
}

static double typical_over_perfect_metric_scalar(int dim) {
  /* power(typical_unit_simplex_size / equilateral_simplex_size, 2, dim) */
  double typical = dim == 3 ? 0.0838934100219 : 0.377645136635;
  double perfect = dim == 3 ? 0.1178511301977579 : 0.4330127018922193;
  double s = typical / perfect;
  return dim == 3 ? om_cbrt(s * s) : s;
}

static double max_norm(int n, om_mat const* a) {
  double x = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double v = fabs(a->m[j][i]);
      x = (x < v) ? v : x;
    }
  return x;
}

/* element_implied_length_metrics (src/Omega_h_metric.cpp:257-276) times a
 * scalar (1 for length metrics, the size scaling of
 * get_element_implied_size_metrics, src/Omega_h_metric.cpp:295-301) */
void om_element_implied_metrics(om_mesh* m, double scalar, double* e2m) {
  int dim = m->dim;
  int nc = om_symm_ncomps(dim);
  int ne = m->nents[dim];
  int const* ev2v = om_verts_of(m, dim);
  for (int e = 0; e < ne; ++e) {
    double const* p[4];
    for (int k = 0; k <= dim; ++k) p[k] = m->coords + (size_t)ev2v[e * (dim + 1) + k] * dim;
    double v[6];
    om_element_implied_metric(dim, p, v);
    for (int k = 0; k < nc; ++k) e2m[(size_t)e * nc + k] = v[k] * scalar;
  }
}

/* project_metrics (src/Omega_h_metric.cpp:221-252): log-Euclidean average
 * of the metrics (dimension mdim) of each vertex's elements, degenerate
 * metrics skipped (average_metric_contrib/_finish with has_degen,
 * src/Omega_h_metric.hpp:125-139) */
void om_project_metrics(om_mesh* m, int mdim, double const* e2m, double* out) {
  int dim = m->dim;
  int nc = om_symm_ncomps(mdim);
  om_graph const* v2e = om_ask_up(m, 0, dim);
  for (int v = 0; v < m->nents[0]; ++v) {
    om_mat am;
    memset(&am, 0, sizeof(am));
    int n = 0;
    for (int j = v2e->a2ab[v]; j < v2e->a2ab[v + 1]; ++j) {
      om_mat em, lm;
      om_get_symm(mdim, e2m + (size_t)v2e->ab2b[j] * nc, &em);
      if (max_norm(mdim, &em) < OM_EPSILON) continue;
      om_log_spd(mdim, &em, &lm);
      for (int c = 0; c < mdim; ++c)
        for (int r = 0; r < mdim; ++r) am.m[c][r] = am.m[c][r] + lm.m[c][r];
      ++n;
    }
    om_mat r;
    if (n == 0) {
      r = am;
    } else {
      for (int c = 0; c < mdim; ++c)
        for (int rr = 0; rr < mdim; ++rr) am.m[c][rr] = am.m[c][rr] / n;
      om_exp_spd(mdim, &am, &r);
    }
    om_set_symm(mdim, &r, out + (size_t)v * nc);
  }
}

/* get_mident_metrics(mesh, dim, v2m, has_degen = true),
 * src/Omega_h_metric.cpp:56-104: the average of each element's vertex
 * metrics, degenerate ones skipped */
void om_element_mident_metrics(om_mesh* m, int mdim, double const* v2m, double* out) {
  int dim = m->dim;
  int nc = om_symm_ncomps(mdim);
  int const* ev2v = om_verts_of(m, dim);
  for (int e = 0; e < m->nents[dim]; ++e) {
    om_mat am;
    memset(&am, 0, sizeof(am));
    int n = 0;
    for (int k = 0; k <= dim; ++k) {
      om_mat vm, lm;
      om_get_symm(mdim, v2m + (size_t)ev2v[e * (dim + 1) + k] * nc, &vm);
      if (max_norm(mdim, &vm) < OM_EPSILON) continue;
      om_log_spd(mdim, &vm, &lm);
      for (int c = 0; c < mdim; ++c)
        for (int r = 0; r < mdim; ++r) am.m[c][r] = am.m[c][r] + lm.m[c][r];
      ++n;
    }
    om_mat r;
    if (n == 0) {
      r = am;
    } else {
      for (int c = 0; c < mdim; ++c)
        for (int rr = 0; rr < mdim; ++rr) am.m[c][rr] = am.m[c][rr] / n;
      om_exp_spd(mdim, &am, &r);
    }
    om_set_symm(mdim, &r, out + (size_t)e * nc);
  }
}

/* get_implied_metrics: project the size-scaled element implied metrics to
 * the vertices by log-Euclidean averaging over each vertex's elements */
void om_get_implied_metrics(om_mesh* m, double* out) {
  int dim = m->dim;
  int nc = om_symm_ncomps(dim);
  int ne = m->nents[dim];
  double* e2m = (double*)om_malloc(sizeof(double) * (size_t)(ne + 1) * nc);
  om_element_implied_metrics(m, typical_over_perfect_metric_scalar(dim), e2m);
  om_project_metrics(m, dim, e2m, out);
  free(e2m);
}
