/* Oracle (test infrastructure only): metric-space geometry of Omega_h's
 * adapt() path, restated in C with the reference's exact operation order
 * (the reference's Vector/Matrix templates evaluate left to right, and
 * matrices are column-major, m[col][row]: src/Omega_h_matrix.hpp:14).
 *
 *   determinant          src/Omega_h_matrix.hpp:271-293
 *   find_polynomial_roots src/Omega_h_eigen.hpp:21-124
 *   characteristic_polynomial src/Omega_h_eigen.hpp:127-148
 *   single/double_eigenvector, get_1d_row_space src/Omega_h_eigen.hpp:175-226
 *   decompose_eigen      src/Omega_h_eigen.hpp:228-319
 *   compose_ortho        src/Omega_h_eigen.hpp:331-334
 *   log_spd_old/exp_spd_old src/Omega_h_eigen.hpp:473-488
 *   form_ortho_basis     src/Omega_h_matrix.hpp:553-575
 *   metric_product/maxdet_metric/average_metric src/Omega_h_metric.hpp:8-181
 *   metric_element_quality src/Omega_h_quality.hpp:8-37
 *   metric_edge_length   src/Omega_h_shape.hpp:112-125
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

#define OM_EPSILON 1e-10
#define OM_PI 3.14159265358979323

static double max2(double a, double b) { return (a < b) ? b : a; }
static double min2(double a, double b) { return (b < a) ? b : a; }

static void mat_zero(om_mat* a) { memset(a, 0, sizeof(*a)); }

/* c = a * v, src/Omega_h_matrix.hpp:89-95 */
static void matvec(int n, om_mat const* a, double const* v, double* c) {
  double t[3];
  for (int i = 0; i < n; ++i) t[i] = a->m[0][i] * v[0];
  for (int j = 1; j < n; ++j)
    for (int i = 0; i < n; ++i) t[i] = t[i] + a->m[j][i] * v[j];
  for (int i = 0; i < n; ++i) c[i] = t[i];
}

/* c = a * b, src/Omega_h_matrix.hpp:97-101 */
static void matmul(int n, om_mat const* a, om_mat const* b, om_mat* c) {
  om_mat t;
  mat_zero(&t);
  for (int j = 0; j < n; ++j) matvec(n, a, b->m[j], t.m[j]);
  *c = t;
}

static void transpose(int n, om_mat const* a, om_mat* b) {
  om_mat t;
  mat_zero(&t);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) t.m[i][j] = a->m[j][i];
  *b = t;
}

static double dot(int n, double const* a, double const* b) {
  double out = a[0] * b[0];
  for (int i = 1; i < n; ++i) out = out + (a[i] * b[i]);
  return out;
}

static double vnorm(int n, double const* v) { return om_sqrt(dot(n, v, v)); }

static void cross3(double const* a, double const* b, double* c) {
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  c[0] = t0;
  c[1] = t1;
  c[2] = t2;
}

double om_det(int dim, om_mat const* m) {
  if (dim == 1) return m->m[0][0];
  if (dim == 2) {
    double a = m->m[0][0], b = m->m[1][0], c = m->m[0][1], d = m->m[1][1];
    return a * d - b * c;
  }
  double a = m->m[0][0], b = m->m[1][0], c = m->m[2][0];
  double d = m->m[0][1], e = m->m[1][1], f = m->m[2][1];
  double g = m->m[0][2], h = m->m[1][2], i = m->m[2][2];
  return (a * e * i) + (b * f * g) + (c * d * h) - (c * e * g) - (b * d * i) -
         (a * f * h);
}

static double trace(int n, om_mat const* a) {
  double t = a->m[0][0];
  for (int i = 1; i < n; ++i) t += a->m[i][i];
  return t;
}

static double max_norm(int n, om_mat const* a) {
  double x = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) x = max2(x, fabs(a->m[j][i]));
  return x;
}

static void identity(int n, om_mat* a) {
  mat_zero(a);
  for (int i = 0; i < n; ++i) a->m[i][i] = 1.0;
}

/* cubic x^3 + a2 x^2 + a1 x + a0, src/Omega_h_eigen.hpp:21-85 */
int om_roots3(double const coeffs[3], double eps, double roots[3], int mults[3]) {
  double a_0 = coeffs[0], a_1 = coeffs[1], a_2 = coeffs[2];
  roots[0] = roots[1] = roots[2] = 0.0;
  mults[0] = mults[1] = mults[2] = 0;
  double p = (3. * a_1 - a_2 * a_2) / 3.;
  double q = (9. * a_1 * a_2 - 27. * a_0 - 2. * (a_2 * (a_2 * a_2))) / 27.;
  double Q = p / 3.;
  double R = q / 2.;
  double D = (Q * (Q * Q)) + R * R;
  double shift = -a_2 / 3.;
  if (D >= 0.0) {
    double S = om_cbrt(R + om_sqrt(D));
    double T = om_cbrt(R - om_sqrt(D));
    double B = S + T;
    double z_1 = shift + B;
    double z_23_real = shift - (1. / 2.) * B;
    roots[0] = z_1;
    roots[1] = roots[2] = z_23_real;
  } else {
    double cos_theta = R / om_sqrt(-(Q * (Q * Q)));
    double theta = om_acos(min2(max2(cos_theta, -1.0), 1.0));
    double radius = 2. * om_sqrt(-Q);
    double z_1 = radius * om_cos((theta) / 3.) + shift;
    double z_2 = radius * om_cos((theta + 2. * OM_PI) / 3.) + shift;
    double z_3 = radius * om_cos((theta - 2. * OM_PI) / 3.) + shift;
    roots[0] = z_1;
    roots[1] = z_2;
    roots[2] = z_3;
  }
  mults[0] = mults[1] = mults[2] = 1;
  if (fabs(roots[0] - roots[1]) < eps) {
    double t = roots[0];
    roots[0] = roots[2];
    roots[2] = t;
  } else if (fabs(roots[0] - roots[2]) < eps) {
    double t = roots[0];
    roots[0] = roots[1];
    roots[1] = t;
  } else if (fabs(roots[1] - roots[2]) < eps) {
  } else {
    return 3;
  }
  roots[1] = (roots[1] + roots[2]) / 2;
  mults[1] = 2;
  if (fabs(roots[0] - roots[1]) < eps) {
    roots[0] = (1. / 3.) * roots[0] + (2. / 3.) * roots[1];
    mults[0] = 3;
    return 1;
  }
  return 2;
}

/* quadratic x^2 + a x + b, src/Omega_h_eigen.hpp:88-112 */
int om_roots2(double const coeffs[2], double eps, double roots[2], int mults[2]) {
  double a = coeffs[1];
  double b = coeffs[0];
  roots[0] = roots[1] = 0.0;
  mults[0] = mults[1] = 0;
  double disc = a * a - 4. * b;
  if (fabs(disc) < eps) {
    mults[0] = 2;
    roots[0] = -a / 2.;
    roots[1] = roots[0];
    mults[1] = mults[0];
    return 1;
  }
  if (disc > 0.0) {
    mults[0] = 1;
    mults[1] = 1;
    roots[0] = (-a + om_sqrt(disc)) / 2.;
    roots[1] = (-a - om_sqrt(disc)) / 2.;
    return 2;
  }
  return 0;
}

static void subtract_from_diag(int n, om_mat* a, double mu) {
  for (int i = 0; i < n; ++i) a->m[i][i] -= mu;
}

static int single_eigenvector3(om_mat const* m, double l, double* out) {
  om_mat s0 = *m;
  subtract_from_diag(3, &s0, l);
  om_mat s;
  transpose(3, &s0, &s);
  double v[3], c[3];
  cross3(s.m[0], s.m[1], v);
  double v_norm = vnorm(3, v);
  cross3(s.m[1], s.m[2], c);
  double c_norm = vnorm(3, c);
  if (c_norm > v_norm) {
    memcpy(v, c, sizeof(v));
    v_norm = c_norm;
  }
  cross3(s.m[0], s.m[2], c);
  c_norm = vnorm(3, c);
  if (c_norm > v_norm) {
    memcpy(v, c, sizeof(v));
    v_norm = c_norm;
  }
  if (!(v_norm > OM_EPSILON)) return -1;
  for (int i = 0; i < 3; ++i) out[i] = v[i] / v_norm;
  return 0;
}

static int row_space_1d(int n, om_mat const* a, double* out) {
  om_mat ta;
  transpose(n, a, &ta);
  int best_row = 0;
  double best_norm = vnorm(n, ta.m[0]);
  for (int i = 1; i < n; ++i) {
    double row_norm = vnorm(n, ta.m[i]);
    if (row_norm > best_norm) {
      best_row = i;
      best_norm = row_norm;
    }
  }
  if (!(best_norm > OM_EPSILON)) return -1;
  for (int i = 0; i < n; ++i) out[i] = ta.m[best_row][i] / best_norm;
  return 0;
}

/* Duff et al. 2017, src/Omega_h_matrix.hpp:567-575 */
static void ortho_basis3(double const* v, om_mat* a) {
  memcpy(a->m[0], v, 3 * sizeof(double));
  double sign = copysign(1.0, v[2]);
  double aa = -1.0 / (sign + v[2]);
  double b = v[0] * v[1] * aa;
  a->m[1][0] = 1.0 + sign * v[0] * v[0] * aa;
  a->m[1][1] = sign * b;
  a->m[1][2] = -sign * v[0];
  a->m[2][0] = b;
  a->m[2][1] = sign + v[1] * v[1] * aa;
  a->m[2][2] = -v[1];
}

/* the reference's OMEGA_H_CHECKs in single_eigenvector and
 * get_1d_row_space (src/Omega_h_eigen.hpp:191,211) abort */
static void eigen_check(int status) {
  if (status != 0) {
    fprintf(stderr, "oracle: no eigenvector for an eigenvalue\n");
    abort();
  }
}

static void decompose_dim(int dim, om_mat const* m, om_mat* q, double* l) {
  mat_zero(q);
  if (dim == 3) {
    double coeffs[3];
    double tA = trace(3, m);
    om_mat AA;
    matmul(3, m, m, &AA);
    coeffs[2] = -tA;
    coeffs[1] = (1. / 2.) * ((tA * tA) - trace(3, &AA));
    coeffs[0] = -om_det(3, m);
    double roots[3];
    int mults[3];
    int nroots = om_roots3(coeffs, 5e-5, roots, mults);
    if (nroots == 3) {
      for (int i = 0; i < 3; ++i) {
        eigen_check(single_eigenvector3(m, roots[i], q->m[i]));
        l[i] = roots[i];
      }
    } else if (nroots == 2 && mults[1] == 2) {
      eigen_check(single_eigenvector3(m, roots[0], q->m[0]));
      l[0] = roots[0];
      om_mat s = *m;
      subtract_from_diag(3, &s, roots[1]);
      double n[3] = {0.0, 0.0, 0.0};
      eigen_check(row_space_1d(3, &s, n));
      om_mat b;
      ortho_basis3(n, &b);
      memcpy(q->m[1], b.m[1], 3 * sizeof(double));
      memcpy(q->m[2], b.m[2], 3 * sizeof(double));
      l[1] = l[2] = roots[1];
    } else {
      l[0] = l[1] = l[2] = roots[0];
      identity(3, q);
    }
  } else if (dim == 2) {
    double coeffs[2];
    coeffs[1] = -trace(2, m);
    coeffs[0] = om_det(2, m);
    double roots[2];
    int mults[2];
    int nroots = om_roots2(coeffs, 5e-5, roots, mults);
    if (nroots == 2) {
      for (int i = 0; i < 2; ++i) {
        om_mat s = *m;
        subtract_from_diag(2, &s, roots[i]);
        double r[2] = {0.0, 0.0};
        eigen_check(row_space_1d(2, &s, r));
        q->m[i][0] = -r[1];
        q->m[i][1] = r[0];
        l[i] = roots[i];
      }
    } else {
      l[0] = l[1] = roots[0];
      identity(2, q);
    }
  } else {
    double root = -(-om_det(1, m));
    q->m[0][0] = 1.0;
    l[0] = root;
  }
}

/* get_max_eigenvalues, src/Omega_h_eigen.cpp:7-40: the root of largest
 * magnitude of the characteristic polynomial (src/Omega_h_eigen.hpp:127-156,
 * tolerance 5e-5, the matrix not normalised); a 1x1 metric is its value */
double om_max_eigenvalue(int dim, om_mat const* m) {
  if (dim == 1) return m->m[0][0];
  double roots[3] = {0.0, 0.0, 0.0};
  int mults[3];
  int n;
  if (dim == 3) {
    double coeffs[3];
    double tA = trace(3, m);
    om_mat AA;
    matmul(3, m, m, &AA);
    coeffs[2] = -tA;
    coeffs[1] = (1. / 2.) * ((tA * tA) - trace(3, &AA));
    coeffs[0] = -om_det(3, m);
    n = om_roots3(coeffs, 5e-5, roots, mults);
  } else {
    double coeffs[2];
    coeffs[1] = -trace(2, m);
    coeffs[0] = om_det(2, m);
    n = om_roots2(coeffs, 5e-5, roots, mults);
  }
  double out = roots[0];
  double max_mag = fabs(roots[0]);
  for (int i = 1; i < n; ++i) {
    double mag = fabs(roots[i]);
    if (mag > max_mag) {
      max_mag = mag;
      out = roots[i];
    }
  }
  return out;
}

void om_decompose_eigen(int dim, om_mat const* m0, om_mat* q, double* l) {
  double nm = max_norm(dim, m0);
  if (nm <= OM_EPSILON) {
    identity(dim, q);
    for (int i = 0; i < dim; ++i) l[i] = 0.0;
    return;
  }
  om_mat m;
  mat_zero(&m);
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < dim; ++i) m.m[j][i] = m0->m[j][i] / nm;
  decompose_dim(dim, &m, q, l);
  for (int i = 0; i < dim; ++i) l[i] = l[i] * nm;
}

void om_compose_ortho(int dim, om_mat const* q, double const* l, om_mat* out) {
  om_mat d;
  mat_zero(&d);
  for (int i = 0; i < dim; ++i) d.m[i][i] = l[i];
  om_mat qd, qt;
  matmul(dim, q, &d, &qd);
  transpose(dim, q, &qt);
  matmul(dim, &qd, &qt, out);
}

void om_log_spd(int dim, om_mat const* m, om_mat* out) {
  om_mat q;
  double l[3];
  om_decompose_eigen(dim, m, &q, l);
  for (int i = 0; i < dim; ++i) l[i] = om_log(l[i]);
  om_compose_ortho(dim, &q, l, out);
}

void om_exp_spd(int dim, om_mat const* m, om_mat* out) {
  om_mat q;
  double l[3];
  om_decompose_eigen(dim, m, &q, l);
  for (int i = 0; i < dim; ++i) l[i] = om_exp(l[i]);
  om_compose_ortho(dim, &q, l, out);
}

int om_symm_ncomps(int mdim) { return ((mdim + 1) * mdim) / 2; }

/* vector2symm / symm2vector, src/Omega_h_matrix.hpp:385-431 */
void om_get_symm(int mdim, double const* v, om_mat* m) {
  mat_zero(m);
  if (mdim == 1) {
    m->m[0][0] = v[0];
  } else if (mdim == 2) {
    m->m[0][0] = v[0];
    m->m[1][1] = v[1];
    m->m[1][0] = v[2];
    m->m[0][1] = v[2];
  } else {
    m->m[0][0] = v[0];
    m->m[1][1] = v[1];
    m->m[2][2] = v[2];
    m->m[1][0] = v[3];
    m->m[2][1] = v[4];
    m->m[2][0] = v[5];
    m->m[0][1] = v[3];
    m->m[1][2] = v[4];
    m->m[0][2] = v[5];
  }
}

void om_set_symm(int mdim, om_mat const* m, double* v) {
  if (mdim == 1) {
    v[0] = m->m[0][0];
  } else if (mdim == 2) {
    v[0] = m->m[0][0];
    v[1] = m->m[1][1];
    v[2] = m->m[1][0];
  } else {
    v[0] = m->m[0][0];
    v[1] = m->m[1][1];
    v[2] = m->m[2][2];
    v[3] = m->m[1][0];
    v[4] = m->m[2][1];
    v[5] = m->m[2][0];
  }
}

/* v . (M v); the isotropic case scales v first, src/Omega_h_metric.hpp:8-18 */
static double metric_product(int sdim, int mdim, om_mat const* m,
    double const* v) {
  double mv[3];
  if (mdim == 1) {
    for (int i = 0; i < sdim; ++i) mv[i] = m->m[0][0] * v[i];
  } else {
    matvec(sdim, m, v, mv);
  }
  return dot(sdim, v, mv);
}

static double aniso_length(double l_a, double l_b) {
  if (fabs(l_a - l_b) > 1e-3) return (l_a - l_b) / (om_log(l_a / l_b));
  return (l_a + l_b) / 2.;
}

double om_kat_aniso_length(double la, double lb) { return aniso_length(la, lb); }

double om_metric_edge_length(int sdim, int mdim, double const* p0,
    double const* p1, double const* m0, double const* m1) {
  double v[3];
  for (int i = 0; i < sdim; ++i) v[i] = p1[i] - p0[i];
  om_mat a, b;
  om_get_symm(mdim, m0, &a);
  om_get_symm(mdim, m1, &b);
  double l_a = om_sqrt(metric_product(sdim, mdim, &a, v));
  double l_b = om_sqrt(metric_product(sdim, mdim, &b, v));
  return aniso_length(l_a, l_b);
}

/* power<dim, 2*mdim>(det), src/Omega_h_shape.hpp:399-402 */
static double metric_size_factor(int dim, int mdim, double det) {
  if (dim == 3 && mdim == 3) return om_sqrt(det);
  if (dim == 3 && mdim == 1) return om_sqrt(det * (det * det));
  if (dim == 2 && mdim == 2) return om_sqrt(det);
  return det; /* dim 2, iso */
}

/* real_simplex_size / simplex_size_from_basis, src/Omega_h_shape.hpp:77-100,166-170:
 * the signed area (2D) or volume (3D) of a simplex from its basis vectors */
double om_element_size(int dim, double const* const* p) {
  double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) b[i][j] = p[i + 1][j] - p[0][j];
  if (dim == 3) {
    double c[3];
    cross3(b[0], b[1], c);
    return dot(3, c, b[2]) / 6.0;
  }
  return (b[0][0] * b[1][1] - b[0][1] * b[1][0]) / 2.0;
}

double om_element_quality(int dim, int mdim, double const* const* p,
    double const* const* ms) {
  /* maxdet_metric, src/Omega_h_metric.hpp:170-181 */
  om_mat m;
  om_get_symm(mdim, ms[0], &m);
  double maxdet = om_det(mdim, &m);
  for (int i = 1; i < dim + 1; ++i) {
    om_mat mi;
    om_get_symm(mdim, ms[i], &mi);
    double det = om_det(mdim, &mi);
    if (det > maxdet) {
      m = mi;
      maxdet = det;
    }
  }
  /* metric_element_quality, src/Omega_h_quality.hpp:27-37 */
  double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) b[i][j] = p[i + 1][j] - p[0][j];
  double rs;
  if (dim == 3) {
    double c[3];
    cross3(b[0], b[1], c);
    rs = dot(3, c, b[2]) / 6.0;
  } else {
    rs = (b[0][0] * b[1][1] - b[0][1] * b[1][0]) / 2.0;
  }
  double s = rs * metric_size_factor(dim, mdim, om_det(mdim, &m));
  if (s < 0) return s;
  double ev[6][3];
  int nedges;
  if (dim == 3) {
    nedges = 6;
    for (int j = 0; j < 3; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
      ev[3][j] = b[2][j];
      ev[4][j] = p[3][j] - p[1][j];
      ev[5][j] = p[3][j] - p[2][j];
    }
  } else {
    nedges = 3;
    for (int j = 0; j < 2; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
    }
  }
  double msl = 0;
  for (int i = 0; i < nedges; ++i) msl += metric_product(dim, mdim, &m, ev[i]);
  msl = msl / nedges;
  double eq = (dim == 3) ? 0.1178511301977579 : 0.4330127018922193;
  double x = s / eq;
  double num = (dim == 3) ? om_cbrt(x * x) : x;
  return num / msl;
}

/* average_metric with has_degen == false, src/Omega_h_metric.hpp:125-149 */
void om_average_metric(int mdim, int n, double const* const* ms, double* out) {
  om_mat am;
  mat_zero(&am);
  for (int k = 0; k < n; ++k) {
    om_mat m, lm;
    om_get_symm(mdim, ms[k], &m);
    om_log_spd(mdim, &m, &lm);
    for (int j = 0; j < mdim; ++j)
      for (int i = 0; i < mdim; ++i) am.m[j][i] = am.m[j][i] + lm.m[j][i];
  }
  for (int j = 0; j < mdim; ++j)
    for (int i = 0; i < mdim; ++i) am.m[j][i] = am.m[j][i] / (double)n;
  om_mat e;
  om_exp_spd(mdim, &am, &e);
  om_set_symm(mdim, &e, out);
}

/* barycentric_from_global (src/Omega_h_shape.hpp:46-63) of point x in the
 * simplex p: the basis b[i] = p[i+1] - p[0] inverted by the cofactor
 * formulas of src/Omega_h_matrix.hpp:332-350, times x - p[0] */
void om_barycentric(int dim, double const* x, double const* const* p, double* xi) {
  om_mat b, inv;
  mat_zero(&b);
  mat_zero(&inv);
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) b.m[i][j] = p[i + 1][j] - p[0][j];
  double det = om_det(dim, &b);
  if (dim == 3) {
    double bc[3][3];
    cross3(b.m[1], b.m[2], bc[0]);
    cross3(b.m[2], b.m[0], bc[1]);
    cross3(b.m[0], b.m[1], bc[2]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) inv.m[j][i] = bc[i][j] / det;
  } else {
    double a = b.m[0][0], bb = b.m[1][0], c = b.m[0][1], d = b.m[1][1];
    inv.m[0][0] = d / det;
    inv.m[0][1] = -c / det;
    inv.m[1][0] = -bb / det;
    inv.m[1][1] = a / det;
  }
  double v[3], lam[3];
  for (int i = 0; i < dim; ++i) v[i] = x[i] - p[0][i];
  for (int i = 0; i < dim; ++i) lam[i] = inv.m[0][i] * v[0];
  for (int j = 1; j < dim; ++j)
    for (int i = 0; i < dim; ++i) lam[i] = lam[i] + inv.m[j][i] * v[j];
  xi[0] = 1.0;
  for (int i = 1; i < dim + 1; ++i) {
    xi[i] = lam[i - 1];
    xi[0] -= lam[i - 1];
  }
}

/* ---- known-answer hooks ---- */

void om_kat_barycentric(int dim, double const* x, double const* simplex, double* xi) {
  double const* p[4];
  for (int i = 0; i <= dim; ++i) p[i] = simplex + i * dim;
  om_barycentric(dim, x, p, xi);
}

double om_kat_max_eigenvalue(int dim, double const* symm) {
  om_mat m;
  om_get_symm(dim, symm, &m);
  return om_max_eigenvalue(dim, &m);
}

int om_kat_roots3(double const* coeffs, double eps, double* roots, int* mults) {
  return om_roots3(coeffs, eps, roots, mults);
}

static void load_colmajor(int dim, double const* a, om_mat* m) {
  mat_zero(m);
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < dim; ++i) m->m[j][i] = a[j * dim + i];
}
static void store_colmajor(int dim, om_mat const* m, double* a) {
  for (int j = 0; j < dim; ++j)
    for (int i = 0; i < dim; ++i) a[j * dim + i] = m->m[j][i];
}

void om_kat_decompose_eigen(int dim, double const* a, double* q, double* l) {
  om_mat m, qq;
  load_colmajor(dim, a, &m);
  om_decompose_eigen(dim, &m, &qq, l);
  store_colmajor(dim, &qq, q);
}

void om_kat_log_spd(int dim, double const* a, double* out) {
  om_mat m, r;
  load_colmajor(dim, a, &m);
  om_log_spd(dim, &m, &r);
  store_colmajor(dim, &r, out);
}

void om_kat_exp_spd(int dim, double const* a, double* out) {
  om_mat m, r;
  load_colmajor(dim, a, &m);
  om_exp_spd(dim, &m, &r);
  store_colmajor(dim, &r, out);
}

double om_kat_quality(int dim, int mdim, double const* pts, double const* ms) {
  double const* p[4];
  double const* mm[4];
  int nc = om_symm_ncomps(mdim);
  for (int i = 0; i < dim + 1; ++i) {
    p[i] = pts + i * dim;
    mm[i] = ms + i * nc;
  }
  return om_element_quality(dim, mdim, p, mm);
}

double om_kat_edge_length(int sdim, int mdim, double const* p0,
    double const* p1, double const* m0, double const* m1) {
  return om_metric_edge_length(sdim, mdim, p0, p1, m0, m1);
}

void om_kat_average_metric(int mdim, int n, double const* ms, double* out) {
  double const* p[16];
  int nc = om_symm_ncomps(mdim);
  for (int i = 0; i < n && i < 16; ++i) p[i] = ms + i * nc;
  om_average_metric(mdim, n, p, out);
}

/* ------------------------------------------------ side normals, most normal */

static double dot3(double const* a, double const* b, int n) {
  double out = a[0] * b[0];
  for (int i = 1; i < n; ++i) out = out + (a[i] * b[i]);
  return out;
}

/* normalize, src/Omega_h_vector.hpp:152-154: each component divided by the
 * norm sqrt(v . v) */
static void normalize_n(double* v, int n) {
  double l = om_sqrt(dot3(v, v, n));
  for (int i = 0; i < n; ++i) v[i] = v[i] / l;
}

/* get_side_vector, src/Omega_h_shape.hpp:276-298: in 3D the cross product
 * (b - a) x (c - a) of the triangle (a, b, c); in 2D -perp(b - a) */
void om_side_vector(int dim, double const* const* p, double* out) {
  if (dim == 3) {
    double u[3], w[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = p[1][i] - p[0][i];
      w[i] = p[2][i] - p[0][i];
    }
    out[0] = u[1] * w[2] - u[2] * w[1];
    out[1] = u[2] * w[0] - u[0] * w[2];
    out[2] = u[0] * w[1] - u[1] * w[0];
  } else {
    double v0 = p[1][0] - p[0][0], v1 = p[1][1] - p[0][1];
    out[0] = -(-v1);
    out[1] = -v0;
  }
}

/* get_most_normal_normal, src/Omega_h_most_normal.hpp:15-88 (Aubry and
 * Lohner, "On the 'most normal' normal", 2008): the centre of the smallest
 * circle on the unit sphere holding the n unit vectors N, tried first
 * through pairs (the normalized bisector), then through triplets (the
 * direction with equal dot products with all three, from the inverse of the
 * matrix whose rows are the three vectors).  In 2D it is the normalized sum
 * of the first two (src/Omega_h_most_normal.hpp:90-93).  Returns 0 on
 * success, 1 where the reference's checks fail (a triplet centre not
 * equidistant, or no circle found). */
int om_most_normal(int dim, int n, double const* N, double* c) {
  double const inside_tol = 1e-10;
  if (dim == 2) {
    c[0] = N[0] + N[2];
    c[1] = N[1] + N[3];
    normalize_n(c, 2);
    return 0;
  }
  double scalmin = -1.0;
  c[0] = c[1] = c[2] = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      double b[3];
      for (int k = 0; k < 3; ++k) b[k] = N[i * 3 + k] + N[j * 3 + k];
      normalize_n(b, 3);
      double scal = dot3(b, N + i * 3, 3);
      if (scal < scalmin) continue;
      int l;
      for (l = 0; l < n; ++l) {
        double scalt = dot3(N + l * 3, b, 3);
        if (scalt + inside_tol < scal) break;
      }
      if (l < n) continue;
      for (int k = 0; k < 3; ++k) c[k] = b[k];
      scalmin = scal;
    }
  }
  if (scalmin != -1.0) return 0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      for (int k = j + 1; k < n; ++k) {
        /* M's columns are N_i, N_j, N_k (src/Omega_h_matrix.hpp:14) */
        om_mat M;
        for (int r = 0; r < 3; ++r) {
          M.m[0][r] = N[i * 3 + r];
          M.m[1][r] = N[j * 3 + r];
          M.m[2][r] = N[k * 3 + r];
        }
        if (fabs(om_det(3, &M)) < 1e-6) continue;
        /* invert(transpose(M)) * (1, 1, 1), src/Omega_h_matrix.hpp:89-94,
         * 281-293,344-350: T = transpose(M); b = rows of cross products of
         * T's columns; inverse = transpose(b) / det(T) */
        om_mat T;
        for (int a = 0; a < 3; ++a)
          for (int r = 0; r < 3; ++r) T.m[a][r] = M.m[r][a];
        double bm[3][3];
        for (int a = 0; a < 3; ++a) {
          double const* x = T.m[(a + 1) % 3];
          double const* y = T.m[(a + 2) % 3];
          bm[a][0] = x[1] * y[2] - x[2] * y[1];
          bm[a][1] = x[2] * y[0] - x[0] * y[2];
          bm[a][2] = x[0] * y[1] - x[1] * y[0];
        }
        double dt = om_det(3, &T);
        double nc[3];
        for (int r = 0; r < 3; ++r) {
          /* column q of the inverse is (b[0][q], b[1][q], b[2][q]) / det;
           * the product with (1, 1, 1) sums the columns in order */
          double s = (bm[r][0] / dt) * 1.0;
          s = s + (bm[r][1] / dt) * 1.0;
          s = s + (bm[r][2] / dt) * 1.0;
          nc[r] = s;
        }
        /* the sum above is column-wise: component r of A * v is
         * A[0][r] v0 + A[1][r] v1 + A[2][r] v2 with A[q][r] = b[r][q] / det */
        normalize_n(nc, 3);
        double scal = dot3(nc, N + i * 3, 3);
        double sj = dot3(nc, N + j * 3, 3), sk = dot3(nc, N + k * 3, 3);
        if (!om_are_close(1, &scal, &sj, 1e-10, 1e-10) ||
            !om_are_close(1, &scal, &sk, 1e-10, 1e-10))
          return 1;
        if (scal < scalmin) continue;
        int l;
        for (l = 0; l < n; ++l) {
          double scalt = dot3(N + l * 3, nc, 3);
          if (scalt + inside_tol < scal) break;
        }
        if (l < n) continue;
        for (int q = 0; q < 3; ++q) c[q] = nc[q];
        scalmin = scal;
      }
    }
  }
  return (scalmin > -1.0) ? 0 : 1;
}
