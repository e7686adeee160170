/* Oracle (test infrastructure only): metric intersection and gradation
 * limiting.
 *
 *   intersect_metrics              src/Omega_h_metric_intersect.hpp:215-269
 *   intersect_degenerate_metrics   src/Omega_h_metric_intersect.hpp:18-213
 *     (Barral's cases 2-5: partially degenerate metrics)
 *   invert (2x2, 3x3)              src/Omega_h_matrix.hpp:336-350
 *   limit_gradation_once_tmpl      src/Omega_h_metric.cpp:156-180
 *   limit_metric_gradation         src/Omega_h_metric.cpp:200-220
 *   are_close / rel_diff_with_floor src/Omega_h_scalar.hpp:266-277,
 *                                  src/Omega_h_array_ops.cpp:75-87
 * Matrices are column-major (m[col][row]); products follow the reference's
 * column-by-column order (c[j] = a * b[j], a * v summed from column 0). */
#include <math.h>
#include <string.h>

#include "om_internal.h"

#define OM_EPSILON 1e-10

static double max2(double a, double b) { return (a < b) ? b : a; }

/* c (m x n) = a (m x p) * b (p x n) */
static void matmul_r(int m, int p, int n, om_mat const* a, om_mat const* b, om_mat* c) {
  om_mat t;
  memset(&t, 0, sizeof(t));
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) t.m[j][i] = a->m[0][i] * b->m[j][0];
    for (int k = 1; k < p; ++k)
      for (int i = 0; i < m; ++i) t.m[j][i] = t.m[j][i] + a->m[k][i] * b->m[j][k];
  }
  *c = t;
}

/* b (n x m) = transpose of a (m x n) */
static void transpose_r(int m, int n, om_mat const* a, om_mat* b) {
  om_mat t;
  memset(&t, 0, sizeof(t));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) t.m[i][j] = a->m[j][i];
  *b = t;
}

static void matvec_n(int n, om_mat const* a, double const* v, double* c) {
  double t[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; ++i) t[i] = a->m[0][i] * v[0];
  for (int j = 1; j < n; ++j)
    for (int i = 0; i < n; ++i) t[i] = t[i] + a->m[j][i] * v[j];
  for (int i = 0; i < n; ++i) c[i] = t[i];
}

static double dot_n(int n, double const* a, double const* b) {
  double r = a[0] * b[0];
  for (int i = 1; i < n; ++i) r = r + (a[i] * b[i]);
  return r;
}

/* v . (m v) */
static double vmv(int n, om_mat const* m, double const* v) {
  double mv[3] = {0.0, 0.0, 0.0};
  matvec_n(n, m, v, mv);
  return dot_n(n, v, mv);
}

static void cross(double const* a, double const* b, double* c) {
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  c[0] = t0;
  c[1] = t1;
  c[2] = t2;
}

static void normalize3(double* v) {
  double n = om_sqrt(dot_n(3, v, v));
  for (int i = 0; i < 3; ++i) v[i] = v[i] / n;
}

static void scale(int m, int n, om_mat* a, double s) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) a->m[j][i] = a->m[j][i] * s;
}

static void outer(int n, double const* u, om_mat* c) {
  memset(c, 0, sizeof(*c));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) c->m[j][i] = u[i] * u[j];
}

/* src/Omega_h_matrix.hpp:336-350 */
static void invert_n(int n, om_mat const* a, om_mat* out) {
  om_mat r;
  memset(&r, 0, sizeof(r));
  double d = om_det(n, a);
  if (n == 1) {
    r.m[0][0] = 1.0 / a->m[0][0];
  } else if (n == 2) {
    double A = a->m[0][0], B = a->m[1][0], C = a->m[0][1], D = a->m[1][1];
    /* matrix_2x2(D, -B, -C, A): columns (D, -C) and (-B, A) */
    r.m[0][0] = D;
    r.m[0][1] = -C;
    r.m[1][0] = -B;
    r.m[1][1] = A;
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) r.m[j][i] = r.m[j][i] / d;
  } else {
    om_mat b;
    cross(a->m[1], a->m[2], b.m[0]);
    cross(a->m[2], a->m[0], b.m[1]);
    cross(a->m[0], a->m[1], b.m[2]);
    transpose_r(3, 3, &b, &r);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) r.m[j][i] = r.m[j][i] / d;
  }
  *out = r;
}

/* transpose(pinv) * diagonal(l) * pinv with pinv = invert(P) */
static void congruence_diag(int n, om_mat const* P, double const* l, om_mat* out) {
  om_mat pinv, pinvt, dg, t;
  invert_n(n, P, &pinv);
  transpose_r(n, n, &pinv, &pinvt);
  memset(&dg, 0, sizeof(dg));
  for (int i = 0; i < n; ++i) dg.m[i][i] = l[i];
  matmul_r(n, n, n, &pinvt, &dg, &t);
  matmul_r(n, n, n, &t, &pinv, out);
}

static void intersect_m(int dim, om_mat const* m1_in, om_mat const* m2_in, om_mat* out);

/* 3D cases 3.a / 4.a: reduce to the plane spanned by u1, v1 */
static void intersect_in_plane(om_mat const* m1, om_mat const* m2, double const* u1,
    double const* v1, om_mat* out) {
  om_mat P, PT, t, m1b, m2b, mint;
  memset(&P, 0, sizeof(P));
  for (int i = 0; i < 3; ++i) {
    P.m[0][i] = u1[i];
    P.m[1][i] = v1[i];
  }
  transpose_r(3, 2, &P, &PT);
  matmul_r(3, 3, 2, m1, &P, &t);
  matmul_r(2, 3, 2, &PT, &t, &m1b);
  matmul_r(3, 3, 2, m2, &P, &t);
  matmul_r(2, 3, 2, &PT, &t, &m2b);
  intersect_m(2, &m1b, &m2b, &mint);
  matmul_r(2, 2, 3, &mint, &PT, &t);
  matmul_r(3, 2, 3, &P, &t, out);
}

static void intersect_degenerate(int dim, om_mat const* m1, om_mat const* q1,
    double const* l1v, int const* d1, int n1, om_mat const* m2, om_mat const* q2,
    double const* l2v, int const* d2, int n2, om_mat* out) {
  if (dim == 1) {
    *out = *m1;
    return;
  }
  if (dim == 2) {
    /* case 5 */
    double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0};
    double l1 = -1.0, l2 = -1.0;
    for (int i = 0; i < 2; ++i) {
      if (d1[i])
        memcpy(v1, q1->m[i], sizeof(v1));
      else {
        memcpy(u1, q1->m[i], sizeof(u1));
        l1 = l1v[i];
      }
      if (d2[i])
        memcpy(v2, q2->m[i], sizeof(v2));
      else
        l2 = l2v[i];
    }
    if (dot_n(2, v1, v2) < OM_EPSILON) {
      outer(2, u1, out);
      scale(2, 2, out, max2(l1, l2));
    } else {
      om_mat P;
      memset(&P, 0, sizeof(P));
      for (int i = 0; i < 2; ++i) {
        P.m[0][i] = v1[i];
        P.m[1][i] = v2[i];
      }
      double l[2] = {vmv(2, m2, v1), vmv(2, m1, v2)};
      congruence_diag(2, &P, l, out);
    }
    return;
  }
  if (n1 == 2 && n2 == 2) {
    /* case 2 */
    double u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
    double l1 = -1.0, l2 = -1.0;
    for (int i = 0; i < 3; ++i) {
      if (!d1[i]) {
        memcpy(u1, q1->m[i], sizeof(u1));
        l1 = l1v[i];
      }
      if (!d2[i]) {
        memcpy(u2, q2->m[i], sizeof(u2));
        l2 = l2v[i];
      }
    }
    double u[3];
    cross(u1, u2, u);
    if (dot_n(3, u, u) < OM_EPSILON) {
      outer(3, u1, out);
      scale(3, 3, out, max2(l1, l2));
    } else {
      normalize3(u);
      double e1[3], e2[3];
      cross(u1, u, e1);
      cross(u2, u, e2);
      om_mat P;
      memcpy(P.m[0], e1, sizeof(e1));
      memcpy(P.m[1], e2, sizeof(e2));
      memcpy(P.m[2], u, sizeof(u));
      double l[3] = {vmv(3, m2, e1), vmv(3, m1, e2), 0.0};
      congruence_diag(3, &P, l, out);
    }
    return;
  }
  if (n1 == 1 && n2 == 2) {
    /* case 3 (the roles of 1 and 2 swapped relative to Barral) */
    double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, w1[3] = {0, 0, 0};
    double u2[3] = {0, 0, 0}, v2[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
    int found_u1 = 0, found_v2 = 0;
    for (int i = 0; i < 3; ++i) {
      if (d1[i])
        memcpy(w1, q1->m[i], sizeof(w1));
      else if (found_u1)
        memcpy(v1, q1->m[i], sizeof(v1));
      else {
        memcpy(u1, q1->m[i], sizeof(u1));
        found_u1 = 1;
      }
      if (!d2[i])
        memcpy(u2, q2->m[i], sizeof(u2));
      else if (found_v2)
        memcpy(w2, q2->m[i], sizeof(w2));
      else {
        memcpy(v2, q2->m[i], sizeof(v2));
        found_v2 = 1;
      }
    }
    if (dot_n(3, u2, w1) < OM_EPSILON) {
      intersect_in_plane(m1, m2, u1, v1, out);
    } else {
      om_mat P;
      memcpy(P.m[0], v2, sizeof(v2));
      memcpy(P.m[1], w2, sizeof(w2));
      memcpy(P.m[2], w1, sizeof(w1));
      double l[3] = {vmv(3, m1, P.m[0]), vmv(3, m1, P.m[1]), vmv(3, m2, P.m[2])};
      congruence_diag(3, &P, l, out);
    }
    return;
  }
  if (n1 == 1 && n2 == 1) {
    /* case 4 */
    double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, w1[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
    int found_u1 = 0;
    for (int i = 0; i < 3; ++i) {
      if (d1[i])
        memcpy(w1, q1->m[i], sizeof(w1));
      else if (found_u1)
        memcpy(v1, q1->m[i], sizeof(v1));
      else {
        memcpy(u1, q1->m[i], sizeof(u1));
        found_u1 = 1;
      }
      if (d2[i]) memcpy(w2, q2->m[i], sizeof(w2));
    }
    double w[3];
    cross(w1, w2, w);
    if (dot_n(3, w, w) < OM_EPSILON) {
      intersect_in_plane(m1, m2, u1, v1, out);
    } else {
      normalize3(w);
      om_mat P;
      memcpy(P.m[0], w1, sizeof(w1));
      memcpy(P.m[1], w2, sizeof(w2));
      memcpy(P.m[2], w, sizeof(w));
      double l[3] = {vmv(3, m2, P.m[0]), vmv(3, m1, P.m[1]),
          max2(vmv(3, m1, P.m[2]), vmv(3, m2, P.m[2]))};
      congruence_diag(3, &P, l, out);
    }
    return;
  }
  *out = *m1;
}

static void intersect_m(int dim, om_mat const* m1_in, om_mat const* m2_in, om_mat* out) {
  om_mat m1 = *m1_in, m2 = *m2_in, q1, q2;
  double l1[3], l2[3];
  om_decompose_eigen(dim, &m1, &q1, l1);
  om_decompose_eigen(dim, &m2, &q2, l2);
  int d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0};
  int n1 = 0, n2 = 0;
  for (int i = 0; i < dim; ++i) {
    d1[i] = l1[i] < OM_EPSILON;
    d2[i] = l2[i] < OM_EPSILON;
    n1 += d1[i];
    n2 += d2[i];
  }
  if (n1 > n2) {
    om_mat t = m1; m1 = m2; m2 = t;
    t = q1; q1 = q2; q2 = t;
    for (int i = 0; i < 3; ++i) {
      double x = l1[i]; l1[i] = l2[i]; l2[i] = x;
      int y = d1[i]; d1[i] = d2[i]; d2[i] = y;
    }
    int x = n1; n1 = n2; n2 = x;
  }
  if (n1 == dim) {
    memset(out, 0, sizeof(*out));
    return;
  }
  if (n1 == 0) {
    double ls[3], lis[3];
    for (int i = 0; i < dim; ++i) {
      ls[i] = om_sqrt(l1[i]);
      lis[i] = 1.0 / ls[i];
    }
    om_mat s, is, ist, a, m2bar, p, st, b;
    om_compose_ortho(dim, &q1, ls, &s);
    om_compose_ortho(dim, &q1, lis, &is);
    transpose_r(dim, dim, &is, &ist);
    matmul_r(dim, dim, dim, &ist, &m2, &a);
    matmul_r(dim, dim, dim, &a, &is, &m2bar);
    double lb[3];
    om_decompose_eigen(dim, &m2bar, &p, lb);
    for (int i = 0; i < dim; ++i) lb[i] = max2(lb[i], 1.0);
    om_mat mib;
    om_compose_ortho(dim, &p, lb, &mib);
    transpose_r(dim, dim, &s, &st);
    matmul_r(dim, dim, dim, &st, &mib, &b);
    matmul_r(dim, dim, dim, &b, &s, out);
    return;
  }
  intersect_degenerate(dim, &m1, &q1, l1, d1, n1, &m2, &q2, l2, d2, n2, out);
}

void om_intersect_metrics(int dim, double const* v1, double const* v2, double* out) {
  om_mat m1, m2, r;
  om_get_symm(dim, v1, &m1);
  om_get_symm(dim, v2, &m2);
  intersect_m(dim, &m1, &m2, &r);
  om_set_symm(dim, &r, out);
}

void om_kat_intersect(int dim, double const* m1, double const* m2, double* out) {
  om_intersect_metrics(dim, m1, m2, out);
}

/* one sweep of gradation limiting, src/Omega_h_metric.cpp:156-180: each
 * vertex metric is intersected, in star order, with every neighbour's
 * metric scaled by 1 / (1 + rate * metric distance)^2 */
void om_limit_gradation_once(int sdim, int mdim, int nverts, int const* a2ab, int const* ab2b,
    double const* coords, double const* in, double max_rate, double* out) {
  int nc = om_symm_ncomps(mdim);
  for (int v = 0; v < nverts; ++v) {
    om_mat m;
    om_get_symm(mdim, in + (size_t)v * nc, &m);
    double const* x = coords + (size_t)v * sdim;
    for (int j = a2ab[v]; j < a2ab[v + 1]; ++j) {
      int av = ab2b[j];
      om_mat am;
      om_get_symm(mdim, in + (size_t)av * nc, &am);
      double vec[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < sdim; ++i) vec[i] = coords[(size_t)av * sdim + i] - x[i];
      double prod;
      if (mdim == 1) {
        double mv[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < sdim; ++i) mv[i] = am.m[0][0] * vec[i];
        prod = dot_n(sdim, vec, mv);
      } else {
        prod = vmv(sdim, &am, vec);
      }
      double dist = om_sqrt(prod);
      double h = 1.0 + dist * max_rate;
      double factor = 1.0 / (h * h);
      scale(mdim, mdim, &am, factor);
      om_mat lim;
      intersect_m(mdim, &m, &am, &lim);
      m = lim;
    }
    om_set_symm(mdim, &m, out + (size_t)v * nc);
  }
}

static int are_close_all(int n, double const* a, double const* b, double tol, double floor) {
  for (int i = 0; i < n; ++i) {
    double am = fabs(a[i]), bm = fabs(b[i]);
    double rd = (am <= floor && bm <= floor) ? 0.0 : fabs(b[i] - a[i]) / max2(am, bm);
    if (!(rd <= tol)) return 0;
  }
  return 1;
}

int om_are_close(int n, double const* a, double const* b, double tol, double floor) {
  return are_close_all(n, a, b, tol, floor);
}
