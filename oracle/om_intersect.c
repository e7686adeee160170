/* Oracle (test infrastructure only): metric intersection,
 * intersect_metrics src/Omega_h_metric_intersect.hpp:215-269, for the case
 * where at least one metric is non-degenerate (and the all-null case).
 * The partially degenerate branches (Barral's cases 2-5,
 * src/Omega_h_metric_intersect.hpp:18-213) return m1 here: parity unpinned. */
#include <math.h>
#include <string.h>

#include "om_internal.h"

static void matmul_n(int n, om_mat const* a, om_mat const* b, om_mat* c) {
  om_mat t;
  memset(&t, 0, sizeof(t));
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) t.m[j][i] = a->m[0][i] * b->m[j][0];
    for (int k = 1; k < n; ++k)
      for (int i = 0; i < n; ++i) t.m[j][i] = t.m[j][i] + a->m[k][i] * b->m[j][k];
  }
  *c = t;
}

static void transpose_n(int n, om_mat const* a, om_mat* b) {
  om_mat t;
  memset(&t, 0, sizeof(t));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) t.m[i][j] = a->m[j][i];
  *b = t;
}

void om_intersect_metrics(int dim, double const* v1, double const* v2, double* out) {
  om_mat m1, m2, q1, q2;
  double l1[3], l2[3];
  om_get_symm(dim, v1, &m1);
  om_get_symm(dim, v2, &m2);
  om_decompose_eigen(dim, &m1, &q1, l1);
  om_decompose_eigen(dim, &m2, &q2, l2);
  int n1 = 0, n2 = 0;
  for (int i = 0; i < dim; ++i) {
    n1 += l1[i] < 1e-10;
    n2 += l2[i] < 1e-10;
  }
  if (n1 > n2) {
    om_mat t = m1; m1 = m2; m2 = t;
    t = q1; q1 = q2; q2 = t;
    for (int i = 0; i < dim; ++i) { double x = l1[i]; l1[i] = l2[i]; l2[i] = x; }
    int x = n1; n1 = n2; n2 = x;
  }
  om_mat r;
  if (n1 == dim) {
    memset(&r, 0, sizeof(r));
  } else if (n1 == 0) {
    double ls[3], lis[3];
    for (int i = 0; i < dim; ++i) {
      ls[i] = om_sqrt(l1[i]);
      lis[i] = 1.0 / ls[i];
    }
    om_mat s, is, ist, a, m2bar, p, st, b;
    om_compose_ortho(dim, &q1, ls, &s);
    om_compose_ortho(dim, &q1, lis, &is);
    transpose_n(dim, &is, &ist);
    matmul_n(dim, &ist, &m2, &a);
    matmul_n(dim, &a, &is, &m2bar);
    double lb[3];
    om_decompose_eigen(dim, &m2bar, &p, lb);
    for (int i = 0; i < dim; ++i) lb[i] = (lb[i] < 1.0) ? 1.0 : lb[i];
    om_mat mib;
    om_compose_ortho(dim, &p, lb, &mib);
    transpose_n(dim, &s, &st);
    matmul_n(dim, &st, &mib, &b);
    matmul_n(dim, &b, &s, &r);
  } else {
    r = m1;
  }
  om_set_symm(dim, &r, out);
}

void om_kat_intersect(int dim, double const* m1, double const* m2, double* out) {
  om_intersect_metrics(dim, m1, m2, out);
}
