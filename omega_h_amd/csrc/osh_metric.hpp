// Metric intersection (device and host), including the partially degenerate
// cases.
//
//   intersect_metrics             src/Omega_h_metric_intersect.hpp:215-269
//   intersect_degenerate_metrics  src/Omega_h_metric_intersect.hpp:18-213
//                                 (Barral's cases 2-5)
//   invert                        src/Omega_h_matrix.hpp:336-350
// Products keep the reference's order: c[j] = a * b[j], with a * v summed
// from column 0, so the bits match a host evaluation of the same steps.
#pragma once

#include "osh_math.hpp"

namespace osh {

constexpr double kMetricEpsilon = 1e-10;

// c (R x C) = a (R x P) * b (P x C)
template <int R, int P, int C>
OSH_HD Mat matmul_rect(Mat const& a, Mat const& b) {
  Mat t;
  mat_zero(t);
  for (int j = 0; j < C; ++j) {
    for (int i = 0; i < R; ++i) t.m[j][i] = a.m[0][i] * b.m[j][0];
    for (int k = 1; k < P; ++k)
      for (int i = 0; i < R; ++i) t.m[j][i] = t.m[j][i] + a.m[k][i] * b.m[j][k];
  }
  return t;
}

// transpose of an R x C matrix
template <int R, int C>
OSH_HD Mat transpose_rect(Mat const& a) {
  Mat t;
  mat_zero(t);
  for (int j = 0; j < C; ++j)
    for (int i = 0; i < R; ++i) t.m[i][j] = a.m[j][i];
  return t;
}

template <int N>
OSH_HD double vmv(Mat const& m, double const* v) {
  double mv[3] = {0.0, 0.0, 0.0};
  matvec<N>(m, v, mv);
  return dot<N>(v, mv);
}

template <int N>
OSH_HD Mat invert(Mat const& a) {
  Mat r;
  mat_zero(r);
  double d = det<N>(a);
  if (N == 1) {
    r.m[0][0] = 1.0 / a.m[0][0];
  } else if (N == 2) {
    double A = a.m[0][0], B = a.m[1][0], C = a.m[0][1], D = a.m[1][1];
    r.m[0][0] = D;
    r.m[0][1] = -C;
    r.m[1][0] = -B;
    r.m[1][1] = A;
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) r.m[j][i] = r.m[j][i] / d;
  } else {
    Mat b;
    cross3(a.m[1], a.m[2], b.m[0]);
    cross3(a.m[2], a.m[0], b.m[1]);
    cross3(a.m[0], a.m[1], b.m[2]);
    r = transpose_rect<3, 3>(b);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) r.m[j][i] = r.m[j][i] / d;
  }
  return r;
}

// transpose(P^-1) * diagonal(l) * P^-1
template <int N>
OSH_HD Mat congruence_diag(Mat const& p, double const* l) {
  Mat pinv = invert<N>(p);
  Mat pinvt = transpose_rect<N, N>(pinv);
  Mat dg;
  mat_zero(dg);
  for (int i = 0; i < N; ++i) dg.m[i][i] = l[i];
  return matmul_rect<N, N, N>(matmul_rect<N, N, N>(pinvt, dg), pinv);
}

OSH_HD Mat scaled_outer(int n, double const* u, double s) {
  Mat c;
  mat_zero(c);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) c.m[j][i] = (u[i] * u[j]) * s;
  return c;
}

OSH_HD void copy3(double* d, double const* s) {
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
}

OSH_HD void normalize3(double* v) {
  double n = math::sqrt_(dot<3>(v, v));
  for (int i = 0; i < 3; ++i) v[i] = v[i] / n;
}

// *bad (when given) is set when an eigendecomposition fails (the reference
// aborts there)
template <int N>
OSH_HD Mat intersect_metrics(Mat const& m1_in, Mat const& m2_in, int* bad = nullptr);

// cases 3.a and 4.a: the intersection within the plane of u1 and v1
OSH_HD Mat intersect_in_plane(Mat const& m1, Mat const& m2, double const* u1,
    double const* v1);

OSH_HD Mat intersect_degenerate2(Mat const& m1, Mat const& q1, double const* l1v,
    int const* d1, Mat const& m2, Mat const& q2, double const* l2v, int const* d2) {
  // case 5
  double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0};
  double l1 = -1.0, l2 = -1.0;
  for (int i = 0; i < 2; ++i) {
    if (d1[i]) {
      copy3(v1, q1.m[i]);
    } else {
      copy3(u1, q1.m[i]);
      l1 = l1v[i];
    }
    if (d2[i])
      copy3(v2, q2.m[i]);
    else
      l2 = l2v[i];
  }
  if (dot<2>(v1, v2) < kMetricEpsilon) return scaled_outer(2, u1, math::max2(l1, l2));
  Mat p;
  mat_zero(p);
  for (int i = 0; i < 2; ++i) {
    p.m[0][i] = v1[i];
    p.m[1][i] = v2[i];
  }
  double l[2] = {vmv<2>(m2, v1), vmv<2>(m1, v2)};
  return congruence_diag<2>(p, l);
}

OSH_HD Mat intersect_degenerate3(Mat const& m1, Mat const& q1, double const* l1v,
    int const* d1, int n1, Mat const& m2, Mat const& q2, double const* l2v, int const* d2,
    int n2) {
  if (n1 == 2 && n2 == 2) {
    // case 2
    double u1[3] = {0, 0, 0}, u2[3] = {0, 0, 0};
    double l1 = -1.0, l2 = -1.0;
    for (int i = 0; i < 3; ++i) {
      if (!d1[i]) {
        copy3(u1, q1.m[i]);
        l1 = l1v[i];
      }
      if (!d2[i]) {
        copy3(u2, q2.m[i]);
        l2 = l2v[i];
      }
    }
    double u[3];
    cross3(u1, u2, u);
    if (dot<3>(u, u) < kMetricEpsilon) return scaled_outer(3, u1, math::max2(l1, l2));
    normalize3(u);
    Mat p;
    cross3(u1, u, p.m[0]);
    cross3(u2, u, p.m[1]);
    copy3(p.m[2], u);
    double l[3] = {vmv<3>(m2, p.m[0]), vmv<3>(m1, p.m[1]), 0.0};
    return congruence_diag<3>(p, l);
  }
  if (n1 == 1 && n2 == 2) {
    // case 3, with the roles of the two metrics swapped relative to Barral
    double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, w1[3] = {0, 0, 0};
    double u2[3] = {0, 0, 0}, v2[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
    bool found_u1 = false, found_v2 = false;
    for (int i = 0; i < 3; ++i) {
      if (d1[i]) {
        copy3(w1, q1.m[i]);
      } else if (found_u1) {
        copy3(v1, q1.m[i]);
      } else {
        copy3(u1, q1.m[i]);
        found_u1 = true;
      }
      if (!d2[i]) {
        copy3(u2, q2.m[i]);
      } else if (found_v2) {
        copy3(w2, q2.m[i]);
      } else {
        copy3(v2, q2.m[i]);
        found_v2 = true;
      }
    }
    if (dot<3>(u2, w1) < kMetricEpsilon) return intersect_in_plane(m1, m2, u1, v1);
    Mat p;
    copy3(p.m[0], v2);
    copy3(p.m[1], w2);
    copy3(p.m[2], w1);
    double l[3] = {vmv<3>(m1, p.m[0]), vmv<3>(m1, p.m[1]), vmv<3>(m2, p.m[2])};
    return congruence_diag<3>(p, l);
  }
  if (n1 == 1 && n2 == 1) {
    // case 4
    double u1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, w1[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
    bool found_u1 = false;
    for (int i = 0; i < 3; ++i) {
      if (d1[i]) {
        copy3(w1, q1.m[i]);
      } else if (found_u1) {
        copy3(v1, q1.m[i]);
      } else {
        copy3(u1, q1.m[i]);
        found_u1 = true;
      }
      if (d2[i]) copy3(w2, q2.m[i]);
    }
    double w[3];
    cross3(w1, w2, w);
    if (dot<3>(w, w) < kMetricEpsilon) return intersect_in_plane(m1, m2, u1, v1);
    normalize3(w);
    Mat p;
    copy3(p.m[0], w1);
    copy3(p.m[1], w2);
    copy3(p.m[2], w);
    double l[3] = {vmv<3>(m2, p.m[0]), vmv<3>(m1, p.m[1]),
        math::max2(vmv<3>(m1, p.m[2]), vmv<3>(m2, p.m[2]))};
    return congruence_diag<3>(p, l);
  }
  return m1;
}

template <int N>
OSH_HD Mat intersect_metrics(Mat const& m1_in, Mat const& m2_in, int* bad) {
  Mat m1 = m1_in, m2 = m2_in, q1, q2;
  double l1[3] = {0, 0, 0}, l2[3] = {0, 0, 0};
  decompose_eigen<N>(m1, q1, l1, bad);
  decompose_eigen<N>(m2, q2, l2, bad);
  int d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0};
  int n1 = 0, n2 = 0;
  for (int i = 0; i < N; ++i) {
    d1[i] = l1[i] < kMetricEpsilon;
    d2[i] = l2[i] < kMetricEpsilon;
    n1 += d1[i];
    n2 += d2[i];
  }
  if (n1 > n2) {
    Mat t = m1;
    m1 = m2;
    m2 = t;
    t = q1;
    q1 = q2;
    q2 = t;
    for (int i = 0; i < 3; ++i) {
      double x = l1[i];
      l1[i] = l2[i];
      l2[i] = x;
      int y = d1[i];
      d1[i] = d2[i];
      d2[i] = y;
    }
    int x = n1;
    n1 = n2;
    n2 = x;
  }
  if (n1 == N) {
    Mat z;
    mat_zero(z);
    return z;
  }
  if (n1 == 0) {
    double ls[3] = {0, 0, 0}, lis[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i) {
      ls[i] = math::sqrt_(l1[i]);
      lis[i] = 1.0 / ls[i];
    }
    Mat s = compose_ortho<N>(q1, ls);
    Mat is = compose_ortho<N>(q1, lis);
    Mat m2bar = matmul_rect<N, N, N>(matmul_rect<N, N, N>(transpose_rect<N, N>(is), m2), is);
    Mat p;
    double lb[3] = {0, 0, 0};
    decompose_eigen<N>(m2bar, p, lb, bad);
    for (int i = 0; i < N; ++i) lb[i] = math::max2(lb[i], 1.0);
    Mat mib = compose_ortho<N>(p, lb);
    return matmul_rect<N, N, N>(matmul_rect<N, N, N>(transpose_rect<N, N>(s), mib), s);
  }
  if constexpr (N == 3) {
    return intersect_degenerate3(m1, q1, l1, d1, n1, m2, q2, l2, d2, n2);
  } else if constexpr (N == 2) {
    return intersect_degenerate2(m1, q1, l1, d1, m2, q2, l2, d2);
  } else {
    return m1;
  }
}

OSH_HD Mat intersect_in_plane(Mat const& m1, Mat const& m2, double const* u1,
    double const* v1) {
  Mat p;
  mat_zero(p);
  for (int i = 0; i < 3; ++i) {
    p.m[0][i] = u1[i];
    p.m[1][i] = v1[i];
  }
  Mat pt = transpose_rect<3, 2>(p);
  Mat m1b = matmul_rect<2, 3, 2>(pt, matmul_rect<3, 3, 2>(m1, p));
  Mat m2b = matmul_rect<2, 3, 2>(pt, matmul_rect<3, 3, 2>(m2, p));
  Mat mint = intersect_metrics<2>(m1b, m2b);
  return matmul_rect<3, 2, 3>(p, matmul_rect<2, 2, 3>(mint, pt));
}

// Householder least-squares / square solve of an m x n system (n <= m <= 6),
// src/Omega_h_qr.hpp:18-149; a[col][row]
OSH_HD void solve_using_qr(int m, int n, double (*a)[6], double const* b_in, double* x) {
  double v[6][6];
  for (int k = 0; k < n; ++k) {
    double norm_x = 0.0;
    for (int i = k; i < m; ++i) norm_x += a[k][i] * a[k][i];
    norm_x = math::sqrt_(norm_x);
    for (int i = k; i < m; ++i) v[k][i] = a[k][i];
    v[k][k] += ((a[k][k] < 0.0) ? -1.0 : 1.0) * norm_x;
    double norm_v = 0.0;
    for (int i = k; i < m; ++i) norm_v += v[k][i] * v[k][i];
    norm_v = math::sqrt_(norm_v);
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

// element_implied_metric, src/Omega_h_shape.hpp:240-273: the metric in
// which every edge of the element has unit length, in symmetric storage
template <int D>
OSH_HD void element_implied_metric(double const (*p)[3], double* out) {
  double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double ev[6][3];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b[i][j] = p[i + 1][j] - p[0][j];
  if constexpr (D == 3) {
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
    Mat a;
    mat_zero(a);
    for (int i = 0; i < 3; ++i) {
      a.m[0][i] = ev[i][0] * ev[i][0];
      a.m[1][i] = ev[i][1] * ev[i][1];
      a.m[2][i] = 2 * ev[i][0] * ev[i][1];
    }
    Mat inv = invert<3>(a);
    double rhs[3] = {1.0, 1.0, 1.0};
    matvec<3>(inv, rhs, out);
  }
}

}  // namespace osh
