// Scalar transcendentals, small dense algebra and metric geometry for the
// adapt() kernels, in double precision with a fixed operation order.
//
// Every function evaluates in the order of the reference's C++ templates
// (left-to-right sums, column-major matrices m[col][row],
// src/Omega_h_matrix.hpp:14), and the library is built with
// -ffp-contract=off, so each operation is one correctly rounded IEEE op.
// The transcendentals use the published fdlibm methods (Sun Microsystems
// fdlibm 5.3: e_log.c, e_exp.c, s_cbrt.c, e_acos.c, k_cos.c, k_sin.c) with
// their minimax coefficients, so the host and gfx950 agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define OSH_HD __host__ __device__ __forceinline__

namespace osh {
namespace math {

OSH_HD std::uint64_t bits(double x) { return __builtin_bit_cast(std::uint64_t, x); }
OSH_HD double from_bits(std::uint64_t u) { return __builtin_bit_cast(double, u); }
OSH_HD std::int32_t hi_word(double x) { return std::int32_t(bits(x) >> 32); }
OSH_HD std::uint32_t lo_word(double x) { return std::uint32_t(bits(x)); }
OSH_HD double make_words(std::int32_t hi, std::uint32_t lo) {
  return from_bits((std::uint64_t(std::uint32_t(hi)) << 32) | std::uint64_t(lo));
}
OSH_HD double sqrt_(double x) { return __builtin_sqrt(x); }
OSH_HD double fabs_(double x) { return __builtin_fabs(x); }
OSH_HD double copysign_(double x, double y) { return __builtin_copysign(x, y); }
OSH_HD double min2(double a, double b) { return (b < a) ? b : a; }
OSH_HD double max2(double a, double b) { return (a < b) ? b : a; }

// natural logarithm (fdlibm e_log.c method)
OSH_HD double log_(double x) {
  constexpr double ln2_hi = 6.93147180369123816490e-01;
  constexpr double ln2_lo = 1.90821492927058770002e-10;
  constexpr double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                   Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                   Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                   Lg7 = 1.479819860511658591e-01;
  std::int32_t hx = hi_word(x);
  std::uint32_t lx = lo_word(x);
  std::int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | std::int32_t(lx)) == 0) return -__builtin_inf();
    if (hx < 0) return __builtin_nan("");
    k -= 54;
    x *= 1.80143985094819840000e+16;
    hx = hi_word(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  std::int32_t i = (hx + 0x95f64) & 0x100000;
  x = make_words(hx | (i ^ 0x3ff00000), lo_word(x));
  k += (i >> 20);
  double f = x - 1.0;
  double dk = double(k);
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  double r = t2 + t1;
  double hfsq = 0.5 * f * f;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + r) + dk * ln2_lo)) - f);
}

// exponential (fdlibm e_exp.c method)
OSH_HD double exp_(double x) {
  constexpr double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                   P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                   P5 = 4.13813679705723846039e-08;
  constexpr double ln2_hi = 6.93147180369123816490e-01;
  constexpr double ln2_lo = 1.90821492927058770002e-10;
  constexpr double inv_ln2 = 1.44269504088896338700e+00;
  if (x != x) return x + x;
  if (x > 7.09782712893383973096e+02) return __builtin_inf();
  if (x < -7.45133219101941108420e+02) return 0.0;
  double ax = x < 0.0 ? -x : x;
  std::int32_t k = 0;
  double hi = 0.0, lo = 0.0;
  if (ax > 0.5 * 6.93147180559945286227e-01) {
    double kd = x * inv_ln2 + (x < 0.0 ? -0.5 : 0.5);
    k = std::int32_t(kd);
    double fk = double(k);
    hi = x - fk * ln2_hi;
    lo = fk * ln2_lo;
    x = hi - lo;
  } else if (ax < 3.7252902984e-09) {
    return 1.0 + x;
  }
  double t = x * x;
  double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
  if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
  double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
  if (k >= -1021) return from_bits(bits(y) + (std::uint64_t(std::int64_t(k)) << 52));
  return from_bits(bits(y) + (std::uint64_t(std::int64_t(k + 1000)) << 52)) *
         9.33263618503218878990e-302;
}

// cube root (fdlibm s_cbrt.c method)
OSH_HD double cbrt_(double x) {
  constexpr std::uint32_t B1 = 715094163u, B2 = 696219795u;
  constexpr double C = 5.42857142857142815906e-01, D = -7.05306122448979611050e-01,
                   E = 1.41428571428571436819e+00, F = 1.60714285714285720630e+00,
                   G = 3.57142857142857150787e-01;
  std::int32_t hx = hi_word(x);
  std::uint32_t lx = lo_word(x);
  std::uint32_t sign = std::uint32_t(hx) & 0x80000000u;
  hx &= 0x7fffffff;
  if (hx >= 0x7ff00000) return x + x;
  if ((hx | std::int32_t(lx)) == 0) return x;
  x = make_words(hx, lx);
  double t;
  if (hx < 0x00100000) {
    t = make_words(0x43500000, 0) * x;
    t = make_words(std::int32_t(std::uint32_t(hi_word(t)) / 3u + B2), 0);
  } else {
    t = make_words(std::int32_t(std::uint32_t(hx) / 3u + B1), 0);
  }
  double r = t * t / x;
  double s = C + r * t;
  t *= G + F / (s + E + D / s);
  t = make_words(hi_word(t) + 1, 0);
  s = t * t;
  r = x / s;
  double w = t + t;
  r = (r - t) / (w + r);
  t = t + t * r;
  return sign ? -t : t;
}

OSH_HD double acos_ratio(double z) {
  constexpr double pS0 = 1.66666666666666657415e-01, pS1 = -3.25565818622400915405e-01,
                   pS2 = 2.01212532134862925881e-01, pS3 = -4.00555345006794114027e-02,
                   pS4 = 7.91534994289814532176e-04, pS5 = 3.47933107596021167570e-05;
  constexpr double qS1 = -2.40339491173441421878e+00, qS2 = 2.02094576023350569471e+00,
                   qS3 = -6.88283971605453293030e-01, qS4 = 7.70381505559019352791e-02;
  double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
  double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
  return p / q;
}

// arc cosine (fdlibm e_acos.c method)
OSH_HD double acos_(double x) {
  constexpr double pio2_hi = 1.57079632679489655800e+00;
  constexpr double pio2_lo = 6.12323399573676603587e-17;
  constexpr double pi = 3.14159265358979311600e+00;
  std::int32_t hx = hi_word(x);
  std::int32_t ix = hx & 0x7fffffff;
  if (ix >= 0x3ff00000) {
    if (((ix - 0x3ff00000) | std::int32_t(lo_word(x))) == 0) {
      if (hx > 0) return 0.0;
      return pi + 2.0 * pio2_lo;
    }
    return __builtin_nan("");
  }
  if (ix < 0x3fe00000) {
    if (ix <= 0x3c600000) return pio2_hi + pio2_lo;
    double z = x * x;
    double r = acos_ratio(z);
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (hx < 0) {
    double z = (1.0 + x) * 0.5;
    double s = sqrt_(z);
    double r = acos_ratio(z);
    double w = r * s - pio2_lo;
    return pi - 2.0 * (s + w);
  }
  double z = (1.0 - x) * 0.5;
  double s = sqrt_(z);
  double df = make_words(hi_word(s), 0);
  double c = (z - df * df) / (s + df);
  double r = acos_ratio(z);
  double w = r * s + c;
  return 2.0 * (df + w);
}

OSH_HD double kernel_cos(double x, double y) {
  constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                   C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                   C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  std::int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if (int(x) == 0) return 1.0;
  }
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  double qx;
  if (ix > 0x3fe90000) {
    qx = 0.28125;
  } else {
    qx = make_words(ix - 0x00200000, 0);
  }
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

OSH_HD double kernel_sin(double x, double y) {
  constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                   S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                   S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// cosine for |x| <= 8 (fdlibm k_cos.c/k_sin.c with a two-part pi/2 reduction)
OSH_HD double cos_(double x) {
  constexpr double inv_pio2 = 6.36619772367581382433e-01;
  constexpr double pio2_1 = 1.57079632673412561417e+00;
  constexpr double pio2_1t = 6.07710050650619224932e-11;
  std::int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  double fn = double(std::int32_t(x * inv_pio2 + (x < 0.0 ? -0.5 : 0.5)));
  std::int32_t n = std::int32_t(fn);
  double r = x - fn * pio2_1;
  double w = fn * pio2_1t;
  double y0 = r - w;
  double y1 = (r - y0) - w;
  switch (n & 3) {
    case 0: return kernel_cos(y0, y1);
    case 1: return -kernel_sin(y0, y1);
    case 2: return -kernel_cos(y0, y1);
    default: return kernel_sin(y0, y1);
  }
}

}  // namespace math

// ---------------------------------------------------------------- algebra
struct Mat {
  double m[3][3];  // m[col][row]
};

OSH_HD void mat_zero(Mat& a) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) a.m[j][i] = 0.0;
}

// c = a * v, src/Omega_h_matrix.hpp:89-95
template <int N>
OSH_HD void matvec(Mat const& a, double const* v, double* c) {
  double t[3];
  for (int i = 0; i < N; ++i) t[i] = a.m[0][i] * v[0];
  for (int j = 1; j < N; ++j)
    for (int i = 0; i < N; ++i) t[i] = t[i] + a.m[j][i] * v[j];
  for (int i = 0; i < N; ++i) c[i] = t[i];
}

template <int N>
OSH_HD Mat matmul(Mat const& a, Mat const& b) {
  Mat t;
  mat_zero(t);
  for (int j = 0; j < N; ++j) matvec<N>(a, b.m[j], t.m[j]);
  return t;
}

template <int N>
OSH_HD Mat transpose(Mat const& a) {
  Mat t;
  mat_zero(t);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) t.m[i][j] = a.m[j][i];
  return t;
}

template <int N>
OSH_HD double dot(double const* a, double const* b) {
  double out = a[0] * b[0];
  for (int i = 1; i < N; ++i) out = out + (a[i] * b[i]);
  return out;
}

template <int N>
OSH_HD double vnorm(double const* v) { return math::sqrt_(dot<N>(v, v)); }

OSH_HD void cross3(double const* a, double const* b, double* c) {
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  c[0] = t0;
  c[1] = t1;
  c[2] = t2;
}

template <int N>
OSH_HD double det(Mat const& m) {
  if (N == 1) return m.m[0][0];
  if (N == 2) {
    double a = m.m[0][0], b = m.m[1][0], c = m.m[0][1], d = m.m[1][1];
    return a * d - b * c;
  }
  double a = m.m[0][0], b = m.m[1][0], c = m.m[2][0];
  double d = m.m[0][1], e = m.m[1][1], f = m.m[2][1];
  double g = m.m[0][2], h = m.m[1][2], i = m.m[2][2];
  return (a * e * i) + (b * f * g) + (c * d * h) - (c * e * g) - (b * d * i) - (a * f * h);
}

template <int N>
OSH_HD double trace(Mat const& a) {
  double t = a.m[0][0];
  for (int i = 1; i < N; ++i) t += a.m[i][i];
  return t;
}

template <int M>
OSH_HD int symm_ncomps() { return ((M + 1) * M) / 2; }

// vector2symm, src/Omega_h_matrix.hpp:385-431
template <int M>
OSH_HD Mat get_symm(double const* v) {
  Mat m;
  mat_zero(m);
  if (M == 1) {
    m.m[0][0] = v[0];
  } else if (M == 2) {
    m.m[0][0] = v[0];
    m.m[1][1] = v[1];
    m.m[1][0] = v[2];
    m.m[0][1] = v[2];
  } else {
    m.m[0][0] = v[0];
    m.m[1][1] = v[1];
    m.m[2][2] = v[2];
    m.m[1][0] = v[3];
    m.m[2][1] = v[4];
    m.m[2][0] = v[5];
    m.m[0][1] = v[3];
    m.m[1][2] = v[4];
    m.m[0][2] = v[5];
  }
  return m;
}

template <int M>
OSH_HD void set_symm(Mat const& m, double* v) {
  if (M == 1) {
    v[0] = m.m[0][0];
  } else if (M == 2) {
    v[0] = m.m[0][0];
    v[1] = m.m[1][1];
    v[2] = m.m[1][0];
  } else {
    v[0] = m.m[0][0];
    v[1] = m.m[1][1];
    v[2] = m.m[2][2];
    v[3] = m.m[1][0];
    v[4] = m.m[2][1];
    v[5] = m.m[2][0];
  }
}

// ------------------------------------------------------------ eigensolver
// cubic x^3 + a2 x^2 + a1 x + a0, src/Omega_h_eigen.hpp:21-85
OSH_HD int roots3(double const* coeffs, double eps, double* roots, int* mults) {
  using namespace math;
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
    double S = cbrt_(R + sqrt_(D));
    double T = cbrt_(R - sqrt_(D));
    double B = S + T;
    double z_1 = shift + B;
    double z_23_real = shift - (1. / 2.) * B;
    roots[0] = z_1;
    roots[1] = roots[2] = z_23_real;
  } else {
    constexpr double pi = 3.14159265358979323;
    double cos_theta = R / sqrt_(-(Q * (Q * Q)));
    double theta = acos_(min2(max2(cos_theta, -1.0), 1.0));
    double radius = 2. * sqrt_(-Q);
    double z_1 = radius * cos_((theta) / 3.) + shift;
    double z_2 = radius * cos_((theta + 2. * pi) / 3.) + shift;
    double z_3 = radius * cos_((theta - 2. * pi) / 3.) + shift;
    roots[0] = z_1;
    roots[1] = z_2;
    roots[2] = z_3;
  }
  mults[0] = mults[1] = mults[2] = 1;
  if (fabs_(roots[0] - roots[1]) < eps) {
    double t = roots[0];
    roots[0] = roots[2];
    roots[2] = t;
  } else if (fabs_(roots[0] - roots[2]) < eps) {
    double t = roots[0];
    roots[0] = roots[1];
    roots[1] = t;
  } else if (fabs_(roots[1] - roots[2]) < eps) {
  } else {
    return 3;
  }
  roots[1] = (roots[1] + roots[2]) / 2;
  mults[1] = 2;
  if (fabs_(roots[0] - roots[1]) < eps) {
    roots[0] = (1. / 3.) * roots[0] + (2. / 3.) * roots[1];
    mults[0] = 3;
    return 1;
  }
  return 2;
}

// quadratic x^2 + a x + b, src/Omega_h_eigen.hpp:88-112
OSH_HD int roots2(double const* coeffs, double eps, double* roots, int* mults) {
  using namespace math;
  double a = coeffs[1];
  double b = coeffs[0];
  roots[0] = roots[1] = 0.0;
  mults[0] = mults[1] = 0;
  double disc = a * a - 4. * b;
  if (fabs_(disc) < eps) {
    mults[0] = 2;
    roots[0] = -a / 2.;
    roots[1] = roots[0];
    mults[1] = mults[0];
    return 1;
  }
  if (disc > 0.0) {
    mults[0] = 1;
    mults[1] = 1;
    roots[0] = (-a + sqrt_(disc)) / 2.;
    roots[1] = (-a - sqrt_(disc)) / 2.;
    return 2;
  }
  return 0;
}

// false when no row pair spans a plane (the reference's check fails,
// src/Omega_h_eigen.hpp:191); out is then left as it was
OSH_HD bool single_eigenvector3(Mat const& m, double l, double* out) {
  Mat s0 = m;
  for (int i = 0; i < 3; ++i) s0.m[i][i] -= l;
  Mat s = transpose<3>(s0);
  double v[3], c[3];
  cross3(s.m[0], s.m[1], v);
  double v_norm = vnorm<3>(v);
  cross3(s.m[1], s.m[2], c);
  double c_norm = vnorm<3>(c);
  if (c_norm > v_norm) {
    v[0] = c[0];
    v[1] = c[1];
    v[2] = c[2];
    v_norm = c_norm;
  }
  cross3(s.m[0], s.m[2], c);
  c_norm = vnorm<3>(c);
  if (c_norm > v_norm) {
    v[0] = c[0];
    v[1] = c[1];
    v[2] = c[2];
    v_norm = c_norm;
  }
  if (!(v_norm > 1e-10)) return false;
  for (int i = 0; i < 3; ++i) out[i] = v[i] / v_norm;
  return true;
}

// false when every row is (numerically) zero (the reference's check fails,
// src/Omega_h_eigen.hpp:211); out is then left as it was
template <int N>
OSH_HD bool row_space_1d(Mat const& a, double* out) {
  Mat ta = transpose<N>(a);
  int best_row = 0;
  double best_norm = vnorm<N>(ta.m[0]);
  for (int i = 1; i < N; ++i) {
    double row_norm = vnorm<N>(ta.m[i]);
    if (row_norm > best_norm) {
      best_row = i;
      best_norm = row_norm;
    }
  }
  if (!(best_norm > 1e-10)) return false;
  for (int i = 0; i < N; ++i) out[i] = ta.m[best_row][i] / best_norm;
  return true;
}

OSH_HD void ortho_basis3(double const* v, Mat& a) {
  a.m[0][0] = v[0];
  a.m[0][1] = v[1];
  a.m[0][2] = v[2];
  double sign = math::copysign_(1.0, v[2]);
  double aa = -1.0 / (sign + v[2]);
  double b = v[0] * v[1] * aa;
  a.m[1][0] = 1.0 + sign * v[0] * v[0] * aa;
  a.m[1][1] = sign * b;
  a.m[1][2] = -sign * v[0];
  a.m[2][0] = b;
  a.m[2][1] = sign + v[1] * v[1] * aa;
  a.m[2][2] = -v[1];
}

template <int N>
OSH_HD void identity(Mat& a) {
  mat_zero(a);
  for (int i = 0; i < N; ++i) a.m[i][i] = 1.0;
}

// get_max_eigenvalues, src/Omega_h_eigen.cpp:7-40: the root of largest
// magnitude of the characteristic polynomial (src/Omega_h_eigen.hpp:127-156,
// tolerance 5e-5, the matrix not normalised); a 1x1 metric is its value
template <int N>
OSH_HD double max_eigenvalue(Mat const& m) {
  if (N == 1) return m.m[0][0];
  double roots[3] = {0.0, 0.0, 0.0};
  int mults[3];
  int n;
  if (N == 3) {
    double coeffs[3];
    double tA = trace<3>(m);
    Mat AA = matmul<3>(m, m);
    coeffs[2] = -tA;
    coeffs[1] = (1. / 2.) * ((tA * tA) - trace<3>(AA));
    coeffs[0] = -det<3>(m);
    n = roots3(coeffs, 5e-5, roots, mults);
  } else {
    double coeffs[2];
    coeffs[1] = -trace<2>(m);
    coeffs[0] = det<2>(m);
    n = roots2(coeffs, 5e-5, roots, mults);
  }
  double out = roots[0];
  double max_mag = math::fabs_(roots[0]);
  for (int i = 1; i < n; ++i) {
    double mag = math::fabs_(roots[i]);
    if (mag > max_mag) {
      max_mag = mag;
      out = roots[i];
    }
  }
  return out;
}

// decompose_eigen with max-norm scaling, src/Omega_h_eigen.hpp:228-319;
// *bad (when given) is set when an eigenvector cannot be formed, where the
// reference's checks abort (src/Omega_h_eigen.hpp:191,211)
template <int N>
OSH_HD void decompose_eigen(Mat const& m0, Mat& q, double* l, int* bad = nullptr) {
  double nm = 0.0;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) nm = math::max2(nm, math::fabs_(m0.m[j][i]));
  if (nm <= 1e-10) {
    identity<N>(q);
    for (int i = 0; i < N; ++i) l[i] = 0.0;
    return;
  }
  Mat m;
  mat_zero(m);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) m.m[j][i] = m0.m[j][i] / nm;
  mat_zero(q);
  if (N == 3) {
    double coeffs[3];
    double tA = trace<3>(m);
    Mat AA = matmul<3>(m, m);
    coeffs[2] = -tA;
    coeffs[1] = (1. / 2.) * ((tA * tA) - trace<3>(AA));
    coeffs[0] = -det<3>(m);
    double roots[3];
    int mults[3];
    int nroots = roots3(coeffs, 5e-5, roots, mults);
    bool ok = true;
    if (nroots == 3) {
      for (int i = 0; i < 3; ++i) {
        ok = single_eigenvector3(m, roots[i], q.m[i]) && ok;
        l[i] = roots[i];
      }
    } else if (nroots == 2 && mults[1] == 2) {
      ok = single_eigenvector3(m, roots[0], q.m[0]);
      l[0] = roots[0];
      Mat s = m;
      for (int i = 0; i < 3; ++i) s.m[i][i] -= roots[1];
      double n[3] = {0.0, 0.0, 0.0};
      ok = row_space_1d<3>(s, n) && ok;
      Mat b;
      ortho_basis3(n, b);
      for (int i = 0; i < 3; ++i) {
        q.m[1][i] = b.m[1][i];
        q.m[2][i] = b.m[2][i];
      }
      l[1] = l[2] = roots[1];
    } else {
      l[0] = l[1] = l[2] = roots[0];
      identity<3>(q);
    }
    if (!ok && bad) *bad = 1;
  } else if (N == 2) {
    double coeffs[2];
    coeffs[1] = -trace<2>(m);
    coeffs[0] = det<2>(m);
    double roots[2];
    int mults[2];
    int nroots = roots2(coeffs, 5e-5, roots, mults);
    if (nroots == 2) {
      for (int i = 0; i < 2; ++i) {
        Mat s = m;
        for (int k = 0; k < 2; ++k) s.m[k][k] -= roots[i];
        double r[2] = {0.0, 0.0};
        if (!row_space_1d<2>(s, r) && bad) *bad = 1;
        q.m[i][0] = -r[1];
        q.m[i][1] = r[0];
        l[i] = roots[i];
      }
    } else {
      l[0] = l[1] = roots[0];
      identity<2>(q);
    }
  } else {
    q.m[0][0] = 1.0;
    l[0] = -(-det<1>(m));
  }
  for (int i = 0; i < N; ++i) l[i] = l[i] * nm;
}

template <int N>
OSH_HD Mat compose_ortho(Mat const& q, double const* l) {
  Mat d;
  mat_zero(d);
  for (int i = 0; i < N; ++i) d.m[i][i] = l[i];
  Mat qd = matmul<N>(q, d);
  Mat qt = transpose<N>(q);
  return matmul<N>(qd, qt);
}

template <int N>
OSH_HD Mat log_spd(Mat const& m, int* bad = nullptr) {
  Mat q;
  double l[3];
  decompose_eigen<N>(m, q, l, bad);
  for (int i = 0; i < N; ++i) l[i] = math::log_(l[i]);
  return compose_ortho<N>(q, l);
}

template <int N>
OSH_HD Mat exp_spd(Mat const& m, int* bad = nullptr) {
  Mat q;
  double l[3];
  decompose_eigen<N>(m, q, l, bad);
  for (int i = 0; i < N; ++i) l[i] = math::exp_(l[i]);
  return compose_ortho<N>(q, l);
}

// log-Euclidean average of two metrics, src/Omega_h_metric.hpp:125-149
template <int M>
OSH_HD void average_metric2(double const* a, double const* b, double* out, int* bad = nullptr) {
  Mat am;
  mat_zero(am);
  double const* ms[2] = {a, b};
  for (int k = 0; k < 2; ++k) {
    Mat lm = log_spd<M>(get_symm<M>(ms[k]), bad);
    for (int j = 0; j < M; ++j)
      for (int i = 0; i < M; ++i) am.m[j][i] = am.m[j][i] + lm.m[j][i];
  }
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < M; ++i) am.m[j][i] = am.m[j][i] / 2.0;
  set_symm<M>(exp_spd<M>(am, bad), out);
}

// ------------------------------------------------------------ geometry
// v . (M v), src/Omega_h_metric.hpp:8-18
template <int S, int M>
OSH_HD double metric_product(Mat const& m, double const* v) {
  double mv[3];
  if (M == 1) {
    for (int i = 0; i < S; ++i) mv[i] = m.m[0][0] * v[i];
  } else {
    matvec<S>(m, v, mv);
  }
  return dot<S>(v, mv);
}

OSH_HD double aniso_length(double l_a, double l_b) {
  if (math::fabs_(l_a - l_b) > 1e-3) return (l_a - l_b) / (math::log_(l_a / l_b));
  return (l_a + l_b) / 2.;
}

// metric_edge_length, src/Omega_h_shape.hpp:112-125
template <int S, int M>
OSH_HD double edge_length(double const* p0, double const* p1, double const* m0,
    double const* m1) {
  double v[3];
  for (int i = 0; i < S; ++i) v[i] = p1[i] - p0[i];
  Mat a = get_symm<M>(m0);
  Mat b = get_symm<M>(m1);
  double l_a = math::sqrt_(metric_product<S, M>(a, v));
  double l_b = math::sqrt_(metric_product<S, M>(b, v));
  return aniso_length(l_a, l_b);
}

// barycentric_from_global (src/Omega_h_shape.hpp:46-63) of point x in the
// simplex p: the basis b[i] = p[i+1] - p[0] inverted by the cofactor formulas
// of src/Omega_h_matrix.hpp:332-350, times x - p[0]
template <int D>
OSH_HD void barycentric(double const* x, double const (*p)[3], double* xi) {
  Mat b, inv;
  mat_zero(b);
  mat_zero(inv);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b.m[i][j] = p[i + 1][j] - p[0][j];
  double dt = det<D>(b);
  if (D == 3) {
    double bc[3][3];
    cross3(b.m[1], b.m[2], bc[0]);
    cross3(b.m[2], b.m[0], bc[1]);
    cross3(b.m[0], b.m[1], bc[2]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) inv.m[j][i] = bc[i][j] / dt;
  } else {
    double a = b.m[0][0], bb = b.m[1][0], c = b.m[0][1], d = b.m[1][1];
    inv.m[0][0] = d / dt;
    inv.m[0][1] = -c / dt;
    inv.m[1][0] = -bb / dt;
    inv.m[1][1] = a / dt;
  }
  double v[3], lam[3];
  for (int i = 0; i < D; ++i) v[i] = x[i] - p[0][i];
  for (int i = 0; i < D; ++i) lam[i] = inv.m[0][i] * v[0];
  for (int j = 1; j < D; ++j)
    for (int i = 0; i < D; ++i) lam[i] = lam[i] + inv.m[j][i] * v[j];
  xi[0] = 1.0;
  for (int i = 1; i < D + 1; ++i) {
    xi[i] = lam[i - 1];
    xi[0] -= lam[i - 1];
  }
}

// maxdet_metric + metric_element_quality, src/Omega_h_metric.hpp:170-181,
// src/Omega_h_quality.hpp:8-37
template <int D, int M>
OSH_HD double element_quality(double const* const* p, double const* const* ms) {
  Mat m = get_symm<M>(ms[0]);
  double maxdet = det<M>(m);
  for (int i = 1; i < D + 1; ++i) {
    Mat mi = get_symm<M>(ms[i]);
    double d = det<M>(mi);
    if (d > maxdet) {
      m = mi;
      maxdet = d;
    }
  }
  double b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b[i][j] = p[i + 1][j] - p[0][j];
  double rs;
  if (D == 3) {
    double c[3];
    cross3(b[0], b[1], c);
    rs = dot<3>(c, b[2]) / 6.0;
  } else {
    rs = (b[0][0] * b[1][1] - b[0][1] * b[1][0]) / 2.0;
  }
  // the chosen metric's determinant is the maximum found above
  double dm = maxdet;
  double f;
  if (D == 3 && M == 3)
    f = math::sqrt_(dm);
  else if (D == 3 && M == 1)
    f = math::sqrt_(dm * (dm * dm));
  else if (D == 2 && M == 2)
    f = math::sqrt_(dm);
  else
    f = dm;
  double s = rs * f;
  if (s < 0) return s;
  double ev[6][3];
  constexpr int nedges = (D == 3) ? 6 : 3;
  if (D == 3) {
    for (int j = 0; j < 3; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
      ev[3][j] = b[2][j];
      ev[4][j] = p[3][j] - p[1][j];
      ev[5][j] = p[3][j] - p[2][j];
    }
  } else {
    for (int j = 0; j < 2; ++j) {
      ev[0][j] = b[0][j];
      ev[1][j] = p[2][j] - p[1][j];
      ev[2][j] = -b[1][j];
    }
  }
  double msl = 0;
  for (int i = 0; i < nedges; ++i) msl += metric_product<D, M>(m, ev[i]);
  msl = msl / nedges;
  constexpr double eq = (D == 3) ? 0.1178511301977579 : 0.4330127018922193;
  double x = s / eq;
  double num = (D == 3) ? math::cbrt_(x * x) : x;
  return num / msl;
}

// ------------------------------------------------------------ FP32 screens
// Single-precision estimates of element_quality and edge_length for the
// mixed-precision form (no counterpart in the reference).  They only screen:
// a cavity is rejected on an estimate that falls short of its threshold by
// more than kScreenBand (relative), and every other cavity is evaluated
// again in FP64, so decisions are those of the FP64 path.  Vertex
// differences are taken in FP64 and rounded once (coordinates near 1 would
// otherwise lose the digits of short edges), the rest runs in FP32 with the
// hardware reciprocal, square root, log2 and exp2.  Their relative error
// stays below 1e-4 for elements and metrics of condition number up to 1e3,
// a tenth of the band.
namespace f32 {

constexpr float kScreenBand = 1e-3f;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// symmetric metric of M = 1 (isotropic), 2 or 3 from Omega_h's layout
template <int M>
__device__ __forceinline__ void load_symm(double const* v, float* out) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
#pragma unroll
  for (int k = 0; k < nc; ++k) out[k] = float(v[k]);
}

template <int M>
__device__ __forceinline__ float symm_det(float const* m) {
  if (M == 1) return m[0];
  if (M == 2) return m[0] * m[1] - m[2] * m[2];
  // (xx, yy, zz, xy, yz, xz)
  return m[0] * (m[1] * m[2] - m[4] * m[4]) - m[3] * (m[3] * m[2] - m[4] * m[5]) +
         m[5] * (m[3] * m[4] - m[1] * m[5]);
}

template <int S, int M>
__device__ __forceinline__ float metric_product(float const* m, float const* v) {
  if (M == 1) {
    float d = v[0] * v[0];
    for (int i = 1; i < S; ++i) d += v[i] * v[i];
    return m[0] * d;
  }
  if (M == 2) return m[0] * v[0] * v[0] + m[1] * v[1] * v[1] + 2.0f * m[2] * v[0] * v[1];
  return m[0] * v[0] * v[0] + m[1] * v[1] * v[1] + m[2] * v[2] * v[2] +
         2.0f * (m[3] * v[0] * v[1] + m[4] * v[1] * v[2] + m[5] * v[0] * v[2]);
}

// Metrics whose components differ by at most kSameMetric of the largest
// component give qualities within a few kSameMetric of each other (the
// mean squared metric length is linear in the metric and the determinant
// enters through a root), so which of them is the max-determinant metric
// does not move the estimate beyond its error budget.
constexpr float kSameMetric = 1e-5f;

template <int M>
__device__ __forceinline__ bool near_same_metric(float const* a, float const* b) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  float scale = 0.0f, diff = 0.0f;
#pragma unroll
  for (int k = 0; k < nc; ++k) {
    float x = a[k] < 0.0f ? -a[k] : a[k];
    float y = a[k] - b[k];
    y = y < 0.0f ? -y : y;
    scale = x > scale ? x : scale;
    diff = y > diff ? y : diff;
  }
  return diff <= kSameMetric * scale;
}

// element_quality in FP32; *unsure is set when the max-determinant vertex
// metric is not clear (another vertex's metric, not nearly the same, with a
// determinant within the band), since the FP64 choice could then differ
template <int D, int M>
__device__ __forceinline__ float element_quality(double const* const* p, double const* const* ms,
    bool* unsure) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  float m[nc], mi[nc];
  load_symm<M>(ms[0], m);
  float maxdet = symm_det<M>(m);
  float second = -3.4e38f;
  for (int i = 1; i < D + 1; ++i) {
    load_symm<M>(ms[i], mi);
    float d = symm_det<M>(mi);
    if (near_same_metric<M>(m, mi)) continue;
    if (d > maxdet) {
      second = maxdet;
      for (int k = 0; k < nc; ++k) m[k] = mi[k];
      maxdet = d;
    } else if (d > second) {
      second = d;
    }
  }
  if (!(maxdet - second > kScreenBand * maxdet)) *unsure = true;
  float b[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b[i][j] = float(p[i + 1][j] - p[0][j]);
  float rs;
  if (D == 3) {
    float c0 = b[0][1] * b[1][2] - b[0][2] * b[1][1];
    float c1 = b[0][2] * b[1][0] - b[0][0] * b[1][2];
    float c2 = b[0][0] * b[1][1] - b[0][1] * b[1][0];
    rs = (c0 * b[2][0] + c1 * b[2][1] + c2 * b[2][2]) * (1.0f / 6.0f);
  } else {
    rs = (b[0][0] * b[1][1] - b[0][1] * b[1][0]) * 0.5f;
  }
  float dm = maxdet;
  float f;
  if (D == 3 && M == 3)
    f = sqrt(dm);
  else if (D == 3 && M == 1)
    f = sqrt(dm * (dm * dm));
  else if (D == 2 && M == 2)
    f = sqrt(dm);
  else
    f = dm;
  float s = rs * f;
  if (s < 0.0f) return s;
  constexpr int nedges = (D == 3) ? 6 : 3;
  float ev[6][3];
  for (int j = 0; j < 3; ++j) {
    ev[0][j] = b[0][j];
    ev[1][j] = b[1][j] - b[0][j];
    ev[2][j] = -b[1][j];
    ev[3][j] = b[2][j];
    ev[4][j] = b[2][j] - b[0][j];
    ev[5][j] = b[2][j] - b[1][j];
  }
  float msl = 0.0f;
  for (int i = 0; i < nedges; ++i) msl += metric_product<D, M>(m, ev[i]);
  msl = msl * (1.0f / nedges);
  constexpr float eq = (D == 3) ? 0.1178511301977579f : 0.4330127018922193f;
  float x = s * (1.0f / eq);
  // cbrt(x * x) = 2^(2/3 log2 x); x = 0 gives 0
  float num = (D == 3) ? exp2((2.0f / 3.0f) * log2(x)) : x;
  return num * rcp(msl);
}

// metric_edge_length in FP32: near-equal end lengths (within 2 %) take the
// arithmetic mean, which is within 3.4e-5 of the logarithmic mean there
template <int S, int M>
__device__ __forceinline__ float edge_length(double const* p0, double const* p1,
    double const* m0, double const* m1) {
  constexpr int nc = (M == 1) ? 1 : (M == 2 ? 3 : 6);
  float v[3] = {0, 0, 0};
  for (int i = 0; i < S; ++i) v[i] = float(p1[i] - p0[i]);
  float a[nc], b[nc];
  load_symm<M>(m0, a);
  load_symm<M>(m1, b);
  float la = sqrt(metric_product<S, M>(a, v));
  float lb = sqrt(metric_product<S, M>(b, v));
  float hi = la > lb ? la : lb;
  float d = la - lb;
  if (!(d > 0.02f * hi) && !(-d > 0.02f * hi)) return 0.5f * (la + lb);
  return d * rcp(log2(la * rcp(lb)) * 0.6931471805599453f);
}

}  // namespace f32

}  // namespace osh
