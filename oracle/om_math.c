/* Oracle (test infrastructure only): deterministic double-precision
 * transcendentals used by the CPU restatement of Omega_h's adapt() path.
 *
 * The reference calls std::log, std::exp, std::cbrt, std::acos and std::cos
 * (glibc libm) on its decision paths:
 *   anisotropic_edge_length      src/Omega_h_shape.hpp:112-117   (log)
 *   mean_ratio / metric_size     src/Omega_h_quality.hpp:8-12,
 *                                src/Omega_h_scalar.hpp:166-228  (cbrt)
 *   find_polynomial_roots        src/Omega_h_eigen.hpp:21-85     (cbrt, acos, cos)
 *   log_spd_old / exp_spd_old    src/Omega_h_eigen.hpp:473-488   (log, exp)
 * glibc's results cannot be reproduced on the GPU, so the oracle and the HIP
 * kernels both use the fixed operation sequences below (the published
 * fdlibm methods: Sun Microsystems fdlibm 5.3 e_log.c, e_exp.c, s_cbrt.c,
 * e_acos.c, k_cos.c/k_sin.c, with their minimax coefficients).  Compiled
 * with -ffp-contract=off, every operation is a correctly rounded IEEE
 * double operation, so the same sequence gives the same bits on the host
 * and on gfx950.  Accuracy against glibc is checked in tests (< 1 ulp).
 */
#include "om_internal.h"

#include <stdint.h>
#include <string.h>

static inline uint64_t bits_of(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
static inline double from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}
static inline int32_t hi_word(double x) { return (int32_t)(bits_of(x) >> 32); }
static inline uint32_t lo_word(double x) { return (uint32_t)bits_of(x); }
static inline double make_words(int32_t hi, uint32_t lo) {
  return from_bits(((uint64_t)(uint32_t)hi << 32) | (uint64_t)lo);
}

/* ------------------------------------------------------------------ log */
static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;
static const double LG1 = 6.666666666666735130e-01;
static const double LG2 = 3.999999999940941908e-01;
static const double LG3 = 2.857142874366239149e-01;
static const double LG4 = 2.222219843214978396e-01;
static const double LG5 = 1.818357216161805012e-01;
static const double LG6 = 1.531383769920937332e-01;
static const double LG7 = 1.479819860511658591e-01;
static const double TWO54 = 1.80143985094819840000e+16;

double om_log(double x) {
  int32_t hx = hi_word(x);
  uint32_t lx = lo_word(x);
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -TWO54 / 0.0;
    if (hx < 0) return (x - x) / 0.0;
    k -= 54;
    x *= TWO54;
    hx = hi_word(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  /* choose m in [sqrt(2)/2, sqrt(2)) */
  int32_t i = (hx + 0x95f64) & 0x100000;
  x = make_words(hx | (i ^ 0x3ff00000), lo_word(x));
  k += (i >> 20);
  double f = x - 1.0;
  double dk = (double)k;
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * (LG2 + w * (LG4 + w * LG6));
  double t2 = z * (LG1 + w * (LG3 + w * (LG5 + w * LG7)));
  double r = t2 + t1;
  double hfsq = 0.5 * f * f;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + r) + dk * LN2_LO)) - f);
}

/* ------------------------------------------------------------------ exp */
static const double EXP_P1 = 1.66666666666666019037e-01;
static const double EXP_P2 = -2.77777777770155933842e-03;
static const double EXP_P3 = 6.61375632143793436117e-05;
static const double EXP_P4 = -1.65339022054652515390e-06;
static const double EXP_P5 = 4.13813679705723846039e-08;
static const double INV_LN2 = 1.44269504088896338700e+00;
static const double EXP_OVER = 7.09782712893383973096e+02;
static const double EXP_UNDER = -7.45133219101941108420e+02;

double om_exp(double x) {
  if (x != x) return x + x;
  if (x > EXP_OVER) return 1.0e300 * 1.0e300;
  if (x < EXP_UNDER) return 1.0e-300 * 1.0e-300;
  double ax = x < 0.0 ? -x : x;
  int32_t k = 0;
  double hi = 0.0, lo = 0.0;
  if (ax > 0.5 * 6.93147180559945286227e-01) {
    double kd = x * INV_LN2 + (x < 0.0 ? -0.5 : 0.5);
    k = (int32_t)kd;
    double fk = (double)k;
    hi = x - fk * LN2_HI;
    lo = fk * LN2_LO;
    x = hi - lo;
  } else if (ax < 3.7252902984e-09) {
    return 1.0 + x;
  }
  double t = x * x;
  double c = x - t * (EXP_P1 + t * (EXP_P2 + t * (EXP_P3 + t * (EXP_P4 + t * EXP_P5))));
  if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
  double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
  if (k >= -1021) {
    return from_bits(bits_of(y) + ((uint64_t)(int64_t)k << 52));
  }
  /* result is subnormal */
  return from_bits(bits_of(y) + ((uint64_t)(int64_t)(k + 1000) << 52)) *
         9.33263618503218878990e-302; /* 2^-1000 */
}

/* ----------------------------------------------------------------- cbrt */
static const uint32_t CBRT_B1 = 715094163u;
static const uint32_t CBRT_B2 = 696219795u;
static const double CBRT_C = 5.42857142857142815906e-01;  /* 19/35 */
static const double CBRT_D = -7.05306122448979611050e-01; /* -864/1225 */
static const double CBRT_E = 1.41428571428571436819e+00;  /* 99/70 */
static const double CBRT_F = 1.60714285714285720630e+00;  /* 45/28 */
static const double CBRT_G = 3.57142857142857150787e-01;  /* 5/14 */

double om_cbrt(double x) {
  int32_t hx = hi_word(x);
  uint32_t lx = lo_word(x);
  uint32_t sign = (uint32_t)hx & 0x80000000u;
  hx &= 0x7fffffff;
  if (hx >= 0x7ff00000) return x + x;
  if ((hx | (int32_t)lx) == 0) return x;
  x = make_words(hx, lx); /* |x| */
  double t;
  if (hx < 0x00100000) {
    t = make_words(0x43500000, 0) * x; /* 2^54 * x */
    t = make_words((int32_t)((uint32_t)hi_word(t) / 3u + CBRT_B2), 0);
  } else {
    t = make_words((int32_t)((uint32_t)hx / 3u + CBRT_B1), 0);
  }
  double r = t * t / x;
  double s = CBRT_C + r * t;
  t *= CBRT_G + CBRT_F / (s + CBRT_E + CBRT_D / s);
  t = make_words(hi_word(t) + 1, 0);
  s = t * t;
  r = x / s;
  double w = t + t;
  r = (r - t) / (w + r);
  t = t + t * r;
  return sign ? -t : t;
}

/* ----------------------------------------------------------------- acos */
static const double PIO2_HI = 1.57079632679489655800e+00;
static const double PIO2_LO = 6.12323399573676603587e-17;
static const double PI_D = 3.14159265358979311600e+00;
static const double PS0 = 1.66666666666666657415e-01;
static const double PS1 = -3.25565818622400915405e-01;
static const double PS2 = 2.01212532134862925881e-01;
static const double PS3 = -4.00555345006794114027e-02;
static const double PS4 = 7.91534994289814532176e-04;
static const double PS5 = 3.47933107596021167570e-05;
static const double QS1 = -2.40339491173441421878e+00;
static const double QS2 = 2.02094576023350569471e+00;
static const double QS3 = -6.88283971605453293030e-01;
static const double QS4 = 7.70381505559019352791e-02;

static inline double acos_ratio(double z) {
  double p = z * (PS0 + z * (PS1 + z * (PS2 + z * (PS3 + z * (PS4 + z * PS5)))));
  double q = 1.0 + z * (QS1 + z * (QS2 + z * (QS3 + z * QS4)));
  return p / q;
}

double om_acos(double x) {
  int32_t hx = hi_word(x);
  int32_t ix = hx & 0x7fffffff;
  if (ix >= 0x3ff00000) {
    if (((ix - 0x3ff00000) | (int32_t)lo_word(x)) == 0) {
      if (hx > 0) return 0.0;
      return PI_D + 2.0 * PIO2_LO;
    }
    return (x - x) / (x - x);
  }
  if (ix < 0x3fe00000) {
    if (ix <= 0x3c600000) return PIO2_HI + PIO2_LO;
    double z = x * x;
    double r = acos_ratio(z);
    return PIO2_HI - (x - (PIO2_LO - x * r));
  }
  if (hx < 0) {
    double z = (1.0 + x) * 0.5;
    double s = om_sqrt(z);
    double r = acos_ratio(z);
    double w = r * s - PIO2_LO;
    return PI_D - 2.0 * (s + w);
  }
  double z = (1.0 - x) * 0.5;
  double s = om_sqrt(z);
  double df = make_words(hi_word(s), 0);
  double c = (z - df * df) / (s + df);
  double r = acos_ratio(z);
  double w = r * s + c;
  return 2.0 * (df + w);
}

/* ------------------------------------------------------------------ cos */
static const double KC1 = 4.16666666666666019037e-02;
static const double KC2 = -1.38888888888741095749e-03;
static const double KC3 = 2.48015872894767294178e-05;
static const double KC4 = -2.75573143513906633035e-07;
static const double KC5 = 2.08757232129817482790e-09;
static const double KC6 = -1.13596475577881948265e-11;
static const double KS1 = -1.66666666666666324348e-01;
static const double KS2 = 8.33333333332248946124e-03;
static const double KS3 = -1.98412698298579493134e-04;
static const double KS4 = 2.75573137070700676789e-06;
static const double KS5 = -2.50507602534068634195e-08;
static const double KS6 = 1.58969099521155010221e-10;
static const double INV_PIO2 = 6.36619772367581382433e-01;
static const double PIO2_1 = 1.57079632673412561417e+00;
static const double PIO2_1T = 6.07710050650619224932e-11;

static double kernel_cos(double x, double y) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return 1.0;
  }
  double z = x * x;
  double r = z * (KC1 + z * (KC2 + z * (KC3 + z * (KC4 + z * (KC5 + z * KC6)))));
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

static double kernel_sin(double x, double y) {
  double z = x * x;
  double v = z * x;
  double r = KS2 + z * (KS3 + z * (KS4 + z * (KS5 + z * KS6)));
  return x - ((z * (0.5 * y - v * r) - y) - v * KS1);
}

/* cos for |x| <= 8 (the cubic solver only needs |x| <= 4*pi/3). */
double om_cos(double x) {
  int32_t ix = hi_word(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  double fn = (double)(int32_t)(x * INV_PIO2 + (x < 0.0 ? -0.5 : 0.5));
  int32_t n = (int32_t)fn;
  double r = x - fn * PIO2_1;
  double w = fn * PIO2_1T;
  double y0 = r - w;
  double y1 = (r - y0) - w;
  switch (n & 3) {
    case 0: return kernel_cos(y0, y1);
    case 1: return -kernel_sin(y0, y1);
    case 2: return -kernel_cos(y0, y1);
    default: return kernel_sin(y0, y1);
  }
}

/* sqrt is correctly rounded in IEEE-754 on both sides. */
double om_sqrt(double x) { return __builtin_sqrt(x); }
