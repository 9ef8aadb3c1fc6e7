/*
 * orc_math.h -- elementary functions of the CPU oracle.
 *
 * TEST INFRASTRUCTURE: part of the oracle, used only by tests and by the
 * CPU-baseline leg of the benchmark, never by the product library.
 *
 * The transport path needs -log(xi) for every collision distance
 * (src/raykernels/CollisionKernel.C:150) and cos/sin of 2*pi*xi for every
 * azimuth.  libm and the GPU device library do not promise identical bits,
 * so the oracle and the HIP kernels both evaluate these with the same
 * published method, using only IEEE-754 +,-,*,/ (no contraction):
 *
 *   log:  argument reduction x = 2^k (1+f), sqrt(2)/2 <= 1+f < sqrt(2),
 *         s = f/(2+f), log(1+f) = f - (hfsq - s (hfsq + R(s^2))) with the
 *         degree-14 minimax R of Sun Microsystems' fdlibm 5.3 e_log.c (1993).
 *   sin/cos of 2*pi*xi, xi in [0,1): exact quadrant reduction t = 4 xi,
 *         q = nearest integer, r = t - q (exact), a = r*pi/2 in [-pi/4, pi/4],
 *         then fdlibm's __kernel_sin / __kernel_cos polynomials.
 *
 * Accuracy is checked against libm in tests (< 1 ulp for log, < 2 ulp for
 * the trigonometric pair); bit-equality with the device copy is checked on
 * the GPU.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <stdint.h>
#include <string.h>

static inline uint64_t orc_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double orc_from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* Natural logarithm (fdlibm e_log method). */
static inline double orc_log(double x)
{
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double two54 = 1.80143985094819840000e+16;
  const double Lg1 = 6.666666666666735130e-01;
  const double Lg2 = 3.999999999940941908e-01;
  const double Lg3 = 2.857142874366239149e-01;
  const double Lg4 = 2.222219843214978396e-01;
  const double Lg5 = 1.818357216161805012e-01;
  const double Lg6 = 1.531383769920937332e-01;
  const double Lg7 = 1.479819860511658591e-01;

  uint64_t u = orc_bits(x);
  int32_t hx = (int32_t)(u >> 32);
  uint32_t lx = (uint32_t)u;
  int32_t k = 0;
  if (hx < 0x00100000) {                       /* x < 2^-1022 */
    if (((hx & 0x7fffffff) | lx) == 0) return -two54 / 0.0; /* log(+-0) = -inf */
    if (hx < 0) return (x - x) / 0.0;          /* log(-x) = NaN */
    k -= 54;
    x *= two54;                                /* scale up subnormal x */
    u = orc_bits(x);
    hx = (int32_t)(u >> 32);
  }
  if (hx >= 0x7ff00000) return x + x;          /* inf or NaN */
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  int32_t i = (hx + 0x95f64) & 0x100000;
  /* normalize x or x/2 so that 1+f lies in [sqrt(2)/2, sqrt(2)) */
  u = ((uint64_t)(uint32_t)(hx | (i ^ 0x3ff00000)) << 32) | (u & 0xffffffffu);
  x = orc_from_bits(u);
  k += (i >> 20);
  double f = x - 1.0;
  double dk = (double)k;
  if ((0x000fffff & (2 + hx)) < 3) {           /* |f| < 2^-20 */
    if (f == 0.0) {
      if (k == 0) return 0.0;
      return dk * ln2_hi + dk * ln2_lo;
    }
    double R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  double s = f / (2.0 + f);
  double z = s * s;
  i = hx - 0x6147a;
  double w = z * z;
  int32_t j = 0x6b851 - hx;
  double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  double R = t2 + t1;
  if (i > 0) {
    double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

/* fdlibm __kernel_sin(x, 0, 0), |x| <= pi/4 */
static inline double orc_ksin(double x)
{
  const double S1 = -1.66666666666666324348e-01;
  const double S2 = 8.33333333332248946124e-03;
  const double S3 = -1.98412698298579493134e-04;
  const double S4 = 2.75573137070700676789e-06;
  const double S5 = -2.50507602534068634195e-08;
  const double S6 = 1.58969099521155010221e-10;
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  return x + v * (S1 + z * r);
}

/* fdlibm __kernel_cos(x, 0), |x| <= pi/4 */
static inline double orc_kcos(double x)
{
  const double C1 = 4.16666666666666019037e-02;
  const double C2 = -1.38888888888741095749e-03;
  const double C3 = 2.48015872894767294178e-05;
  const double C4 = -2.75573143513906633035e-07;
  const double C5 = 2.08757232129817482790e-09;
  const double C6 = -1.13596475577881948265e-11;
  int32_t ix = (int32_t)(orc_bits(x) >> 32) & 0x7fffffff;
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - z * r); /* |x| < 0.3 */
  double qx;
  if (ix > 0x3fe90000) qx = 0.28125;
  else qx = orc_from_bits((uint64_t)(uint32_t)(ix - 0x00200000) << 32); /* x/4 */
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - z * r);
}

/* sin(2 pi xi) and cos(2 pi xi) for xi in [0, 1). */
static inline void orc_sincos_2pi(double xi, double *s, double *c)
{
  const double pio2 = 1.57079632679489661923;
  double t = 4.0 * xi;                 /* exact */
  double q = (double)(int64_t)(t + 0.5); /* nearest quadrant, t >= 0 */
  double r = t - q;                    /* exact, |r| <= 0.5 */
  double a = r * pio2;
  double sa = orc_ksin(a), ca = orc_kcos(a);
  int iq = ((int)q) & 3;
  switch (iq) {
    case 0: *s = sa; *c = ca; break;
    case 1: *s = ca; *c = -sa; break;
    case 2: *s = -sa; *c = -ca; break;
    default: *s = -ca; *c = sa; break;
  }
}

/* cos(pi/2 * xi) for xi in [0, 1] (Maxwell/Watt sampling). */
static inline double orc_cos_halfpi(double xi)
{
  /* 2 pi * (xi/4): xi/4 is exact */
  double s, c;
  orc_sincos_2pi(0.25 * xi, &s, &c);
  return c;
}

#endif
