/*
 * orc_math.h -- TEST INFRASTRUCTURE ONLY (parity oracle).
 *
 * Serial FP64 elementary functions for the CPU oracle.  The reference kernels
 * call the CUDA device math library (exp, log, sinh, cosh, sincos, rsqrt,
 * fmod; e.g. simulation_kernel.cu:30,176-177, space_charge_kernel.cu:60);
 * that library is closed and not reproduced here.  Instead the oracle and the
 * HIP product share one *published-method specification* and each implements
 * it independently, so that both produce identical bits:
 *
 *   exp    : Cody-Waite reduction x = k*ln2 + r (ln2 split in two parts,
 *            fused multiply-add), Taylor polynomial of degree 13 on
 *            |r| <= ln2/2 (coefficients 1/n! rounded to nearest), scaling by
 *            2^k in two exact power-of-two multiplications.
 *   log    : x = 2^e * m, m in [sqrt(1/2), sqrt(2)], log m = 2 atanh(s),
 *            s = (m-1)/(m+1), odd series in s up to s^23.
 *   sincos : Cody-Waite reduction by pi/2 (three-part pi/2, fma), Taylor
 *            polynomials sin to r^17 and cos to r^18 on |r| <= pi/4.
 *   sinh/cosh : Taylor series (to x^21 / x^20) for |x| < 1, otherwise
 *            (e^|x| -/+ e^-|x|)/2.
 *   fmod   : exact binary long division by repeated subtraction of
 *            power-of-two multiples of the divisor (Sterbenz-exact steps);
 *            equals the IEEE remainder-toward-zero bit for bit.
 *   rsqrt(x) is written 1.0/sqrt(x) everywhere.
 *
 * Only IEEE-exact primitives are used (+, -, *, /, sqrt, fma, floor), and the
 * oracle is compiled with -ffp-contract=off, so no operation is fused behind
 * the source's back.
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint64_t orc_bits(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }
static inline double orc_from_bits(uint64_t u) { double v; memcpy(&v, &u, 8); return v; }

/* 2^k for -1022 <= k <= 1023 */
static inline double orc_pow2i(int k) { return orc_from_bits((uint64_t)(k + 1023) << 52); }

static inline double orc_scale2(double x, int k)
{
  int k1 = k / 2;
  int k2 = k - k1;
  return (x * orc_pow2i(k1)) * orc_pow2i(k2);
}

#define ORC_LN2_HI 0x1.62e42fefa39efp-1
#define ORC_LN2_LO 0x1.abc9e3b39803fp-56
#define ORC_INV_LN2 0x1.71547652b82fep+0
#define ORC_PIO2_1 0x1.921fb54442d18p+0
#define ORC_PIO2_2 0x1.1a62633145c07p-54
#define ORC_PIO2_3 (-0x1.f1976b7ed8fbcp-110)
#define ORC_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define ORC_SQRT2 1.4142135623730951

static inline double orc_exp(double x)
{
  if (x != x) return x;
  if (x > 709.79) return INFINITY;
  if (x < -745.2) return 0.0;
  double kd = floor(x * ORC_INV_LN2 + 0.5);
  double r = fma(-kd, ORC_LN2_HI, x);
  r = fma(-kd, ORC_LN2_LO, r);
  double p = 1.6059043836821613e-10;          /* 1/13! */
  p = fma(p, r, 2.08767569878681e-09);         /* 1/12! */
  p = fma(p, r, 2.505210838544172e-08);        /* 1/11! */
  p = fma(p, r, 2.755731922398589e-07);        /* 1/10! */
  p = fma(p, r, 2.7557319223985893e-06);       /* 1/9!  */
  p = fma(p, r, 2.48015873015873e-05);         /* 1/8!  */
  p = fma(p, r, 0.0001984126984126984);        /* 1/7!  */
  p = fma(p, r, 0.001388888888888889);         /* 1/6!  */
  p = fma(p, r, 0.008333333333333333);         /* 1/5!  */
  p = fma(p, r, 0.041666666666666664);         /* 1/4!  */
  p = fma(p, r, 0.16666666666666666);          /* 1/3!  */
  p = fma(p, r, 0.5);                          /* 1/2!  */
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return orc_scale2(p, (int)kd);
}

static inline double orc_log(double x)
{
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  uint64_t u = orc_bits(x);
  int e = (int)((u >> 52) & 0x7ff);
  if (e == 0) {
    x = x * 0x1p54;
    u = orc_bits(x);
    e = (int)((u >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  double m = orc_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > ORC_SQRT2) { m = m * 0.5; e += 1; }
  double f = m - 1.0;
  double s = f / (m + 1.0);
  double z = s * s;
  double p = 0.043478260869565216;             /* 1/23 */
  p = fma(p, z, 0.047619047619047616);         /* 1/21 */
  p = fma(p, z, 0.05263157894736842);          /* 1/19 */
  p = fma(p, z, 0.058823529411764705);         /* 1/17 */
  p = fma(p, z, 0.06666666666666667);          /* 1/15 */
  p = fma(p, z, 0.07692307692307693);          /* 1/13 */
  p = fma(p, z, 0.09090909090909091);          /* 1/11 */
  p = fma(p, z, 0.1111111111111111);           /* 1/9  */
  p = fma(p, z, 0.14285714285714285);          /* 1/7  */
  p = fma(p, z, 0.2);                          /* 1/5  */
  p = fma(p, z, 0.3333333333333333);           /* 1/3  */
  double two_s = 2.0 * s;
  double lm = fma(two_s * z, p, two_s);
  double ed = (double)e;
  return fma(ed, ORC_LN2_HI, fma(ed, ORC_LN2_LO, lm));
}

static inline void orc_sincos(double x, double* sp, double* cp)
{
  if (!(fabs(x) < INFINITY)) { *sp = x - x; *cp = x - x; return; }
  double kd = floor(x * ORC_TWO_OVER_PI + 0.5);
  double r = fma(-kd, ORC_PIO2_1, x);
  r = fma(-kd, ORC_PIO2_2, r);
  r = fma(-kd, ORC_PIO2_3, r);
  int q = (int)(((int64_t)kd) & 3);
  double z = r * r;
  double ps = 2.8114572543455206e-15;          /*  1/17! */
  ps = fma(ps, z, -7.647163731819816e-13);     /* -1/15! */
  ps = fma(ps, z, 1.6059043836821613e-10);     /*  1/13! */
  ps = fma(ps, z, -2.505210838544172e-08);     /* -1/11! */
  ps = fma(ps, z, 2.7557319223985893e-06);     /*  1/9!  */
  ps = fma(ps, z, -0.0001984126984126984);     /* -1/7!  */
  ps = fma(ps, z, 0.008333333333333333);       /*  1/5!  */
  ps = fma(ps, z, -0.16666666666666666);       /* -1/3!  */
  double s = fma(r * z, ps, r);
  double pc = -1.5619206968586225e-16;         /* -1/18! */
  pc = fma(pc, z, 4.779477332387385e-14);      /*  1/16! */
  pc = fma(pc, z, -1.1470745597729725e-11);    /* -1/14! */
  pc = fma(pc, z, 2.08767569878681e-09);       /*  1/12! */
  pc = fma(pc, z, -2.755731922398589e-07);     /* -1/10! */
  pc = fma(pc, z, 2.48015873015873e-05);       /*  1/8!  */
  pc = fma(pc, z, -0.001388888888888889);      /* -1/6!  */
  pc = fma(pc, z, 0.041666666666666664);       /*  1/4!  */
  pc = fma(pc, z, -0.5);                       /* -1/2!  */
  double c = fma(z, pc, 1.0);
  switch (q) {
    case 0: *sp = s; *cp = c; break;
    case 1: *sp = c; *cp = -s; break;
    case 2: *sp = -s; *cp = -c; break;
    default: *sp = -c; *cp = s; break;
  }
}

static inline double orc_sinh(double x)
{
  double ax = fabs(x);
  if (ax < 1.0) {
    double z = x * x;
    double p = 1.9572941063391263e-20;         /* 1/21! */
    p = fma(p, z, 8.22063524662433e-18);       /* 1/19! */
    p = fma(p, z, 2.8114572543455206e-15);     /* 1/17! */
    p = fma(p, z, 7.647163731819816e-13);      /* 1/15! */
    p = fma(p, z, 1.6059043836821613e-10);     /* 1/13! */
    p = fma(p, z, 2.505210838544172e-08);      /* 1/11! */
    p = fma(p, z, 2.7557319223985893e-06);     /* 1/9!  */
    p = fma(p, z, 0.0001984126984126984);      /* 1/7!  */
    p = fma(p, z, 0.008333333333333333);       /* 1/5!  */
    p = fma(p, z, 0.16666666666666666);        /* 1/3!  */
    return fma(x * z, p, x);
  }
  double e = orc_exp(ax);
  double h = 0.5 * e - 0.5 / e;
  return x < 0.0 ? -h : h;
}

static inline double orc_cosh(double x)
{
  double ax = fabs(x);
  if (ax < 1.0) {
    double z = x * x;
    double p = 4.110317623312165e-19;          /* 1/20! */
    p = fma(p, z, 1.5619206968586225e-16);     /* 1/18! */
    p = fma(p, z, 4.779477332387385e-14);      /* 1/16! */
    p = fma(p, z, 1.1470745597729725e-11);     /* 1/14! */
    p = fma(p, z, 2.08767569878681e-09);       /* 1/12! */
    p = fma(p, z, 2.755731922398589e-07);      /* 1/10! */
    p = fma(p, z, 2.48015873015873e-05);       /* 1/8!  */
    p = fma(p, z, 0.001388888888888889);       /* 1/6!  */
    p = fma(p, z, 0.041666666666666664);       /* 1/4!  */
    p = fma(p, z, 0.5);                        /* 1/2!  */
    return fma(z, p, 1.0);
  }
  double e = orc_exp(ax);
  return 0.5 * e + 0.5 / e;
}

/* tan(x) = sin(x) / cos(x) from orc_sincos (one rounding on top of the
 * two ~1-ulp results, so within a few ulp away from the poles). */
static inline double orc_tan(double x)
{
  double s, c;
  orc_sincos(x, &s, &c);
  return s / c;
}

/* fmod for x >= 0 (or -0.0), y > 0 finite: exact remainder. */
static inline double orc_fmod_pos(double x, double y)
{
  if (!(x < INFINITY)) return NAN;
  if (x < y) return x;
  double yy = y;
  while (yy * 2.0 <= x) yy = yy * 2.0;
  while (yy >= y) {
    if (x >= yy) x = x - yy;
    yy = yy * 0.5;
  }
  return x;
}

/* double -> int with defined behaviour for NaN / out-of-range values
 * (NaN -> 0, saturation), matching the GPU conversion instruction. */
static inline int orc_to_int(double v)
{
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return (int)v;
}

#endif
