// hps_math.hpp -- FP64 elementary functions for the tracking kernels.
//
// The reference kernels use the CUDA device math library (exp, log, sinh,
// cosh, sincos, rsqrt, fmod: simulation_kernel.cu:30,176-177,
// space_charge_kernel.cu:60, beam_kernel.cu:1520).  These functions follow
// published methods instead and are written so that a serial host build of
// the same method gives identical bits:
//   exp    Cody-Waite reduction by ln2 (two-part constant, fma), degree-13
//          Taylor polynomial, exact scaling by two powers of two;
//   log    x = 2^e m, m in [sqrt(1/2), sqrt(2)], 2*atanh((m-1)/(m+1)) series;
//   sincos Cody-Waite reduction by pi/2 (three-part constant, fma), Taylor
//          polynomials (sin to r^17, cos to r^18) on |r| <= pi/4;
//   sinh/cosh  Taylor series for |x| < 1, else (e^|x| -/+ e^-|x|)/2;
//   fmod   exact remainder by subtracting power-of-two multiples of the
//          divisor (each step exact by Sterbenz's lemma).
// Only correctly rounded primitives are used (+ - * /, sqrt, fma, floor);
// the library is compiled with -ffp-contract=off.  The exception is the
// second build of the tracking kernels (hps_push_fast.hip).  It serves
// batches held to error bounds, not to the oracle's bits.  Its maps, kick
// and deposit take psqrt, psincos and psinhcosh as sqrt_rsq (v_rsq_f64 and
// two refinement steps) and shorter Taylor polynomials.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HPS_HD __host__ __device__ __forceinline__

namespace hps {

constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kPio2A = 0x1.921fb54442d18p+0;
constexpr double kPio2B = 0x1.1a62633145c07p-54;
constexpr double kPio2C = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kInf = __builtin_huge_val();

HPS_HD uint64_t bits_of(double v) { return __builtin_bit_cast(uint64_t, v); }
HPS_HD double double_of(uint64_t u) { return __builtin_bit_cast(double, u); }
HPS_HD double pow2i(int k) { return double_of((uint64_t)(k + 1023) << 52); }

HPS_HD double exp(double x)
{
  if (x != x) return x;
  if (x > 709.79) return kInf;
  if (x < -745.2) return 0.0;
  const double kd = ::floor(x * kInvLn2 + 0.5);
  double r = ::fma(-kd, kLn2Hi, x);
  r = ::fma(-kd, kLn2Lo, r);
  // Taylor coefficients 1/13! ... 1/2!, then 1, 1.
  double p = 1.6059043836821613e-10;
  p = ::fma(p, r, 2.08767569878681e-09);
  p = ::fma(p, r, 2.505210838544172e-08);
  p = ::fma(p, r, 2.755731922398589e-07);
  p = ::fma(p, r, 2.7557319223985893e-06);
  p = ::fma(p, r, 2.48015873015873e-05);
  p = ::fma(p, r, 0.0001984126984126984);
  p = ::fma(p, r, 0.001388888888888889);
  p = ::fma(p, r, 0.008333333333333333);
  p = ::fma(p, r, 0.041666666666666664);
  p = ::fma(p, r, 0.16666666666666666);
  p = ::fma(p, r, 0.5);
  p = ::fma(p, r, 1.0);
  p = ::fma(p, r, 1.0);
  const int k = (int)kd;
  const int k1 = k / 2;
  return (p * pow2i(k1)) * pow2i(k - k1);
}

HPS_HD double log(double x)
{
  if (x != x) return x;
  if (x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -kInf;
  if (x == kInf) return x;
  uint64_t u = bits_of(x);
  int e = (int)((u >> 52) & 0x7ff);
  if (e == 0) {               // subnormal: renormalise
    x = x * 0x1p54;
    u = bits_of(x);
    e = (int)((u >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  double m = double_of((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > kSqrt2) { m = m * 0.5; e += 1; }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  // 1/23, 1/21, ..., 1/3
  double p = 0.043478260869565216;
  p = ::fma(p, z, 0.047619047619047616);
  p = ::fma(p, z, 0.05263157894736842);
  p = ::fma(p, z, 0.058823529411764705);
  p = ::fma(p, z, 0.06666666666666667);
  p = ::fma(p, z, 0.07692307692307693);
  p = ::fma(p, z, 0.09090909090909091);
  p = ::fma(p, z, 0.1111111111111111);
  p = ::fma(p, z, 0.14285714285714285);
  p = ::fma(p, z, 0.2);
  p = ::fma(p, z, 0.3333333333333333);
  const double s2 = 2.0 * s;
  const double lm = ::fma(s2 * z, p, s2);
  const double ed = (double)e;
  return ::fma(ed, kLn2Hi, ::fma(ed, kLn2Lo, lm));
}

HPS_HD void sincos(double x, double* sp, double* cp)
{
  if (!(__builtin_fabs(x) < kInf)) { *sp = x - x; *cp = x - x; return; }
  const double kd = ::floor(x * kTwoOverPi + 0.5);
  double r = ::fma(-kd, kPio2A, x);
  r = ::fma(-kd, kPio2B, r);
  r = ::fma(-kd, kPio2C, r);
  const int q = (int)(((int64_t)kd) & 3);
  const double z = r * r;
  // sin: 1/17!, -1/15!, ..., -1/3!
  double ps = 2.8114572543455206e-15;
  ps = ::fma(ps, z, -7.647163731819816e-13);
  ps = ::fma(ps, z, 1.6059043836821613e-10);
  ps = ::fma(ps, z, -2.505210838544172e-08);
  ps = ::fma(ps, z, 2.7557319223985893e-06);
  ps = ::fma(ps, z, -0.0001984126984126984);
  ps = ::fma(ps, z, 0.008333333333333333);
  ps = ::fma(ps, z, -0.16666666666666666);
  const double s = ::fma(r * z, ps, r);
  // cos: -1/18!, 1/16!, ..., -1/2!
  double pc = -1.5619206968586225e-16;
  pc = ::fma(pc, z, 4.779477332387385e-14);
  pc = ::fma(pc, z, -1.1470745597729725e-11);
  pc = ::fma(pc, z, 2.08767569878681e-09);
  pc = ::fma(pc, z, -2.755731922398589e-07);
  pc = ::fma(pc, z, 2.48015873015873e-05);
  pc = ::fma(pc, z, -0.001388888888888889);
  pc = ::fma(pc, z, 0.041666666666666664);
  pc = ::fma(pc, z, -0.5);
  const double c = ::fma(z, pc, 1.0);
  if (q == 0) { *sp = s; *cp = c; }
  else if (q == 1) { *sp = c; *cp = -s; }
  else if (q == 2) { *sp = -s; *cp = -c; }
  else { *sp = -c; *cp = s; }
}

HPS_HD double sinh(double x)
{
  const double ax = __builtin_fabs(x);
  if (ax < 1.0) {
    const double z = x * x;
    double p = 1.9572941063391263e-20;   // 1/21!
    p = ::fma(p, z, 8.22063524662433e-18);
    p = ::fma(p, z, 2.8114572543455206e-15);
    p = ::fma(p, z, 7.647163731819816e-13);
    p = ::fma(p, z, 1.6059043836821613e-10);
    p = ::fma(p, z, 2.505210838544172e-08);
    p = ::fma(p, z, 2.7557319223985893e-06);
    p = ::fma(p, z, 0.0001984126984126984);
    p = ::fma(p, z, 0.008333333333333333);
    p = ::fma(p, z, 0.16666666666666666);
    return ::fma(x * z, p, x);
  }
  const double e = hps::exp(ax);
  const double h = 0.5 * e - 0.5 / e;
  return x < 0.0 ? -h : h;
}

HPS_HD double cosh(double x)
{
  const double ax = __builtin_fabs(x);
  if (ax < 1.0) {
    const double z = x * x;
    double p = 4.110317623312165e-19;    // 1/20!
    p = ::fma(p, z, 1.5619206968586225e-16);
    p = ::fma(p, z, 4.779477332387385e-14);
    p = ::fma(p, z, 1.1470745597729725e-11);
    p = ::fma(p, z, 2.08767569878681e-09);
    p = ::fma(p, z, 2.755731922398589e-07);
    p = ::fma(p, z, 2.48015873015873e-05);
    p = ::fma(p, z, 0.001388888888888889);
    p = ::fma(p, z, 0.041666666666666664);
    p = ::fma(p, z, 0.5);
    return ::fma(z, p, 1.0);
  }
  const double e = hps::exp(ax);
  return 0.5 * e + 0.5 / e;
}

// sinh and cosh of one argument with a single exp: the same two results as
// sinh(x) and cosh(x) above, bit for bit.
HPS_HD void sinhcosh(double x, double* shp, double* chp)
{
  const double ax = __builtin_fabs(x);
  if (ax < 1.0) {
    *shp = hps::sinh(x);
    *chp = hps::cosh(x);
    return;
  }
  const double e = hps::exp(ax);
  const double h = 0.5 * e - 0.5 / e;
  *shp = x < 0.0 ? -h : h;
  *chp = 0.5 * e + 0.5 / e;
}

// tan(x) = sin(x) / cos(x) from sincos.
HPS_HD double tan(double x)
{
  double s, c;
  sincos(x, &s, &c);
  return s / c;
}

// Square root from v_rsq_f64.  One Goldschmidt step refines g ~ sqrt(x)
// and h ~ 1 / (2 sqrt(x)); one Newton correction of g follows.  For
// 2^-767 <= x < inf the result is within about one unit in the last place.
// The correctly rounded ::sqrt adds a second correction, scales x below
// 2^-767, and costs 21 instructions against 12.  +-0 and +inf are returned
// as they are; negative x and NaN give NaN.  Positive x below 2^-767 is
// outside its range.  Device code only.
__device__ __forceinline__ double sqrt_rsq(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = ::fma(-h, g, 0.5);
  g = ::fma(g, r, g);
  h = ::fma(h, r, h);
  const double d = ::fma(-g, g, x);
  g = ::fma(d, h, g);
  return (x == 0.0 || x == kInf) ? x : g;
}

// sincos with the Taylor polynomials cut at r^13 (sin) and r^14 (cos) after
// the same reduction: on |r| <= pi/4 the first omitted terms are below
// 2.1e-14 and 1.1e-15.  Device code only.
__device__ __forceinline__ void sincos_short(double x, double* sp, double* cp)
{
  if (!(__builtin_fabs(x) < kInf)) { *sp = x - x; *cp = x - x; return; }
  const double kd = ::floor(x * kTwoOverPi + 0.5);
  double r = ::fma(-kd, kPio2A, x);
  r = ::fma(-kd, kPio2B, r);
  r = ::fma(-kd, kPio2C, r);
  const int q = (int)(((int64_t)kd) & 3);
  const double z = r * r;
  // sin: 1/13!, -1/11!, ..., -1/3!
  double ps = 1.6059043836821613e-10;
  ps = ::fma(ps, z, -2.505210838544172e-08);
  ps = ::fma(ps, z, 2.7557319223985893e-06);
  ps = ::fma(ps, z, -0.0001984126984126984);
  ps = ::fma(ps, z, 0.008333333333333333);
  ps = ::fma(ps, z, -0.16666666666666666);
  const double s = ::fma(r * z, ps, r);
  // cos: 1/14!, -1/12!, ..., -1/2!
  double pc = -1.1470745597729725e-11;
  pc = ::fma(pc, z, 2.08767569878681e-09);
  pc = ::fma(pc, z, -2.755731922398589e-07);
  pc = ::fma(pc, z, 2.48015873015873e-05);
  pc = ::fma(pc, z, -0.001388888888888889);
  pc = ::fma(pc, z, 0.041666666666666664);
  pc = ::fma(pc, z, -0.5);
  const double c = ::fma(z, pc, 1.0);
  if (q == 0) { *sp = s; *cp = c; }
  else if (q == 1) { *sp = c; *cp = -s; }
  else if (q == 2) { *sp = -s; *cp = -c; }
  else { *sp = -c; *cp = s; }
}

// sinhcosh with the Taylor series for |x| < 1 cut at x^15 (sinh) and x^16
// (cosh): the first omitted terms at |x| = 1 are 2.8e-15 and 1.6e-16 of the
// results.  |x| >= 1 as sinhcosh.  Device code only.
__device__ __forceinline__ void sinhcosh_short(double x, double* shp, double* chp)
{
  const double ax = __builtin_fabs(x);
  if (ax < 1.0) {
    const double z = x * x;
    double p = 7.647163731819816e-13;    // 1/15!
    p = ::fma(p, z, 1.6059043836821613e-10);
    p = ::fma(p, z, 2.505210838544172e-08);
    p = ::fma(p, z, 2.7557319223985893e-06);
    p = ::fma(p, z, 0.0001984126984126984);
    p = ::fma(p, z, 0.008333333333333333);
    p = ::fma(p, z, 0.16666666666666666);
    *shp = ::fma(x * z, p, x);
    double c = 4.779477332387385e-14;     // 1/16!
    c = ::fma(c, z, 1.1470745597729725e-11);
    c = ::fma(c, z, 2.08767569878681e-09);
    c = ::fma(c, z, 2.755731922398589e-07);
    c = ::fma(c, z, 2.48015873015873e-05);
    c = ::fma(c, z, 0.001388888888888889);
    c = ::fma(c, z, 0.041666666666666664);
    c = ::fma(c, z, 0.5);
    *chp = ::fma(z, c, 1.0);
    return;
  }
  sinhcosh(x, shp, chp);
}

// sincos and sinhcosh of the per-particle maps: the full polynomials, or the
// short ones in the build of hps_push_fast.hip (see psqrt below).
static __device__ __forceinline__ void psincos(double x, double* sp, double* cp)
{
#ifdef HPS_FAST_PUSH
  sincos_short(x, sp, cp);
#else
  sincos(x, sp, cp);
#endif
}

static __device__ __forceinline__ void psinhcosh(double x, double* shp, double* chp)
{
#ifdef HPS_FAST_PUSH
  sinhcosh_short(x, shp, chp);
#else
  sinhcosh(x, shp, chp);
#endif
}

// The square root of the per-particle maps, the kick and the deposit: the
// correctly rounded ::sqrt, or sqrt_rsq in the build of hps_push_fast.hip,
// which serves batches whose results are held to error bounds, not to the
// oracle's bits.  Internal to each translation unit, so the two builds keep
// their own.
static __device__ __forceinline__ double psqrt(double x)
{
#ifdef HPS_FAST_PUSH
  return sqrt_rsq(x);
#else
  return ::sqrt(x);
#endif
}

// Remainder of x >= 0 by y > 0 (finite), exact.
HPS_HD double fmod_pos(double x, double y)
{
  if (!(x < kInf)) return __builtin_nan("");
  if (x < y) return x;
  double yy = y;
  while (yy * 2.0 <= x) yy = yy * 2.0;
  while (yy >= y) {
    if (x >= yy) x = x - yy;
    yy = yy * 0.5;
  }
  return x;
}

// double -> int: NaN -> 0, saturating at the int range.
HPS_HD int to_int(double v)
{
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return -2147483647 - 1;
  return (int)v;
}

}  // namespace hps
