// hps_device.hpp -- per-particle maps of the tracking engine, register form.
//
// Each map restates one reference CUDA kernel body for a single particle held
// in registers, so that any sequence of maps between two space-charge kicks
// runs in one pass over HBM.  Expression order follows the reference source
// line by line (the library is built with -ffp-contract=off), which is what
// makes the results bit-identical to a serial host restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hpsim_amd.h"
#include "hps_math.hpp"

namespace hps {

constexpr double CLIGHT = 299.792458;                 // constant.h:7
constexpr double PI = 3.14159265358979323846;        // constant.h:8
constexpr double TWOPI = 3.14159265358979323846 * 2.0;

constexpr uint32_t kLanes = 256;                 // canonical reduction lanes
constexpr uint32_t kPerLane = 4;                 // particles per lane per chunk
constexpr uint32_t kChunk = kLanes * kPerLane;   // particles per chunk
constexpr int kDepositShift = 36;                // fixed-point deposit: 2^-36

struct Particle {
  double x, xp, y, yp, phi, w;
  uint32_t loss, lloss;
};

enum OpKind : int32_t {
  OP_DRIFT = 1,     // SimulatePartialDriftKernel
  OP_QUAD = 2,      // SimulateHalfQuadKernel
  OP_GAP1 = 3,      // SimulateRFGapFirstHalfKernel
  OP_GAP2 = 4,      // SimulateRFGapSecondHalfKernel
  OP_ROT = 5,       // SimulateRotationKernel
  OP_STEER = 6,     // SimulateSteererKernel
  OP_APER_C = 7,    // SimulateCircularApertureKernel
  OP_APER_R = 8,    // SimulateRectangularApertureKernel
  OP_BUNCH = 9,     // SimulateBuncherKernel
  OP_DIP1 = 10,     // SimulateFirstHalfDipoleKernel
  OP_DIP2 = 11      // SimulateSecondHalfDipoleKernel
};

// Map classes by register cost; a k_segment instantiation compiles only the
// maps of its class and below, so segments without RF gaps (or quads) run
// with fewer registers and more wavefronts in flight.
enum OpClass : int32_t {
  OPC_LIGHT = 0,    // drift, rotation, steerer, apertures
  OPC_QUAD = 1,     // + half quad
  OPC_GAP = 2       // + RF gap halves, buncher, dipole halves
};

__host__ __device__ inline int32_t op_class_of(int32_t kind)
{
  if (kind == OP_GAP1 || kind == OP_GAP2 || kind == OP_BUNCH || kind == OP_DIP1 ||
      kind == OP_DIP2)
    return OPC_GAP;
  if (kind == OP_QUAD) return OPC_QUAD;
  return OPC_LIGHT;
}

// Element constants of the two RF-gap half maps: every subexpression of
// SimulateRFGapFirst/SecondHalfKernel that does not depend on the particle,
// evaluated once per map on the host by gap_uniform() with exactly the
// expressions the kernels use (IEEE double, no contraction), so the per
// particle arithmetic below is unchanged bit for bit.
struct GapUniform {
  // first half (simulation_kernel.cu:239-285)
  double cell_len, dd1, wave_len, twopi_dd1, beta_in1, pclbl_dps1, twopi_half_q1, pi_dcl1,
         betag1;
  int32_t dcl1_nonzero, pad0_;
  // second half (simulation_kernel.cu:326-493)
  double dd2, inv_mass, beta_in2, twopi_dd2, betag_ccl, gmg, inv_gmg2, q, beta_out,
         concomm, concomm1, neg_con1, neg_con2nex, con2nex, inv_gmgcl, inv_fkis, fkr, sinps,
         cosps, bgw, twopi_dg, cell_pi, q_efl, two_con1, efl_inv_mass, pclbl_dps2,
         twopi_half_q2, pi_dcl2;
  int32_t dcl2_nonzero, pad1_;
};

// design_w_in, phi_in (first half), phi (second half: phi_out for CCL),
// length, q1, q2: the arguments the reference passes to the two kernels.
HPS_HD GapUniform gap_uniform(const hps_gap_param& gap, double design_w_in, double phi_in1,
                              double phi_in2, double length, double q1, double q2,
                              uint32_t ccl_cell_num, bool ccl, double mass, double charge)
{
  GapUniform u{};
  // first half
  u.cell_len = length + q1 * 0.5 + q2 * 0.5;
  u.dd1 = 0.5 * u.cell_len - gap.dg - 0.5 * q1;
  u.wave_len = CLIGHT / gap.frequency;
  u.twopi_dd1 = TWOPI * u.dd1;
  {
    const double gam = design_w_in / mass;
    u.beta_in1 = ::sqrt((gam + 2.0) * gam) / (gam + 1.0);
  }
  if (!ccl) {
    const double dps = gap.phase_ref - phi_in1;
    const double d_cell_len = u.cell_len - gap.beta_center * u.wave_len;
    u.twopi_half_q1 = TWOPI * 0.5 * q1;
    u.dcl1_nonzero = (d_cell_len > 1e-15 || d_cell_len < -1e-15) ? 1 : 0;
    u.pi_dcl1 = PI * d_cell_len;
    u.pclbl_dps1 = PI * gap.cell_length_over_beta_lambda + dps;
  } else {
    const double dps = gap.phase_ref - phi_in1;
    u.betag1 = u.cell_len / (gap.cell_length_over_beta_lambda * u.wave_len);
    u.pclbl_dps1 = PI * gap.cell_length_over_beta_lambda + dps;
  }
  // second half
  u.dd2 = 0.5 * u.cell_len + gap.dg - q2 * 0.5;
  u.inv_mass = 1.0 / mass;
  const double gm_in = design_w_in * u.inv_mass + 1.0;
  u.beta_in2 = ::sqrt(1.0 - 1.0 / (gm_in * gm_in));
  u.twopi_dd2 = TWOPI * u.dd2;
  const double cl = u.cell_len / gap.cell_length_over_beta_lambda;
  u.betag_ccl = cl / u.wave_len;
  const double betag = ccl ? gap.beta_center : u.betag_ccl;
  const double betag2 = betag * betag;
  u.gmg = 1.0 / ::sqrt(1.0 - betag2);
  u.inv_gmg2 = 1.0 / (u.gmg * u.gmg);
  double q = charge;
  q = q > 0.0 ? q : -q;
  u.q = q;
  const double gm_out = gap.energy_out * u.inv_mass + 1.0;
  u.beta_out = ::sqrt(1.0 - 1.0 / (gm_out * gm_out));
  const double efl = gap.amplitude * u.cell_len;
  u.concomm = q * efl / (mass * betag2);
  u.concomm1 = u.concomm * u.inv_gmg2 / u.gmg;
  const double con1 = PI * u.concomm1;
  u.neg_con1 = -con1;
  u.con2nex = u.concomm * u.inv_gmg2;
  u.neg_con2nex = -u.con2nex;
  if (ccl) {
    u.inv_gmgcl = 1.0 / (u.gmg * betag * u.wave_len);
    u.inv_fkis = betag * u.wave_len / TWOPI;
    u.bgw = u.betag_ccl * u.wave_len;
  } else {
    u.inv_gmgcl = 1.0 / (u.gmg * cl);
    u.inv_fkis = cl / TWOPI;
    u.bgw = u.beta_in2 * u.wave_len;
  }
  u.fkr = TWOPI * u.inv_gmgcl;
  hps::sincos(gap.phase_ref, &u.sinps, &u.cosps);
  u.twopi_dg = TWOPI * gap.dg;
  u.cell_pi = ccl_cell_num * PI;
  u.q_efl = q * efl;
  u.two_con1 = 2.0 * con1;
  u.efl_inv_mass = efl * u.inv_mass;
  if (ccl) {
    const double dps = phi_in2 - gap.phase_ref;
    u.pclbl_dps2 = PI * gap.cell_length_over_beta_lambda + dps;
  } else {
    const double d_cell_len = u.cell_len - gap.beta_center * u.wave_len;
    u.twopi_half_q2 = TWOPI * 0.5 * q2;
    u.dcl2_nonzero = (d_cell_len > 1e-15 || d_cell_len < -1e-15) ? 1 : 0;
    u.pi_dcl2 = PI * d_cell_len;
    const double dps = gap.phase_ref - phi_in2;
    u.pclbl_dps2 = PI * gap.cell_length_over_beta_lambda + dps;
  }
  return u;
}

// One map step.  a[] meaning per kind:
//  DRIFT  h, aperture, wavelength
//  QUAD   length, aperture, gradient, wavelength
//  GAP1   design_w_in, phi_in, length, qlen1, qlen2
//  GAP2   design_w_in, phi_in, length, qlen1, qlen2, wavelength
//  ROT    angle
//  STEER  bl_h, bl_v
//  APER_C aperture        APER_R  xl, xr, yt, yb
//  BUNCH  w_ref, frequency, voltage, phase, aperture, frequency ratio, wavelength
//  DIP1/2 rho, angle, edge angle (in / out), half gap, k1, k2, field index,
//         kinetic energy, wavelength
struct Op {
  int32_t kind;
  uint32_t elem;          // loss code written on loss (element model index)
  int32_t ccl;            // gap form: 0 DTL, 1 CCL
  uint32_t ccl_cell_num;
  double a[10];
  hps_gap_param g;
  GapUniform gu;          // GAP1 / GAP2 element constants (gap_uniform)
};

// An op read through the constant address space: the op table does not
// change while a kernel runs, so its fields (uniform across the workgroup)
// load with scalar loads into SGPRs instead of vector loads into VGPRs,
// which the compiler cannot prove safe for a generic pointer once the kernel
// has stored particles.  The maps take their element arguments by template
// type for this reason.
typedef const __attribute__((address_space(4))) Op OpConst;

// Bessi0c, simulation_kernel.cu:15-36
__device__ __forceinline__ double bessi0c(double r_x)
{
  double ans, y;
  const double ax = __builtin_fabs(r_x);
  if (ax < 3.75) {
    y = r_x / 3.75;
    y = y * y;
    ans = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
          y * (0.360768e-1 + y * 0.45813e-2)))));
  } else {
    y = 3.75 / ax;
    ans = (hps::exp(ax) * (1.0 / psqrt(ax))) * (0.39894228 + y * (0.1328592e-1 +
          y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 +
          y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
  }
  return ans;
}

// Bessi1c, simulation_kernel.cu:38-64
__device__ __forceinline__ double bessi1c(double r_x)
{
  double y, ans;
  const double ax = __builtin_fabs(r_x);
  if (ax < 3.75) {
    y = r_x / 3.75;
    y = y * y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 +
          y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y *
          (-0.1031555e-1 + y * ans))));
    ans = ans * (hps::exp(ax) * (1.0 / psqrt(ax)));
  }
  return r_x < 0.0 ? -ans : ans;
}

// Bessi1p, simulation_kernel.cu:66-73
__device__ __forceinline__ double bessi1p(double r_x)
{
  if (r_x != 0.0) return bessi0c(r_x) - bessi1c(r_x) / r_x;
  return 0.5;
}

// bessi0c(x), bessi1c(x) and bessi1p(x) of one argument, sharing the
// common factors (identical results to the three separate calls).
__device__ __forceinline__ void bessel_all(double r_x, double& b0, double& b1, double& b1p)
{
  const double ax = __builtin_fabs(r_x);
  if (ax < 3.75) {
    double y = r_x / 3.75;
    y = y * y;
    b0 = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
         y * (0.360768e-1 + y * 0.45813e-2)))));
    b1 = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 +
         y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    const double y = 3.75 / ax;
    const double f = hps::exp(ax) * (1.0 / psqrt(ax));
    b0 = f * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
         y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 +
         y * 0.392377e-2))))))));
    double a1 = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    a1 = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y *
         (-0.1031555e-1 + y * a1))));
    b1 = a1 * f;
  }
  if (r_x < 0.0) b1 = -b1;
  b1p = r_x != 0.0 ? b0 - b1 / r_x : 0.5;
}

// GetTransitTimeFactors, simulation_kernel.cu:290-308
template <class G>
__device__ __forceinline__ void ttf(double r_beta, const G& g, double& t, double& tp, double& sp)
{
  double beta = r_beta;
  if (r_beta > 1.1 * g.fit_beta_center) beta = g.fit_beta_center;
  else if (r_beta < g.fit_beta_min) beta = g.fit_beta_min;
  t = g.fit_t0 + beta * (g.fit_t1 + beta * (g.fit_t2 + beta * (g.fit_t3 + beta *
      (g.fit_t4 + beta * g.fit_t5))));
  const double coef = beta * beta / (TWOPI * g.fit_beta_center);
  tp = coef * (g.fit_t1 + beta * (2 * g.fit_t2 + beta * (3 * g.fit_t3 + beta *
       (4 * g.fit_t4 + 5 * beta * g.fit_t5))));
  sp = -coef * (g.fit_s1 + beta * (2 * g.fit_s2 + beta * (3 * g.fit_s3 + beta *
       (4 * g.fit_s4 + 5 * beta * g.fit_s5))));
}

// SimulatePartialDriftKernel body, simulation_kernel.cu:95-115
__device__ __forceinline__ void drift(Particle& p, double h, double aper, double wavelen,
                                      uint32_t elem, double mass)
{
  const double aper2 = aper * aper;
  double x = p.x;
  double y = p.y;
  x += p.xp * h;
  y += p.yp * h;
  if (aper2 != 0.0 && x * x + y * y > aper2) {
    p.loss = elem;
  } else {
    p.x = x;
    p.y = y;
    const double gm1 = p.w / mass;
    const double bt = psqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
    p.phi += TWOPI * h / (wavelen * bt);
  }
}

// SimulateHalfQuadKernel body, simulation_kernel.cu:147-214
__device__ __forceinline__ void half_quad(Particle& p, double length, double aper,
                                          double gradient, double wavelen, uint32_t elem,
                                          double mass, double charge)
{
  double aper2 = aper;
  aper2 *= aper2;
  const double qg = gradient * charge;
  double x = p.x, y = p.y, xp = p.xp, yp = p.yp;
  const double w = p.w;
  if (qg == 0) {
    // the loss test below uses the coordinates before this half-drift
    p.x = x + xp * length * 0.5;
    p.y = y + yp * length * 0.5;
  } else {
    const double gamma1 = w / mass;
    const double bg = psqrt(gamma1 * (gamma1 + 2.0) * (1.0 + xp * xp + yp * yp));
    const double absqg = qg > 0.0 ? qg : -qg;
    const double k = psqrt(absqg / (bg * mass) * CLIGHT);
    const double inv_k = 1.0 / k;
    const double kds = length * 0.25 * k;
    double s, c;
    double sh, ch;
    psinhcosh(kds, &sh, &ch);
    psincos(kds, &s, &c);
    const double sovk = s * inv_k, shovk = sh * inv_k, sk = s * k, shk = sh * k;
    double xx, yy;
#pragma unroll
    for (int rep = 0; rep < 2; ++rep) {
      if (qg > 0.0) {
        xx = c * x + sovk * xp; xp = -sk * x + c * xp; x = xx;
        yy = ch * y + shovk * yp; yp = shk * y + ch * yp; y = yy;
      } else {
        xx = ch * x + shovk * xp; xp = shk * x + ch * xp; x = xx;
        yy = c * y + sovk * yp; yp = -sk * y + c * yp; y = yy;
      }
    }
    p.x = x; p.y = y; p.xp = xp; p.yp = yp;
  }
  if (aper2 != 0 && x * x + y * y > aper2) {
    p.loss = elem;
  } else {
    const double gm1 = w / mass;
    const double bt = psqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
    p.phi += TWOPI * length * 0.5 / (wavelen * bt);
  }
}

// SimulateRFGapFirstHalfKernel body, simulation_kernel.cu:239-285
template <class G, class U>
__device__ __forceinline__ void gap_first_half(Particle& p, const G& gap, const U& u, bool ccl,
                                               double mass)
{
  const double x = p.x, y = p.y, xp = p.xp, yp = p.yp, phi = p.phi, w = p.w;
  p.x = x + u.dd1 * xp;
  p.y = y + u.dd1 * yp;
  if (gap.amplitude == 0.0) {
    const double gm1 = w / mass;
    const double bt = psqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
    p.phi += u.twopi_dd1 / (u.wave_len * bt);
  } else {
    const double beta_in = u.beta_in1;
    const double gam = w / mass;
    const double beta = psqrt((gam + 2.0) * gam) / (gam + 1.0);
    if (!ccl) {
      const double dphi_qlen1 = u.twopi_half_q1 / (beta * u.wave_len);
      double dphi_clen = 0.0;
      if (u.dcl1_nonzero) dphi_clen = u.pi_dcl1 / (beta * u.wave_len);
      p.phi = phi + (1.0 - (beta - beta_in) / beta) * u.pclbl_dps1 + dphi_clen - dphi_qlen1;
    } else {
      p.phi = phi + (1.0 - (beta - beta_in) * u.betag1 / (beta * beta_in)) * u.pclbl_dps1;
    }
  }
}

// SimulateRFGapSecondHalfKernel body, simulation_kernel.cu:326-493
template <class G, class U>
__device__ __forceinline__ void gap_second_half(Particle& p, const G& gap, const U& u,
                                                double wavelen, bool ccl,
                                                double mass)
{
  double x = p.x, y = p.y, xp = p.xp, yp = p.yp;
  const double inv_mass = u.inv_mass;
  const double phi = p.phi;
  double w = p.w;
  const double beta_in = u.beta_in2;
  double gam = w * inv_mass;
  const double bg1 = psqrt(gam * (gam + 2.0));
  double beta = bg1 / (gam + 1.0);
  double t_fac, tp_fac, sp_fac;
  ttf(beta, gap, t_fac, tp_fac, sp_fac);
  if (gap.amplitude == 0.0 || (ccl && beta < gap.fit_beta_min)) {
    p.x = x + u.dd2 * xp;
    p.y = y + u.dd2 * yp;
    const double gm1 = p.w / mass;
    const double bt = psqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
    p.phi += u.twopi_dd2 / (wavelen * bt);
    return;
  }
  const double wave_len = u.wave_len;
  const double gmg = u.gmg, inv_gmg2 = u.inv_gmg2;
  double con5nex, tw;
  if (ccl) {
    con5nex = u.concomm * (t_fac * u.inv_fkis - u.bgw * tp_fac * inv_gmg2);
    tw = t_fac - TWOPI * tp_fac * (beta_in / beta - 1.0) * u.betag_ccl / beta_in;
  } else {
    con5nex = u.concomm * (t_fac * u.inv_fkis - u.bgw * tp_fac * inv_gmg2);
    tw = t_fac;
  }
  const double con5nex1 = u.concomm1 * t_fac;
  const double con5nex2 = u.neg_con2nex * t_fac * u.inv_fkis;
  const double fkr = u.fkr;
  const double sinps = u.sinps, cosps = u.cosps;
  double phi1, sinp1, cosp1;
  if (ccl)
    phi1 = phi + gap.phase_shift - u.twopi_dg / (beta * wave_len) + u.cell_pi;
  else
    phi1 = phi + gap.phase_shift;
  psincos(phi1, &sinp1, &cosp1);
  const double prome1 = u.neg_con1 * (tp_fac * (sinp1 - sinps) + sp_fac * (cosp1 - cosps));
  psincos(phi1 + prome1, &sinp1, &cosp1);
  const double rs_p = x * x + y * y;
  const double rrp = x * xp + y * yp;
  double rpris = 0.0, bes1covr;
  const double ris = psqrt(rs_p);
  const double fkrris = fkr * ris;
  double b0c, b1c, b1p;
  bessel_all(fkrris, b0c, b1c, b1p);
  if (ris != 0.0) bes1covr = b1c / ris;
  else bes1covr = fkr * 0.5;
  const double delte = u.q_efl * (tw * b0c * cosp1 + ((tw - TWOPI * tp_fac) / gmg * bes1covr +
                       tw / gmg * fkr * b1p - tw * gmg * bes1covr) * rrp * sinp1);
  w += delte;
  gam = w * inv_mass;
  const double bg = psqrt(gam * (gam + 2.0));
  beta = bg / (gam + 1.0);
  const double bgr = bg1 / bg;
  const double onema = 1.0 - (con5nex * bes1covr + con5nex1 * b1p + con5nex2 * bes1covr) * cosp1;
  const double dovbl = u.con2nex * bes1covr * tw * sinp1;
  const double bgrovonema = bgr / onema;
  xp = -dovbl * x + bgrovonema * xp;
  x = x * onema + u.dd2 * xp;
  yp = -dovbl * y + bgrovonema * yp;
  y = y * onema + u.dd2 * yp;
  if (rs_p > 0.0) rpris = rrp / ris;
  const double prome = u.two_con1 * (-tp_fac * b0c + ris * tw * b1c * u.inv_gmgcl) *
                       (sinp1 - sinps) - u.efl_inv_mass * tw * b1c * rpris * (cosp1 - cosps);
  double phi_new;
  if (ccl) {
    phi_new = phi + (1.0 - (beta - u.beta_out) * u.betag_ccl / (beta * u.beta_out)) *
              u.pclbl_dps2 + prome;
  } else {
    const double dphi_qlen2 = u.twopi_half_q2 / (beta * wave_len);
    double dphi_clen = 0.0;
    if (u.dcl2_nonzero) dphi_clen = u.pi_dcl2 / (beta * wave_len);
    phi_new = phi + (1.0 - (beta - u.beta_out) / beta) * u.pclbl_dps2 + prome + dphi_clen -
              dphi_qlen2;
  }
  if (w <= 0.0 || x != x || y != y || xp != xp || yp != yp || w != w || phi_new != phi_new) {
    p.loss = 33333333;
  } else {
    p.x = x; p.y = y; p.xp = xp; p.yp = yp; p.w = w; p.phi = phi_new;
  }
}

// SimulateRotationKernel body, simulation_kernel.cu:516-528
__device__ __forceinline__ void rotation(Particle& p, double angle)
{
  double sina, cosa;
  hps::sincos(angle, &sina, &cosa);
  const double x = p.x, y = p.y, xp = p.xp, yp = p.yp;
  p.x = x * cosa + y * sina;
  p.y = y * cosa - x * sina;
  p.xp = xp * cosa + yp * sina;
  p.yp = yp * cosa - xp * sina;
}

// SimulateSteererKernel body, simulation_kernel.cu:128-137
__device__ __forceinline__ void steerer(Particle& p, double blh, double blv, double mass,
                                        double charge)
{
  const double gm1 = p.w / mass;
  const double btgm = psqrt(gm1 * (gm1 + 2.0));
  const double pm = btgm * mass / CLIGHT;
  if (blh > 1e-10) p.xp += blh * charge / pm;
  if (blv > 1e-10) p.yp += blv * charge / pm;
}

// Apply one map step to an alive particle.
// SimulateBuncherKernel body, simulation_kernel.cu:807-874 (rsqrt as 1/sqrt).
template <class D>
__device__ __forceinline__ void buncher(Particle& p, D* a, uint32_t elem, double mass,
                                        double charge)
{
  const double w_ref = a[0], freq = a[1], voltage = a[2], phase = a[3], aper = a[4];
  const double freq_ratio = a[5], wavelen = a[6];
  const double x = p.x, y = p.y, xp = p.xp, yp = p.yp;
  const double phi = p.phi * freq_ratio;
  const double w = p.w;
  const double r2 = x * x + y * y;
  const double aper2 = aper * aper;
  if (aper2 != 0.0 && r2 > aper2) { p.loss = elem; return; }
  const double h = freq * wavelen / CLIGHT;
  const double gm1 = w_ref / mass;
  const double gm = gm1 + 1.0;
  const double btgm = psqrt(gm1 * (gm1 + 2.0));
  const double beta = btgm / gm;
  const double k = h * TWOPI / (btgm * wavelen);
  const double k2 = k * k;
  const double q = charge > 0.0 ? charge : -charge;
  const double con1 = TWOPI * h * voltage * q / (beta * wavelen * mass);
  const double arg = k2 * r2 * 0.25;
  const double ph = h * phi + phase;
  double sph, cph;
  psincos(ph, &sph, &cph);
  const double dw = q * voltage * cph * (1.0 + arg * (1.0 + arg * (0.25 + arg / 36.0)));
  const double wb_par = w + 0.5 * dw;
  const double wf = w + dw;
  if (wb_par <= 0.0 || wf <= 0.0) { p.loss = elem; return; }
  double gmtmp = wb_par / mass;
  const double bav = psqrt(gmtmp * (gmtmp + 2.0)) / (gmtmp + 1.0);
  gmtmp = w / mass;
  const double bg = psqrt(gmtmp * (gmtmp + 2.0));
  gmtmp = wf / mass;
  const double rbgf = 1.0 / psqrt(gmtmp * (gmtmp + 2.0));
  const double del = -con1 * sph * (1.0 - bav * beta) * (0.5 + arg * (0.25 + arg / 24.0)) / bav;
  p.xp = (xp * bg + del * x) * rbgf;
  p.yp = (yp * bg + del * y) * rbgf;
  p.w = wf;
}

// Half a sector dipole (arc rho * angle / 2, field index n) and the pole-face
// edge kick, SimulateFirst/SecondHalfDipoleKernel, simulation_kernel.cu:603-805.
struct DipoleHalf {
  double r11, r12, r16, r21, r22, r26, r33, r34, r43, r44, r51, r52, r56, beta, dl;
};

template <class D>
__device__ __forceinline__ DipoleHalf dipole_half(D* a, double mass)
{
  DipoleHalf o;
  const double rho = a[0], angle = a[1], findx = a[7], design_w = a[8];
  const double h = 1.0 / rho;
  const double h2 = h * h;
  const double dl = 0.5 * rho * angle;
  const double yk2 = findx * h2;
  const double yk = psqrt(__builtin_fabs(yk2));
  const double inv_yk = 1.0 / yk;
  const double ykl = yk * dl;
  const double xk2 = h2 - yk2;
  const double inv_xk2 = 1.0 / xk2;
  const double xk = psqrt(__builtin_fabs(xk2));
  const double inv_xk = 1.0 / xk;
  const double xkl = xk * dl;
  double cxkl = 1.0, sxokx = dl, s, c;
  if (xk2 > 0.0) { hps::sincos(xkl, &s, &c); cxkl = c; sxokx = s * inv_xk; }
  else if (xk2 < 0.0) { double sh; hps::sinhcosh(xkl, &sh, &cxkl); sxokx = sh * inv_xk; }
  double cykl = 1.0, syoky = dl;
  if (yk2 > 0.0) { hps::sincos(ykl, &s, &c); cykl = c; syoky = s * inv_yk; }
  else if (yk2 < 0.0) { double sh; hps::sinhcosh(ykl, &sh, &cykl); syoky = sh * inv_yk; }
  o.r11 = cxkl; o.r12 = sxokx; o.r21 = -xk2 * sxokx; o.r22 = cxkl;
  o.r33 = cykl; o.r34 = syoky; o.r43 = -yk2 * syoky; o.r44 = cykl;
  o.r26 = h * sxokx;
  o.r16 = h * (1.0 - cxkl) * inv_xk2;
  o.r51 = -o.r26;
  o.r52 = -o.r16;
  const double gamma = design_w / mass + 1.0;
  const double beta2 = 1.0 - 1.0 / (gamma * gamma);
  o.beta = psqrt(beta2);
  o.r56 = -h2 * (dl * beta2 - sxokx) * inv_xk2 + dl * (1.0 - h2 * inv_xk2) / (gamma * gamma);
  o.dl = dl;
  return o;
}

template <class D>
__device__ __forceinline__ void dipole_edge(D* a, double b, double& r21, double& r43)
{
  const double h = 1.0 / a[0];
  const double c1 = 2.0 * a[4] * a[5] * h;
  double sinb, cosb;
  hps::sincos(b, &sinb, &cosb);
  const double inv_cosb = 1.0 / cosb;
  const double psi = c1 * (1.0 + sinb * sinb) * inv_cosb * (1.0 - c1 * a[6] * sinb * inv_cosb);
  r21 = hps::tan(b) * h;
  r43 = hps::tan(b - psi) * h;
}

template <class D>
__device__ __forceinline__ void dipole_first_half(Particle& p, D* a, double mass)
{
  double er21, er43;
  dipole_edge(a, a[2], er21, er43);
  const DipoleHalf d = dipole_half(a, mass);
  const double design_w = a[8], wavelen = a[9];
  const double gamma = design_w / mass + 1.0;
  const double x = p.x, y = p.y;
  double xp = p.xp, yp = p.yp;
  xp += er21 * x;
  yp -= er43 * y;
  const double dpop = gamma / (design_w * (gamma + 1.0)) * (p.w - design_w);
  p.x = d.r11 * x + d.r12 * xp + d.r16 * dpop;
  p.xp = d.r21 * x + d.r22 * xp + d.r26 * dpop;
  p.y = d.r33 * y + d.r34 * yp;
  p.yp = d.r43 * y + d.r44 * yp;
  p.phi += TWOPI * (-(d.r51 * x + d.r52 * xp + d.r56 * dpop) + d.dl) / (d.beta * wavelen);
}

template <class D>
__device__ __forceinline__ void dipole_second_half(Particle& p, D* a, double mass)
{
  double er21, er43;
  dipole_edge(a, a[3], er21, er43);
  const DipoleHalf d = dipole_half(a, mass);
  const double design_w = a[8], wavelen = a[9];
  const double gamma = design_w / mass + 1.0;
  const double x = p.x, y = p.y;
  double xp = p.xp, yp = p.yp;
  const double dpop = gamma / (design_w * (gamma + 1.0)) * (p.w - design_w);
  p.phi += TWOPI * (-(d.r51 * x + d.r52 * xp + d.r56 * dpop) + d.dl) / (d.beta * wavelen);
  const double x_new = d.r11 * x + d.r12 * xp + d.r16 * dpop;
  xp = d.r21 * x + d.r22 * xp + d.r26 * dpop;
  const double y_new = d.r33 * y + d.r34 * yp;
  yp = d.r43 * y + d.r44 * yp;
  p.x = x_new;
  p.y = y_new;
  p.xp = xp + er21 * x_new;
  p.yp = yp - er43 * y_new;
}

template <int kClass, class OpT>
__device__ __forceinline__ void apply_op(const OpT& op, Particle& p, uint32_t index, double mass,
                                         double charge, double x_avg, double y_avg)
{
  switch (op.kind) {
    case OP_DRIFT: drift(p, op.a[0], op.a[1], op.a[2], op.elem, mass); break;
    case OP_QUAD:
      if constexpr (kClass >= OPC_QUAD)
        half_quad(p, op.a[0], op.a[1], op.a[2], op.a[3], op.elem, mass, charge);
      break;
    case OP_GAP1:
      if constexpr (kClass >= OPC_GAP) gap_first_half(p, op.g, op.gu, op.ccl != 0, mass);
      break;
    case OP_GAP2:
      if constexpr (kClass >= OPC_GAP)
        gap_second_half(p, op.g, op.gu, op.a[5], op.ccl != 0, mass);
      break;
    case OP_BUNCH:
      if constexpr (kClass >= OPC_GAP) buncher(p, op.a, op.elem, mass, charge);
      break;
    case OP_DIP1:
      if constexpr (kClass >= OPC_GAP) dipole_first_half(p, op.a, mass);
      break;
    case OP_DIP2:
      if constexpr (kClass >= OPC_GAP) dipole_second_half(p, op.a, mass);
      break;
    case OP_ROT: rotation(p, op.a[0]); break;
    case OP_STEER: steerer(p, op.a[0], op.a[1], mass, charge); break;
    case OP_APER_C: {   // simulation_kernel.cu:548-555 (particle 0 exempt)
      if (index == 0) break;
      const double aper2 = op.a[0] * op.a[0];
      const double x = p.x - x_avg, y = p.y - y_avg;
      if (x * x + y * y > aper2) p.loss = op.elem;
      break;
    }
    case OP_APER_R: {   // simulation_kernel.cu:564-574 (particle 0 exempt)
      if (index == 0) break;
      const double xmin = x_avg - op.a[0], xmax = x_avg + op.a[1];
      const double ymin = y_avg - op.a[3], ymax = y_avg + op.a[2];
      if (p.x < xmin || p.x > xmax || p.y < ymin || p.y > ymax) p.loss = op.elem;
      break;
    }
    default: break;
  }
}

// Radial and longitudinal field at radius rp and axial distance d of a ring
// of radius rs, before the c1 charge factor: the complete-elliptic-integral
// fit of UpdateFldTbl1Kernel (space_charge_kernel.cu:52-71, repeated for the
// image rings at :74-94); er is 0 on the axis.
HPS_HD void ring_green(double rs, double rp, double d, double& er, double& ez)
{
  const double d2 = d * d;
  const double c = (rp - rs) * (rp - rs);
  const double b = (rp + rs) * (rp + rs);
  const double rsrp = rs * rp;
  double a = 4.0 * rsrp / (b + d2);
  const double g = 1.0 - a;
  const double h = hps::log(g);
  const double ee = 1.0 + g * (0.4630106 - 0.2452740 * h + g * (0.1077857 - 0.04125321 * h));
  const double ek = 1.38629436 - 0.5 * h + g * (0.1119697 - 0.1213486 * h + g *
                    (0.07253230 - 0.02887472 * h));
  a = ::sqrt(b + d2);
  const double rs2mrp2 = (rs * rs - rp * rp);
  er = 0.0;
  if (rp != 0.0) er = (ek - (rs2mrp2 + d2) * ee / (c + d2)) / (2.0 * rp * a);
  ez = d * ee / (a * (c + d2));
}

// log of the bunch-frame cell aspect dz*gm/dr of a SCHEFF mesh (0 for
// spacings that are not positive and finite): the one parameter of the
// Green's table normalised to dr = 1, which the grouped field solve of a
// batch interpolates in.
HPS_HD double log_aspect(double dr, double dz, double gm)
{
  const double v = dz * gm / dr;
  if (!(v > 0.0) || !(v < 1.0e300)) return 0.0;
  return hps::log(v);
}

// Nearest class c = round(lk / lnr) of a log aspect lk on the grid of step
// lnr and the Lagrange weights of classes c - 1, c, c + 1 for quadratic
// interpolation in lk (t = lk / lnr - c in [-1/2, 1/2)).
HPS_HD int32_t aspect_lagrange(double lk, double lnr, double (&w)[3])
{
  const double x = lk / lnr;
  const double c = ::floor(x + 0.5);
  const double t = x - c;
  w[0] = 0.5 * t * (t - 1.0);
  w[1] = (1.0 - t) * (1.0 + t);
  w[2] = 0.5 * t * (t + 1.0);
  return (int32_t)c;
}

// UpdateRelativePhiKernel body, beam_kernel.cu:1514-1520
__device__ __forceinline__ double relative_phi(double phi, double ref)
{
  double tmp = phi - ref + PI;
  double sign = 1.0;
  if (tmp < 0) { sign = -1.0; tmp = -tmp; }
  return sign * (hps::fmod_pos(tmp, TWOPI) - PI);
}

}  // namespace hps
